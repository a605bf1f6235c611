"""bench.py's contract pieces that run without a GPU: the workloads exist, more than one GPU is
refused with a message, and the byte model of the headline workload."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_workloads_exist():
    for name in ("solve_workload", "learned_workload", "batch_workload", "batch_problem", "newton_workload", "cpu_baseline",
                 "copy_probe"):
        assert callable(getattr(bench, name)), name


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_headline_byte_model():
    # 64^3 hexahedra, p = 6: N_tot = 385^3, N_int = 383^3, 343 Gauss points per element.  The algorithmic bytes
    # of one application with the geometry recomputed from the Q1 vertices (element maps included):
    # A x FP64 1.281 GB, FP32 0.828 GB; J v FP64 1,737,162,272 B, FP32 1.056 GB
    n_tot, n_int, nelem, nvert = 385 ** 3, 383 ** 3, 64 ** 3, 65 ** 3
    assert bench.alg_bytes(8, n_tot, n_int, 7, nelem, nvert, True) == 1737162272
    assert abs(bench.alg_bytes(8, n_tot, n_int, 7, nelem, nvert, False) - 1.281e9) < 1e6
    assert abs(bench.alg_bytes(4, n_tot, n_int, 7, nelem, nvert, False) - 0.828e9) < 1e6
    assert abs(bench.alg_bytes(4, n_tot, n_int, 7, nelem, nvert, True) - 1.056e9) < 1e6
    # what the stored-data form streams (6 values per point): FP64 5.215 GB, FP32 2.607 GB
    assert abs(bench.stored_bytes(8, n_int, 7, nelem) - 5.2148e9) < 1e6
    assert abs(bench.stored_bytes(4, n_int, 7, nelem) - 2.6074e9) < 1e6


def test_committed_traffic_line_is_read():
    assert bench.measured_traffic("nonlinear diffusion J(u0)v, 64^3 hex, p=6, FP64, perturbed mesh") > 1e9
    assert bench.measured_traffic(
        "nonlinear diffusion J(u0)v from stored quadrature data, 64^3 hex, p=6, FP64, perturbed mesh") > 5e9


def test_headline_flop_model():
    # the three-pass contraction counted by running it: every multiply-add of each pass is counted
    import numpy as np

    def counted(P, Q, D):
        A = [np.ones((Q, P))] * 3
        u = np.ones((P, P, P, D))
        macs = 0
        a = np.zeros((Q, P, P, D))
        for i in range(Q):
            for m in range(P):
                a[i] += A[0][i, m] * u[m]
                macs += P * P * D
        b = np.zeros((Q, Q, P, D))
        for j in range(Q):
            for m in range(P):
                b[:, j] += A[1][j, m] * a[:, m]
                macs += Q * P * D
        c = np.zeros((Q, Q, Q, D))
        for k in range(Q):
            for m in range(P):
                c[:, :, k] += A[2][k, m] * b[:, :, m]
                macs += Q * Q * D
        assert c[0, 0, 0, 0] == P ** 3
        return 2 * macs

    for P, Q, D in ((2, 7, 3), (7, 7, 1), (3, 5, 2)):
        assert bench._contraction_flops(P, Q, D) == counted(P, Q, D)
    # the reference's figures at p = 6: 14,406 flops per TensorMult_Hex call, about 22.5 kflop of coordinate
    # contractions per element for the geometry
    assert bench._contraction_flops(7, 7, 1) == 14406
    assert abs(4 * bench._contraction_flops(2, 7, 3) - 22.5e3) < 0.1e3
    # J v at p = 6: six gradient passes, push-forwards, pointwise, geometry = 143,248 flops per element
    assert bench.alg_flops(7, "jv") == 6 * 14406 + 56 * 343 + 4 * bench._contraction_flops(2, 7, 3) + 44 * 343
    assert bench.alg_flops(7, "jv") == 143248
    # the reference's iterator also interpolates the direction's value and its transpose: two more contractions
    assert bench.alg_flops(7, "jv_ref") == bench.alg_flops(7, "jv") + 2 * 14406 == 172060

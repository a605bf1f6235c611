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

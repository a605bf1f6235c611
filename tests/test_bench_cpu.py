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
    # 64^3 hexahedra, p = 6 (n_int = 383^3 owned nodes, 343 Gauss points per element):
    # J v recomputed: x + y + vertices + state = 1.362 GB; linear recomputed: 0.906 GB;
    # from stored data (6 values per point): FP64 5.215 GB, FP32 2.607 GB
    n_tot, n_int, nelem, nvert = 385 ** 3, 383 ** 3, 64 ** 3, 65 ** 3
    assert abs(bench.alg_bytes(8, n_tot, n_int, 7, nelem, nvert, True) - 1.362e9) < 1e6
    assert abs(bench.alg_bytes(8, n_tot, n_int, 7, nelem, nvert, False) - 0.9055e9) < 1e6
    assert abs(bench.alg_bytes(8, n_tot, n_int, 7, nelem, nvert, True, nq=6) - 5.2148e9) < 1e6
    assert abs(bench.alg_bytes(4, n_tot, n_int, 7, nelem, nvert, True, nq=6) - 2.6074e9) < 1e6
    # the block-space kernels of the batch also read index tables
    assert bench.alg_bytes(8, n_tot, n_int, 7, nelem, nvert, False, maps=True) == 905501192 + (4 * 343 + 32) * nelem


def test_committed_traffic_line_is_read():
    assert bench.measured_traffic("nonlinear diffusion J(u0)v, 64^3 hex, p=6, FP64, perturbed mesh") > 1e9
    assert bench.measured_traffic(
        "nonlinear diffusion J(u0)v from stored quadrature data, 64^3 hex, p=6, FP64, perturbed mesh") > 5e9

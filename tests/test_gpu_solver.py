"""Krylov solve with the embedded-Q1 preconditioner: plain FP64 and FP32-inner/FP64-outer."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from dohp_amd.fields import lognormal_field  # noqa: E402
from dohp_amd.mesh import block_vertices, splitmix64_uniform  # noqa: E402


@pytest.fixture(scope="module")
def C():
    assert torch.cuda.is_available()
    import dohp_amd.capi as capi
    return capi


@pytest.mark.parametrize("M,deg", [((6, 6, 6), 5), ((4, 5, 3), 3)])
def test_solve_plain_and_mixed_reach_1e10(C, M, deg):
    print(f"solve M={M} p={deg}", flush=True)
    V = block_vertices(M, box=(0, 1, 0, 1, 0, 1), perturb=0.1, seed=3)
    fs = C.FS(M, deg, V)
    k = lognormal_field(V, seed=7)
    op = C.Ellip(fs, p=2.0, kvert=k)
    xs = torch.as_tensor(2 * splitmix64_uniform(5, fs.nowned) - 1, dtype=torch.float64, device="cuda")
    b = torch.zeros_like(xs)
    op.mat_mult(xs, b)
    q1 = op.assemble_q1(torch.zeros_like(xs))
    res = {}
    for mixed in (False, True):
        x = torch.zeros_like(xs)
        st = C.Ellip.solve(op, q1, b, x, mixed=mixed, rtol=1e-10)
        r = torch.zeros_like(xs)
        op.mat_mult(x, r)
        relres = float(torch.linalg.norm(b - r) / torch.linalg.norm(b))
        err = float(torch.linalg.norm(x - xs) / torch.linalg.norm(xs))
        print(f"mixed={mixed} stats={st} relres={relres:.2e} err={err:.2e}")
        assert relres <= 1.01e-10
        assert err < 1e-6
        res[mixed] = st
    assert res[False]["inner_its"] < 1000

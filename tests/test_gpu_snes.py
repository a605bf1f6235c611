"""Whole nonlinear solves of the reference's ellip tests on the GPU path: Newton with GMRES on the library's
residual and Jacobian action (stored quadrature data, the default), the embedded-Q1 matrix assembled by the
library at every iterate, against the SNES and KSP residual histories the reference recorded
(src/fs/tests/refout/ellip-*.refout, run as src/fs/tests/CMakeLists.txt:53-56: 4^3 block mesh, exact solution
0, -pc_type lu, KSP defaults GMRES(30) rtol 1e-5, SNES rtol 1e-10, line search basic or the default
backtracking).  The LU factorization of the Q1 matrix is the test's stand-in for PETSc's -pc_type lu (scipy,
on the host: the preconditioner of these runs is a direct solve).

Printed norms carry six significant digits; "< 1.e-11" is recorded as 0.  Every recorded norm is matched:
relative 2e-5 above 1e-6 of the first norm (rounding of the printout), 2e-3 below it (where the KSP's
inexact steps and the last digits of the arithmetic enter), and below 1e-11 where the reference printed 0.
"""
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from dohp_amd.mesh import block_vertices  # noqa: E402
from oracle import oracle as O  # noqa: E402

G = os.path.join(os.path.dirname(__file__), "golden")
HIST = json.load(open(os.path.join(G, "ellip_histories.json")))
# name: (bdeg, p, lambda, (a, b, c), line search)
CASES = {"ellip-e0-b1-p16": (1, 1.6, 0.0, (1.6, 1.7, 1.8), "basic"),
         "ellip-e0-b3-p16": (3, 1.6, 0.0, (1.6, 1.7, 1.8), "basic"),
         "ellip-e0-b3-p14-sse": (3, 1.4, 0.0, (1.6, 1.7, 1.8), "bt"),
         "ellip-e0-b3-p17-lam1-sse": (3, 1.7, 1.0, (1.0, 1.0, 1.0), "bt")}


@pytest.fixture(scope="module")
def C():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import dohp_amd.capi as capi
    return capi


def lu_pc(op, x):
    import scipy.sparse.linalg as spla
    lu = spla.splu(op.assemble_q1(x).to_scipy().tocsc())

    def M(v, out):
        out.copy_(torch.as_tensor(lu.solve(v.cpu().numpy()), device=out.device))
    return M


def match(ours, ref, scale):
    for k, (a, r) in enumerate(zip(ours, ref)):
        if r == 0.0:
            assert a < 1e-11, (k, a)
        else:
            tol = 2e-5 if r > 1e-6 * scale else 2e-3
            assert a == pytest.approx(r, rel=tol), (k, a, r)


@pytest.mark.parametrize("name", list(CASES))
def test_newton_gmres_histories_match_the_reference(C, name):
    from dohp_amd import snes
    print(name, flush=True)
    bdeg, ep, lam, (a, b, c), ls = CASES[name]
    M = (4, 4, 4)
    V = block_vertices(M)
    fs = C.FS(M, bdeg, V)
    op = C.Ellip(fs, p=ep, eps=1.0, lam=lam, exact=0, a=a, b=b, c=c)
    S = O.BlockSpace(M, bdeg, V)
    fs.dirichlet_commit(O.Ellip(S, p=ep, eps=1.0, lam=lam, exact=0, a=a, b=b, c=c).dirichlet_values())
    x = torch.zeros(fs.nowned, dtype=torch.float64, device="cuda")
    x, hist, ksp, reason = snes.newton(op, x, pc=lu_pc, linesearch=ls)
    ref = HIST[name]
    print("snes", ["%.6g" % h for h in hist])
    print("ksp lengths", [len(k) for k in ksp], "reference", [len(k) for k in ref["ksp"]])
    assert reason > 0
    assert len(hist) == len(ref["snes"])
    match(hist, ref["snes"], ref["snes"][0])
    assert len(ksp) == len(ref["ksp"])
    for k, (kh, kr) in enumerate(zip(ksp, ref["ksp"])):
        assert len(kh) == len(kr), (k, len(kh), len(kr))
        match(kh, kr, kr[0])


@pytest.mark.parametrize("q1_mode,mg_degree,mg_fp32", [(0, 2, False), (0, 2, True), (1, 1, True)])
def test_newton_with_multigrid_preconditioner(C, q1_mode, mg_degree, mg_fp32):
    """The device-only configuration: Newton (backtracking) with GMRES preconditioned by the multigrid V-cycle on
    the embedded-Q1 matrix of each iterate (reference or corrected arithmetic; FP64, or FP32 on FP16 coefficients),
    on a perturbed 6^3 mesh at p = 4 (the line search backtracks in the first steps from u = 0); converges to the
    SNES tolerance, and the discrete solution agrees with the manufactured one to discretization accuracy at the
    nodes."""
    from dohp_amd import snes
    print(f"newton q1_mode={q1_mode} mg_degree={mg_degree} mg_fp32={mg_fp32}", flush=True)
    M, bdeg = (6, 6, 6), 4
    V = block_vertices(M, box=(0, 1, 0, 1, 0, 1), perturb=0.1, seed=3)
    fs = C.FS(M, bdeg, V)
    prm = dict(p=1.6, eps=1.0, lam=0.0, exact=0, a=1.2, b=0.9, c=1.1)
    op = C.Ellip(fs, **prm)
    S = O.BlockSpace(M, bdeg, V)
    E = O.Ellip(S, **prm)
    fs.dirichlet_commit(E.dirichlet_values())
    op.set_q1_mode(q1_mode)
    x = torch.zeros(fs.nowned, dtype=torch.float64, device="cuda")
    x, hist, ksp, reason = snes.newton(op, x, pc="mg", linesearch="bt", mg_degree=mg_degree, mg_fp32=mg_fp32)
    print("snes", ["%.3g" % h for h in hist], "ksp its", [len(k) - 1 for k in ksp])
    assert reason > 0 and hist[-1] <= 1e-10 * hist[0]
    exact = E.exact_solution(S.nodal_coordinates()[:fs.nowned])
    err = np.abs(x.cpu().numpy() - exact).max() / np.abs(exact).max()
    print("max nodal error", err)  # 3.7e-9 for the same Newton iteration on the oracle with LU
    assert err < 1e-6


def test_cg_reproduces_the_morph_history(C):
    """ellip-morph (src/fs/tests/CMakeLists.txt:57: -morph 1 -twist 0.4 -snes_max_it 1 -ksp_type cg -ksp_rtol 1e-8
    -pc_type lu): the 31 preconditioned residual norms of the CG solve of the one Newton step, with the GPU
    operator on the curved, twisted mesh."""
    from dohp_amd import snes
    print("ellip-morph cg", flush=True)
    M = (4, 4, 4)
    V = block_vertices(M, morph=1.0, twist=0.4)
    fs = C.FS(M, 3, V)
    prm = dict(p=2.0, eps=1.0, lam=0.0, exact=0, a=1.0, b=1.0, c=1.0)
    op = C.Ellip(fs, **prm)
    fs.dirichlet_commit(O.Ellip(O.BlockSpace(M, 3, V), **prm).dirichlet_values())
    x = torch.zeros(fs.nowned, dtype=torch.float64, device="cuda")
    x, hist, ksp, reason = snes.newton(op, x, pc=lu_pc, max_it=1, ksp_type="cg", ksp_rtol=1e-8)
    ref = HIST["ellip-morph"]["ksp"][0]
    print("cg", ["%.6g" % h for h in ksp[0]])
    assert len(ksp[0]) == len(ref)
    match(ksp[0], ref, ref[0])

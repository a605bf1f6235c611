"""Pins of the oracle by whole nonlinear solves: Newton with GMRES (dohp_amd/snes.py) on the oracle's residual,
Jacobian action and embedded-Q1 matrix, preconditioned by an LU factorization of that matrix at every iterate,
reproduces the SNES and KSP residual histories the reference recorded for its ellip tests
(src/fs/tests/refout/ellip-*.refout; runs as src/fs/tests/CMakeLists.txt:53-56).  Every Newton iterate after the
first depends on the previous ones, so this checks the residual, the Jacobian and the Q1 matrix away from the
initial state too.  Tolerances as in tests/test_gpu_snes.py."""
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from dohp_amd import snes  # noqa: E402
from dohp_amd.mesh import block_vertices  # noqa: E402
from oracle import oracle as O  # noqa: E402

HIST = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ellip_histories.json")))
CASES = {"ellip-e0-b1-p16": (1, 1.6, 0.0, (1.6, 1.7, 1.8), "basic"),
         "ellip-e0-b3-p16": (3, 1.6, 0.0, (1.6, 1.7, 1.8), "basic"),
         "ellip-e0-b3-p14-sse": (3, 1.4, 0.0, (1.6, 1.7, 1.8), "bt"),
         "ellip-e0-b3-p17-lam1-sse": (3, 1.7, 1.0, (1.0, 1.0, 1.0), "bt")}


class OracleOp:
    """The oracle behind the operator interface the Newton layer uses (function records the state)."""

    def __init__(self, E, S, dv):
        self.E, self.S, self.dv = E, S, dv

    def function(self, x, f):
        self.uc = self.E.closure(x.numpy(), self.dv)
        f.copy_(torch.as_tensor(self.E.function(self.uc)[:self.S.nowned]))

    def mat_mult(self, v, y):
        vc = np.concatenate([v.numpy(), np.zeros(self.S.nclosure - self.S.nowned)])
        y.copy_(torch.as_tensor(self.E.jacobian_apply(self.uc, vc)[:self.S.nowned]))


def lu_pc(op, x):
    import scipy.sparse.linalg as spla
    lu = spla.splu(op.E.q1_matrix(op.E.closure(x.numpy(), op.dv)).tocsc())
    return lambda v, out: out.copy_(torch.as_tensor(lu.solve(v.numpy())))


def match(ours, ref, scale):
    for k, (a, r) in enumerate(zip(ours, ref)):
        if r == 0.0:
            assert a < 1e-11, (k, a)
        else:
            assert a == pytest.approx(r, rel=2e-5 if r > 1e-6 * scale else 2e-3), (k, a, r)


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_newton_reproduces_the_reference_histories(name):
    bdeg, ep, lam, (a, b, c), ls = CASES[name]
    M = (4, 4, 4)
    V = block_vertices(M)
    S = O.BlockSpace(M, bdeg, V)
    E = O.Ellip(S, p=ep, eps=1.0, lam=lam, exact=0, a=a, b=b, c=c)
    op = OracleOp(E, S, E.dirichlet_values())
    x, hist, ksp, reason = snes.newton(op, torch.zeros(S.nowned, dtype=torch.float64), pc=lu_pc, linesearch=ls)
    ref = HIST[name]
    assert reason > 0 and len(hist) == len(ref["snes"])
    match(hist, ref["snes"], ref["snes"][0])
    assert [len(k) for k in ksp] == [len(k) for k in ref["ksp"]]
    for kh, kr in zip(ksp, ref["ksp"]):
        match(kh, kr, kr[0])


def test_oracle_cg_reproduces_the_morph_history():
    """ellip-morph (src/fs/tests/CMakeLists.txt:57: -morph 1 -twist 0.4 -snes_max_it 1 -ksp_type cg -ksp_rtol 1e-8
    -pc_type lu): the 31 preconditioned residual norms of the one Newton step's CG solve on the curved mesh."""
    M = (4, 4, 4)
    V = block_vertices(M, morph=1.0, twist=0.4)
    S = O.BlockSpace(M, 3, V)
    E = O.Ellip(S, p=2.0, eps=1.0, lam=0.0, exact=0, a=1.0, b=1.0, c=1.0)
    op = OracleOp(E, S, E.dirichlet_values())
    x, hist, ksp, reason = snes.newton(op, torch.zeros(S.nowned, dtype=torch.float64), pc=lu_pc, max_it=1,
                                       ksp_type="cg", ksp_rtol=1e-8)
    ref = HIST["ellip-morph"]["ksp"][0]
    assert len(ksp[0]) == len(ref)
    match(ksp[0], ref, ref[0])

"""The CPU oracle against the reference's own golden outputs and against mathematics.

Fixtures in tests/golden/ are numbers printed by the reference's tests
(src/jacobi/tests/refout, src/fs/tests/refout), extracted by
tests/golden/make_fixtures.py.
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from dohp_amd.mesh import block_vertices
from oracle import oracle as O

G = os.path.join(os.path.dirname(__file__), "golden")


# ------------------------------------------------------------------ 1-D rules
@pytest.mark.parametrize("n", range(1, 13))
def test_gauss_legendre_matches_numpy(n):
    z, w = O.gauss_legendre(n)
    zr, wr = np.polynomial.legendre.leggauss(n)
    np.testing.assert_allclose(z, zr, atol=1e-14)
    np.testing.assert_allclose(w, wr, atol=1e-14)


@pytest.mark.parametrize("n", range(2, 13))
def test_gauss_lobatto_closed_form(n):
    z, w = O.gauss_lobatto(n)
    # interior nodes are roots of P'_{n-1}; weights 2/(n(n-1)P_{n-1}^2)
    P = np.polynomial.legendre.Legendre.basis(n - 1)
    np.testing.assert_allclose(P.deriv()(z[1:-1]), 0, atol=1e-12)
    np.testing.assert_allclose(w, 2.0 / (n * (n - 1) * P(z) ** 2), rtol=1e-13)
    # exact for polynomials of degree 2n-3
    for k in range(2 * n - 2):
        np.testing.assert_allclose(np.dot(w, z ** k), (1 - (-1) ** (k + 1)) / (k + 1), atol=1e-13)


@pytest.mark.parametrize("P", range(2, 12))
def test_tensor_basis_exactness(P):
    zq, _ = O.gauss_legendre(P + 1)
    B, D, node, _ = O.tensor_basis(P, zq)
    np.testing.assert_allclose(B.sum(axis=1), 1.0, atol=1e-12)
    np.testing.assert_allclose(D.sum(axis=1), 0.0, atol=1e-10)
    for k in range(P):  # polynomials of degree <= p = P-1 are reproduced exactly
        np.testing.assert_allclose(B @ node ** k, zq ** k, atol=1e-12)
        np.testing.assert_allclose(D @ node ** k, k * zq ** max(k - 1, 0) if k else 0 * zq, atol=1e-9)


# ------------------------------------------------------ jacobi-ex3 tables
def _explicit(P, rc, w1):
    Q = len(rc)
    np3 = P ** 3
    interp = np.zeros((Q ** 3, np3))
    grad = np.zeros((Q ** 3, np3 * 3))
    for i in range(np3):
        e = np.zeros(np3)
        e[i] = 1
        interp[:, i] = O.efs_apply_hex(P, rc, None, 1, e, O.APPLY_INTERP)
        grad[:, 3 * i:3 * i + 3] = O.efs_apply_hex(P, rc, None, 1, e, O.APPLY_GRAD).reshape(Q ** 3, 3)
    w = np.einsum("i,j,k->ijk", w1, w1, w1).reshape(-1)
    mass = interp.T @ (w[:, None] * interp)
    return interp, grad, mass


@pytest.mark.parametrize("family,rdeg,fname", [(O.GAUSS_LEGENDRE, 8, "jacobi3_tensor_gauss.npz"),
                                               (O.GAUSS_LOBATTO, 6, "jacobi3_tensor_lobatto.npz")])
def test_jacobi3_explicit_tables(family, rdeg, fname):
    """jacobi-3-tensor-{gauss,lobatto}: ex3.c TestExplicitBases, orders 1..4 on a 5-point rule."""
    ref = np.load(os.path.join(G, fname))
    rc, rw = O.rule_1d(O.QUAD_FAST, family, rdeg)
    for e in range(4):
        P = e + 2
        interp, grad, mass = _explicit(P, rc, rw)
        Q, Pn, rows = ref[f"e{e}_shape"]
        assert (Q, Pn) == interp.shape
        np.testing.assert_allclose(interp[:rows], ref[f"e{e}_interp"], atol=2e-6)
        np.testing.assert_allclose(grad[:rows], ref[f"e{e}_grad"], atol=2e-6)
        np.testing.assert_allclose(mass, ref[f"e{e}_mass"], atol=2e-6)


@pytest.mark.parametrize("rdeg", [4, 6])
def test_jacobi3_sparse_q1_tables(rdeg):
    """jacobi-3-tensor-sparsex-r4/r6: explicit sparse Q1 tables incl. the +-1 derivative."""
    ref = np.load(os.path.join(G, f"jacobi3_sparsex_r{rdeg}.npz"))
    rc, rw = O.rule_1d(O.QUAD_SPARSE, O.GAUSS_LEGENDRE, rdeg)
    npatch, psize, qidx, _ = O.rule_patches(len(rc), rw, True)
    for e in range(3):
        P = e + 2
        np.testing.assert_array_equal(qidx, ref[f"e{e}_qidx"])
        Pn, eidx, interp, deriv = O.explicit_sparse_hex(P, len(rc), rc, npatch, psize, qidx)
        np.testing.assert_array_equal(eidx, ref[f"e{e}_bidx"])
        for pt in range(npatch):
            np.testing.assert_allclose(interp[pt], ref[f"e{e}_interp_p{pt}"], atol=6e-6)
            np.testing.assert_allclose(deriv[pt].reshape(psize, -1), ref[f"e{e}_deriv_p{pt}"], atol=6e-6)
    # the +-1 quirk: corner derivative of the P=3 element equals -interp1^2 (efstopo.c:665-666)
    if rdeg == 4:
        assert abs(ref["e1_deriv_p0"][0, 0] - (-0.78867513459481287 ** 2)) < 1e-5


@pytest.mark.parametrize("method,fname", [(O.QUAD_FAST, "jacobi4_fast.npz"), (O.QUAD_SPARSE, "jacobi4_sparse.npz")])
def test_jacobi4_patches(method, fname):
    ref = np.load(os.path.join(G, fname))
    for e in range(4):
        rc, rw = O.rule_1d(method, O.GAUSS_LEGENDRE, 2 * (e + 1) - 2)
        npatch, psize, ind, pw = O.rule_patches(len(rc), rw, method == O.QUAD_SPARSE)
        assert [npatch, psize] == list(ref[f"e{e}_np"])
        np.testing.assert_array_equal(ind, ref[f"e{e}_indices"])
        np.testing.assert_allclose(pw, ref[f"e{e}_weights"], atol=6e-6)


# ------------------------------------------------------- element invariants
@pytest.mark.parametrize("P,Q,D", [(2, 2, 1), (3, 3, 1), (3, 4, 3), (4, 3, 2), (5, 5, 1)])
def test_efs_apply_brute_force_kronecker(P, Q, D):
    rng = np.random.default_rng(P * 100 + Q * 10 + D)
    rc, _ = O.gauss_legendre(Q)
    B, Dm, _, _ = O.tensor_basis(P, rc)
    x = rng.uniform(-1, 1, P ** 3 * D)
    Kb = np.kron(np.kron(B, B), B)
    X = x.reshape(P ** 3, D)
    np.testing.assert_allclose(O.efs_apply_hex(P, rc, None, D, x, O.APPLY_INTERP).reshape(-1, D), Kb @ X, atol=1e-13)
    G3 = [np.kron(np.kron(Dm, B), B), np.kron(np.kron(B, Dm), B), np.kron(np.kron(B, B), Dm)]
    jinv = rng.uniform(-1, 1, (Q ** 3, 3, 3))
    ref = np.einsum("kqd,qke->qde", np.stack([g @ X for g in G3]), jinv)
    got = O.efs_apply_hex(P, rc, jinv, D, x, O.APPLY_GRAD).reshape(Q ** 3, D, 3)
    np.testing.assert_allclose(got, ref, atol=1e-12)
    # adjointness <A x, y> = <x, A^T y> for INTERP and GRAD
    y = rng.uniform(-1, 1, Q ** 3 * D)
    np.testing.assert_allclose(np.dot(O.efs_apply_hex(P, rc, None, D, x, O.APPLY_INTERP), y),
                               np.dot(x, O.efs_apply_hex(P, rc, None, D, y, O.APPLY_INTERP_TRANSPOSE)), rtol=1e-12)
    yg = rng.uniform(-1, 1, Q ** 3 * D * 3)
    np.testing.assert_allclose(np.dot(got.reshape(-1), yg),
                               np.dot(x, O.efs_apply_hex(P, rc, jinv, D, yg, O.APPLY_GRAD_TRANSPOSE)), rtol=1e-11)
    # ADD_VALUES accumulates
    base = rng.uniform(-1, 1, Q ** 3 * D)
    np.testing.assert_allclose(O.efs_apply_hex(P, rc, None, D, x, O.APPLY_INTERP, O.ADD_VALUES, out=base),
                               base + Kb.dot(X).reshape(-1), atol=1e-13)


# ------------------------------------------------------------ ellip goldens
HIST = json.load(open(os.path.join(G, "ellip_histories.json")))


def _ellip(bdeg, ep, lam, abc, morph=0.0, twist=0.0, exact=0):
    V = block_vertices((4, 4, 4), morph=morph, twist=twist)
    S = O.BlockSpace((4, 4, 4), bdeg, V)
    E = O.Ellip(S, p=ep, eps=1.0, lam=lam, exact=exact, a=abc[0], b=abc[1], c=abc[2])
    return S, E, E.dirichlet_values()


CASES = {"ellip-e0-b1-p16": (1, 1.6, 0.0, (1.6, 1.7, 1.8)), "ellip-e0-b3-p16": (3, 1.6, 0.0, (1.6, 1.7, 1.8)),
         "ellip-e0-b3-p14-sse": (3, 1.4, 0.0, (1.6, 1.7, 1.8)), "ellip-e0-b3-p17-lam1-sse": (3, 1.7, 1.0, (1, 1, 1))}


@pytest.mark.parametrize("name", list(CASES))
def test_ellip_initial_residual_and_q1_preconditioned_norm(name):
    S, E, d = _ellip(*CASES[name])
    uc = E.closure(np.zeros(S.nowned), d)
    F = E.function(uc)[:S.nowned]
    ref = HIST[name]
    assert np.linalg.norm(F) == pytest.approx(ref["snes"][0], rel=6e-6)
    Jp = E.q1_matrix(uc)
    r0 = spla.splu(Jp.tocsc()).solve(F)
    assert np.linalg.norm(r0) == pytest.approx(ref["ksp"][0][0], rel=6e-6)


def test_ellip_b1_newton_history_pins_jacobian():
    """p=1: the Q1 matrix is exactly 4x the Jacobian (the +-1 quirk), so one GMRES step is exact."""
    S, E, d = _ellip(*CASES["ellip-e0-b1-p16"])
    ref = HIST["ellip-e0-b1-p16"]
    x = np.zeros(S.nowned)
    for it in range(4):
        uc = E.closure(x, d)
        F = E.function(uc)[:S.nowned]
        assert np.linalg.norm(F) == pytest.approx(ref["snes"][it], rel=2e-5)
        r0 = spla.splu(E.q1_matrix(uc).tocsc()).solve(F)
        assert np.linalg.norm(r0) == pytest.approx(ref["ksp"][it][0], rel=2e-5)
        Jr = E.jacobian_apply(uc, np.concatenate([r0, np.zeros(len(d))]))[:S.nowned]
        np.testing.assert_allclose(Jr, F / 4, atol=1e-13 * np.linalg.norm(F))
        x = x - 4 * r0


def _gmres_left(A, M, b, rtol, restart=30, maxit=200):
    """Left-preconditioned GMRES(restart) reporting preconditioned residual norms (PETSc KSPGMRES)."""
    n = b.size
    x = np.zeros(n)
    hist = []
    r = M(b)
    beta = np.linalg.norm(r)
    r0 = beta
    hist.append(beta)
    it = 0
    while it < maxit:
        V = np.zeros((restart + 1, n))
        H = np.zeros((restart + 1, restart))
        V[0] = r / beta
        g = np.zeros(restart + 1)
        g[0] = beta
        cs, sn = np.zeros(restart), np.zeros(restart)
        for j in range(restart):
            w = M(A(V[j]))
            for i in range(j + 1):
                H[i, j] = np.dot(w, V[i])
                w = w - H[i, j] * V[i]
            H[j + 1, j] = np.linalg.norm(w)
            V[j + 1] = w / H[j + 1, j]
            for i in range(j):
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = t
            den = np.hypot(H[j, j], H[j + 1, j])
            cs[j], sn[j] = H[j, j] / den, H[j + 1, j] / den
            H[j, j] = den
            H[j + 1, j] = 0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            it += 1
            hist.append(abs(g[j + 1]))
            if abs(g[j + 1]) <= rtol * r0:
                y = np.linalg.solve(np.triu(H[:j + 1, :j + 1]), g[:j + 1])
                return x + V[:j + 1].T @ y, hist
        y = np.linalg.solve(np.triu(H[:restart, :restart]), g[:restart])
        x = x + V[:restart].T @ y
        r = M(b - A(x))
        beta = np.linalg.norm(r)
    return x, hist


def test_ellip_b3_gmres_and_newton_step_pin_high_order_jacobian():
    """ellip-e0-b3-p16: GMRES(30) with LU(Q1), rtol 1e-5, basic line search (the refout's run)."""
    S, E, d = _ellip(*CASES["ellip-e0-b3-p16"])
    ref = HIST["ellip-e0-b3-p16"]
    x = np.zeros(S.nowned)
    nd = len(d)
    for it in range(2):
        uc = E.closure(x, d)
        F = E.function(uc)[:S.nowned]
        assert np.linalg.norm(F) == pytest.approx(ref["snes"][it], rel=5e-5)
        lu = spla.splu(E.q1_matrix(uc).tocsc())
        A = lambda v, uc=uc: E.jacobian_apply(uc, np.concatenate([v, np.zeros(nd)]))[:S.nowned]
        dx, hist = _gmres_left(A, lu.solve, F, 1e-5)
        kref = ref["ksp"][it]
        assert len(hist) == len(kref)
        np.testing.assert_allclose(hist, kref, rtol=2e-4)
        x = x - dx
    uc = E.closure(x, d)
    assert np.linalg.norm(E.function(uc)[:S.nowned]) == pytest.approx(ref["snes"][2], rel=5e-4)


def test_ellip_morph_q1_pins_element_ordered_inverse_jacobian():
    """ellip-morph (-morph 1 -twist 0.4): only the element-ordered cjinv gives 111.034."""
    S, E, d = _ellip(3, 2.0, 0.0, (1, 1, 1), morph=1.0, twist=0.4)
    uc = E.closure(np.zeros(S.nowned), d)
    F = E.function(uc)[:S.nowned]
    ref = HIST["ellip-morph"]["ksp"][0]
    r0 = spla.splu(E.q1_matrix(uc).tocsc()).solve(F)
    assert np.linalg.norm(r0) == pytest.approx(ref[0], rel=6e-6)
    O.lib().orc_set_q1_cjinv_patch_order(1)
    try:
        r1 = spla.splu(E.q1_matrix(uc).tocsc()).solve(F)
    finally:
        O.lib().orc_set_q1_cjinv_patch_order(0)
    # the patch-ordered variant differs by ~3.5e-4 relative, far above the printed 6 digits
    assert abs(np.linalg.norm(r1) - ref[0]) > 2e-5 * ref[0]


def test_gather_scatter_sum_invariant():
    """E^T 1 counts element multiplicity: sum = expanded size (fs/tests/ex1.c:226-236 pattern)."""
    V = block_vertices((3, 2, 2))
    S = O.BlockSpace((3, 2, 2), 4, V, dirichlet_mask=0)
    cnt = np.bincount(S.elem_nodes.reshape(-1), minlength=S.nclosure)
    assert cnt.sum() == S.nelem * 125
    assert cnt.max() == 8 and cnt.min() == 1


NORM_CASES = {"ellip-e0-b1-p16": CASES["ellip-e0-b1-p16"], "ellip-e0-b3-p16": CASES["ellip-e0-b3-p16"],
              "ellip-e0-b3-p14-sse": CASES["ellip-e0-b3-p14-sse"],
              "ellip-e0-b3-p17-lam1-sse": CASES["ellip-e0-b3-p17-lam1-sse"]}


@pytest.mark.parametrize("name", list(NORM_CASES))
def test_ellip_converged_error_norms(name):
    """Newton to the discrete solution, then the printed pointwise solution/gradient error norms
    (util.c:110-168 through ellip.c:538-564), three significant digits in the refout."""
    S, E, d = _ellip(*NORM_CASES[name])
    nd = len(d)
    x = np.zeros(S.nowned)
    fn = lambda x: E.function(E.closure(x, d))[:S.nowned]
    F = fn(x)
    for it in range(40):
        if np.linalg.norm(F) < 1e-11:
            break
        uc = E.closure(x, d)
        lu = spla.splu(E.q1_matrix(uc).tocsc())
        A = lambda v, uc=uc: E.jacobian_apply(uc, np.concatenate([v, np.zeros(nd)]))[:S.nowned]
        dx, _ = _gmres_left(A, lu.solve, F, 1e-12, maxit=400)
        lam = 1.0  # backtracking on ||F|| (the discrete solution does not depend on the path)
        while True:
            Fn = fn(x - lam * dx)
            if np.linalg.norm(Fn) <= (1 - 1e-4 * lam) * np.linalg.norm(F) or lam < 1e-3:
                break
            lam /= 2
        x, F = x - lam * dx, Fn
    assert np.linalg.norm(F) < 1e-10
    en, gn = E.error_norms(E.closure(x, d))
    ref = HIST[name]["error_norms"]
    np.testing.assert_allclose(en, ref["solution"], rtol=5e-3)
    np.testing.assert_allclose(gn, ref["gradient"], rtol=5e-3)


def test_ellip_morph_linear_solve_error_norms():
    S, E, d = _ellip(3, 2.0, 0.0, (1, 1, 1), morph=1.0, twist=0.4)
    uc = E.closure(np.zeros(S.nowned), d)
    F = E.function(uc)[:S.nowned]
    lu = spla.splu(E.q1_matrix(uc).tocsc())
    A = lambda v: E.jacobian_apply(uc, np.concatenate([v, np.zeros(len(d))]))[:S.nowned]
    dx, _ = _gmres_left(A, lu.solve, F, 1e-12, maxit=400)
    en, gn = E.error_norms(E.closure(-dx, d))
    ref = HIST["ellip-morph"]["error_norms"]
    np.testing.assert_allclose(en, ref["solution"], rtol=5e-3)
    np.testing.assert_allclose(gn, ref["gradient"], rtol=5e-3)


def test_rule_compute_physical_closed_form():
    """orc_rule_compute_physical (jacobi.c:383-414 with dGeomInvert3, dohpgeom.h:57-75) against numpy's inverse
    and determinant of non-symmetric matrices (a transposed or sign-flipped cofactor fails), jw = w det, and the
    negative-determinant error."""
    import ctypes
    rng = np.random.default_rng(7)
    n = 40
    jac = rng.uniform(-1, 1, (n, 3, 3)) + 2.5 * np.eye(3)
    w = rng.uniform(0.1, 1.0, n)
    jinv, jw = np.zeros((n, 3, 3)), np.zeros(n)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = O.lib().orc_rule_compute_physical(n, w.ctypes.data_as(dp), jac.ctypes.data_as(dp), jinv.ctypes.data_as(dp),
                                           jw.ctypes.data_as(dp))
    assert rc == 0
    np.testing.assert_allclose(jinv, np.linalg.inv(jac), rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(jw, w * np.linalg.det(jac), rtol=1e-13)
    bad = jac.copy()
    bad[5] = -bad[5]  # det < 0
    assert O.lib().orc_rule_compute_physical(n, w.ctypes.data_as(dp), bad.ctypes.data_as(dp), jinv.ctypes.data_as(dp),
                                             jw.ctypes.data_as(dp)) == 2

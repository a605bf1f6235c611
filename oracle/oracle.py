"""TEST INFRASTRUCTURE -- ctypes wrapper around the CPU oracle (oracle/liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module, and only as the checker / timed CPU baseline.  It also builds,
independently of the product library, the element->closure numbering of a
structured block (the E matrix of src/fs/impls/cont/cont.c:200-249 filled by
src/jacobi/impls/tensor/tensor.c:360-468 for a conforming mesh) and the
Dirichlet data of ellip (src/fs/tests/ellip.c:566-592, fsgeom.c:72-109).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

APPLY_INTERP, APPLY_INTERP_TRANSPOSE, APPLY_GRAD, APPLY_GRAD_TRANSPOSE = 40001, 40002, 40003, 40004
INSERT_VALUES, ADD_VALUES = 1, 2
QUAD_FAST, QUAD_SPARSE, QUAD_SELF = 0, 1, 2
GAUSS_LEGENDRE, GAUSS_LOBATTO = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class _Problem(ctypes.Structure):
    _fields_ = [("nelem", ctypes.c_int), ("p", ctypes.c_int), ("nclosure", ctypes.c_long),
                ("nowned", ctypes.c_long), ("elem_nodes", _ip), ("elem_vx", _dp),
                ("ep", ctypes.c_double), ("eps", ctypes.c_double), ("lam", ctypes.c_double),
                ("exact", ctypes.c_int), ("ea", ctypes.c_double), ("eb", ctypes.c_double),
                ("ec", ctypes.c_double), ("fmethod", ctypes.c_int), ("nthreads", ctypes.c_int),
                ("kvert", _dp)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "-B"], check=True)


def lib():
    if not hasattr(lib, "_h"):
        if not os.path.exists(_LIB):
            build()
        h = ctypes.CDLL(_LIB)
        h.orc_ellip_q1_assemble.restype = ctypes.c_long
        lib._h = h
    return lib._h


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


# ---------------------------------------------------------------- 1-D tables
def gauss_legendre(n):
    z, w = np.zeros(n), np.zeros(n)
    assert lib().orc_gauss_legendre(n, _p(z), _p(w)) == 0
    return z, w


def gauss_lobatto(n):
    z, w = np.zeros(n), np.zeros(n)
    assert lib().orc_gauss_lobatto(n, _p(z), _p(w)) == 0
    return z, w


def rule_1d(method, family, order):
    sz = ctypes.c_int()
    z, w = np.zeros(order + 4), np.zeros(order + 4)
    assert lib().orc_rule_1d(method, family, order, ctypes.byref(sz), _p(z), _p(w)) == 0
    return z[:sz.value].copy(), w[:sz.value].copy()


def tensor_basis(P, rcoord):
    rcoord = np.ascontiguousarray(rcoord, dtype=np.float64)
    Q = rcoord.size
    B, D, BT, DT = (np.zeros(P * Q) for _ in range(4))
    node, weight = np.zeros(P), np.zeros(P)
    assert lib().orc_tensor_basis(P, Q, _p(rcoord), _p(B), _p(D), _p(BT), _p(DT), _p(node), _p(weight)) == 0
    return B.reshape(Q, P), D.reshape(Q, P), node, weight


def rule_patches(Q1d, w1d, sparse):
    npatch, psize = ctypes.c_int(), ctypes.c_int()
    ind = np.zeros(Q1d ** 3, dtype=np.int32)
    pw = np.zeros(Q1d ** 3)
    w1d = np.ascontiguousarray(w1d, dtype=np.float64)
    assert lib().orc_rule_patches(Q1d, _p(w1d), int(sparse), ctypes.byref(npatch), ctypes.byref(psize),
                                  _p(ind, _ip), _p(pw)) == 0
    return npatch.value, psize.value, ind.reshape(npatch.value, psize.value), pw.reshape(npatch.value, psize.value)


def efs_apply_hex(P, rcoord, jinv, D, x, amode, imode=INSERT_VALUES, out=None):
    rcoord = np.ascontiguousarray(rcoord, dtype=np.float64)
    Q = rcoord.size
    nout = {APPLY_INTERP: Q ** 3 * D, APPLY_GRAD: Q ** 3 * D * 3,
            APPLY_INTERP_TRANSPOSE: P ** 3 * D, APPLY_GRAD_TRANSPOSE: P ** 3 * D}[amode]
    out = np.zeros(nout) if out is None else np.ascontiguousarray(out, dtype=np.float64).copy()
    x = np.ascontiguousarray(x, dtype=np.float64)
    jp = _p(np.ascontiguousarray(jinv, dtype=np.float64)) if jinv is not None else None
    rc = lib().orc_efs_apply_hex(P, Q, _p(rcoord), jp, D, _p(x), _p(out), amode, imode)
    assert rc == 0, rc
    return out


def explicit_sparse_hex(P, Q1d, rcoord, npatches, Q, qidx, cjinv=None):
    rcoord = np.ascontiguousarray(rcoord, dtype=np.float64)
    qidx = np.ascontiguousarray(qidx, dtype=np.int32)
    Pmax = P ** 3
    eidx = np.zeros(npatches * Pmax, dtype=np.int32)
    interp = np.zeros(npatches * Q * Pmax)
    deriv = np.zeros(npatches * Q * Pmax * 3)
    Pout = ctypes.c_int()
    cj = _p(np.ascontiguousarray(cjinv, dtype=np.float64)) if cjinv is not None else None
    rc = lib().orc_explicit_sparse_hex(P, Q1d, _p(rcoord), npatches, Q, _p(qidx, _ip), cj, 0, ctypes.byref(Pout),
                                      _p(eidx, _ip), _p(interp), _p(deriv))
    assert rc == 0, rc
    Pn = Pout.value
    return (Pn, eidx[:npatches * Pn].reshape(npatches, Pn), interp[:npatches * Q * Pn].reshape(npatches, Q, Pn),
            deriv[:npatches * Q * Pn * 3].reshape(npatches, Q, Pn, 3))


# ------------------------------------------------------- structured space
class BlockSpace:
    """Conforming continuous space of degree p on an M0 x M1 x M2 block (oracle-side numbering).

    Closure numbering: owned (non-Dirichlet) lattice nodes first, then Dirichlet nodes,
    each in lexicographic (I, J, K) order with K fastest.
    """

    def __init__(self, M, p, vcoords, dirichlet_mask=7):
        self.M = tuple(int(m) for m in M)
        self.p = int(p)
        n = self.p + 1
        M0, M1, M2 = self.M
        L = [self.p * m + 1 for m in self.M]
        self.L = L
        V = np.asarray(vcoords, dtype=np.float64).reshape(M0 + 1, M1 + 1, M2 + 1, 3)
        I, J, K = np.meshgrid(np.arange(L[0]), np.arange(L[1]), np.arange(L[2]), indexing="ij")
        dmask = np.zeros(I.shape, dtype=bool)
        if dirichlet_mask & 1:
            dmask |= (I == 0) | (I == L[0] - 1) | (J == 0) | (J == L[1] - 1)
        if dirichlet_mask & 2:
            dmask |= K == L[2] - 1
        if dirichlet_mask & 4:
            dmask |= K == 0
        dmask = dmask.reshape(-1)
        self.nowned = int((~dmask).sum())
        self.nclosure = dmask.size
        l2c = np.empty(dmask.size, dtype=np.int64)
        l2c[~dmask] = np.arange(self.nowned)
        l2c[dmask] = self.nowned + np.arange(dmask.size - self.nowned)
        self.l2c = l2c
        self.dirichlet = dmask
        e0, e1, e2 = np.meshgrid(np.arange(M0), np.arange(M1), np.arange(M2), indexing="ij")
        e0, e1, e2 = e0.reshape(-1), e1.reshape(-1), e2.reshape(-1)
        a = np.arange(n)
        A, B, C = np.meshgrid(a, a, a, indexing="ij")
        LI = (e0[:, None] * self.p + A.reshape(-1)[None, :])
        LJ = (e1[:, None] * self.p + B.reshape(-1)[None, :])
        LK = (e2[:, None] * self.p + C.reshape(-1)[None, :])
        self.elem_nodes = np.ascontiguousarray(l2c[(LI * L[1] + LJ) * L[2] + LK].astype(np.int32))
        t = np.array([(ta, tb, tc) for ta in (0, 1) for tb in (0, 1) for tc in (0, 1)])
        self.elem_vx = np.ascontiguousarray(
            V[e0[:, None] + t[None, :, 0], e1[:, None] + t[None, :, 1], e2[:, None] + t[None, :, 2]])
        self.nelem = e0.size
        # lattice coordinate per closure node (for numbering-invariant comparisons)
        self.closure_lattice = np.empty((self.nclosure, 3), dtype=np.int64)
        self.closure_lattice[l2c] = np.stack([I.reshape(-1), J.reshape(-1), K.reshape(-1)], axis=1)

    def nodal_coordinates(self):
        """Closure nodal coordinates averaged over elements (fsgeom.c:72-109)."""
        n = self.p + 1
        z, _ = gauss_lobatto(n)
        lo, hi = (1 - z) / 2, (1 + z) / 2
        N1 = np.stack([lo, hi], axis=1)  # [n, 2]
        N = np.einsum("ax,by,cz->abcxyz", N1, N1, N1).reshape(n ** 3, 8)
        X = np.einsum("qt,etd->eqd", N, self.elem_vx)
        acc = np.zeros((self.nclosure, 3))
        cnt = np.zeros(self.nclosure)
        np.add.at(acc, self.elem_nodes.reshape(-1), X.reshape(-1, 3))
        np.add.at(cnt, self.elem_nodes.reshape(-1), 1.0)
        return acc / cnt[:, None]


class Ellip:
    """Oracle for ellip's residual, Jacobian action and embedded-Q1 matrix on a BlockSpace."""

    def __init__(self, space: BlockSpace, p=2.0, eps=1.0, lam=0.0, exact=0, a=1.0, b=1.0, c=1.0,
                 fmethod=QUAD_FAST, nthreads=0, kvert=None):
        self.space = space
        self._keep = [space.elem_nodes, space.elem_vx]
        self.kvert = None if kvert is None else np.ascontiguousarray(kvert, dtype=np.float64)
        self.pb = _Problem(space.nelem, space.p, space.nclosure, space.nowned,
                           _p(space.elem_nodes, _ip), _p(space.elem_vx), float(p), float(eps), float(lam),
                           int(exact), float(a), float(b), float(c), int(fmethod), int(nthreads),
                           _p(self.kvert) if self.kvert is not None else None)

    def exact_solution(self, X):
        X = np.asarray(X, dtype=np.float64)
        u = np.zeros(len(X))
        du = np.zeros(3)
        for i, x in enumerate(X):
            xx = np.ascontiguousarray(x)
            uu = ctypes.c_double()
            lib().orc_exact_solution(ctypes.byref(self.pb), _p(xx), ctypes.byref(uu), _p(du))
            u[i] = uu.value
        return u

    def dirichlet_values(self):
        """Exact solution at averaged nodal coordinates of the Dirichlet nodes (ellip.c:717-723)."""
        X = self.space.nodal_coordinates()[self.space.nowned:]
        return self.exact_solution(X)

    def closure(self, g, dvals):
        return np.concatenate([np.asarray(g, dtype=np.float64), np.asarray(dvals, dtype=np.float64)])

    def function(self, uc):
        uc = np.ascontiguousarray(uc, dtype=np.float64)
        yc = np.zeros(self.space.nclosure)
        rc = lib().orc_ellip_function(ctypes.byref(self.pb), _p(uc), _p(yc))
        assert rc == 0, rc
        return yc

    def jacobian_apply(self, uc, xc):
        uc = np.ascontiguousarray(uc, dtype=np.float64)
        xc = np.ascontiguousarray(xc, dtype=np.float64)
        yc = np.zeros(self.space.nclosure)
        rc = lib().orc_ellip_jacobian_apply(ctypes.byref(self.pb), _p(uc), _p(xc), _p(yc))
        assert rc == 0, rc
        return yc

    def q1_matrix(self, uc, qmethod=QUAD_SPARSE):
        import scipy.sparse as sp
        uc = np.ascontiguousarray(uc, dtype=np.float64)
        need = lib().orc_ellip_q1_assemble(ctypes.byref(self.pb), _p(uc), qmethod, 0, None, None, None)
        assert need != -1
        cap = abs(need)
        rows, cols, vals = np.zeros(cap, np.int32), np.zeros(cap, np.int32), np.zeros(cap)
        got = lib().orc_ellip_q1_assemble(ctypes.byref(self.pb), _p(uc), qmethod, cap, _p(rows, _ip),
                                          _p(cols, _ip), _p(vals))
        assert got == cap, (got, cap)
        n = self.space.nowned
        return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))

    def error_norms(self, uc):
        uc = np.ascontiguousarray(uc, dtype=np.float64)
        en, gn = np.zeros(3), np.zeros(3)
        assert lib().orc_ellip_error_norms(ctypes.byref(self.pb), _p(uc), _p(en), _p(gn)) == 0
        return en, gn


def expanded_poisson_apply(degrees, elem_vx, x, kvert=None, nthreads=0):
    """K_e x_e for independent hexahedra of per-element degree (element slices consecutive,
    offsets sum (degree+1)^3): the Jacobian action of ellip at p = 2 (ellip.c:442-472) on
    one-element spaces without Dirichlet nodes.  elem_vx [nelem][8][3], kvert [nelem][8]."""
    d = np.ascontiguousarray(degrees, dtype=np.int32)
    v = np.ascontiguousarray(elem_vx, dtype=np.float64)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    assert v.size == 24 * d.size and xx.size == int(((d + 1) ** 3).sum())
    k = None if kvert is None else np.ascontiguousarray(kvert, dtype=np.float64)
    y = np.zeros_like(xx)
    rc = lib().orc_expanded_poisson_apply(int(d.size), _p(d, _ip), _p(v), _p(k) if k is not None else None,
                                          _p(xx), _p(y), int(nthreads))
    assert rc == 0, rc
    return y

"""TEST INFRASTRUCTURE, NOT PRODUCT CODE.

numpy restatement of the embedded-Q1 matrix of the linear operator (ellip at p = 2, lambda = 0, coefficient k)
with learned per-sub-cell derivative scales, of the gradient of its bilinear form in those scales, and of the
sub-cell features the network reads.  The reference arithmetic follows EllipJacobian (src/fs/tests/ellip.c:474-536)
with dEFSGetExplicitSparse_Tensor_Hex (src/jacobi/impls/tensor/efstopo.c:660-717): sub-cell derivatives +-1 per
element-reference unit, interpolants (1 +- 1/sqrt 3)/2 at the 2x2x2 Gauss points of each GLL sub-cell, the inverse
Jacobian of the element-natural point patch*8+q (src/fs/interface/fsrulesetit.c:684).  With every scale 1 the
matrix is the oracle's orc_ellip_q1_assemble (which the refout norms pin); the tests check that equality.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import oracle as O

_I1 = np.array([[0.78867513459481287, 0.21132486540518713], [0.21132486540518713, 0.78867513459481287]])
_D1 = np.array([[-1.0, 1.0], [-1.0, 1.0]])


def _jac(X, x0, x1, x2):
    """Jacobian dx/dX of the trilinear map of vertices X [8][3] (tensor order) at reference point (x0, x1, x2)."""
    f = [[0.5 * (1 - x0), 0.5 * (1 + x0)], [0.5 * (1 - x1), 0.5 * (1 + x1)], [0.5 * (1 - x2), 0.5 * (1 + x2)]]
    J = np.zeros((3, 3))
    for t in range(8):
        a, b, c = t >> 2, (t >> 1) & 1, t & 1
        s = [(0.5 if a else -0.5) * f[1][b] * f[2][c], f[0][a] * (0.5 if b else -0.5) * f[2][c],
             f[0][a] * f[1][b] * (0.5 if c else -0.5)]
        J += np.outer(X[t], s)
    return J


class ScaledQ1:
    """Per-sub-cell pieces of the embedded-Q1 matrix of a BlockSpace: K_cell(sigma) = sum_q W_q D_q D_q^T with
    D_q = sum_a sigma_a U_qa.  kvert: [nelem][8] coefficient at the element vertices (None = 1)."""

    def __init__(self, space, kvert=None):
        self.space = S = space
        p, n1 = S.p, S.p + 1
        zg, _ = O.gauss_lobatto(n1)
        self.h = np.diff(zg)
        zq, wq = O.rule_1d(O.QUAD_SPARSE, O.GAUSS_LEGENDRE, 2 * p)
        NQ, P3 = 2 * p, p ** 3
        nc = S.nelem * P3
        self.U = np.zeros((nc, 8, 3, 8, 3))
        self.W = np.zeros((nc, 8))
        self.gid = np.zeros((nc, 8), dtype=np.int64)
        self.hsub = np.zeros((nc, 3))
        for e in range(S.nelem):
            X = S.elem_vx[e]
            ke = np.ones(8) if kvert is None else np.asarray(kvert)[e]
            for pidx in range(P3):
                pa, pb, pc = pidx // (p * p), (pidx // p) % p, pidx % p
                ci = e * P3 + pidx
                self.hsub[ci] = [self.h[pa], self.h[pb], self.h[pc]]
                for q in range(8):
                    ii, jj, kk = q >> 2, (q >> 1) & 1, q & 1
                    qi, qj, qk = 2 * pa + ii, 2 * pb + jj, 2 * pc + kk
                    det = np.linalg.det(_jac(X, zq[qi], zq[qj], zq[qk]))
                    qe = pidx * 8 + q
                    ei, ej, ek = qe // (NQ * NQ), (qe // NQ) % NQ, qe % NQ
                    cj = np.linalg.inv(_jac(X, zq[ei], zq[ej], zq[ek]))
                    f = [[0.5 * (1 - zq[qi]), 0.5 * (1 + zq[qi])], [0.5 * (1 - zq[qj]), 0.5 * (1 + zq[qj])],
                         [0.5 * (1 - zq[qk]), 0.5 * (1 + zq[qk])]]
                    kq = sum(ke[t] * f[0][t >> 2] * f[1][(t >> 1) & 1] * f[2][t & 1] for t in range(8))
                    self.W[ci, q] = wq[qi] * wq[qj] * wq[qk] * det * kq
                    for b in range(8):
                        a0, a1, a2 = b >> 2, (b >> 1) & 1, b & 1
                        dd = [_D1[ii][a0] * _I1[jj][a1] * _I1[kk][a2], _I1[ii][a0] * _D1[jj][a1] * _I1[kk][a2],
                              _I1[ii][a0] * _I1[jj][a1] * _D1[kk][a2]]
                        for a in range(3):
                            self.U[ci, q, a, b] = dd[a] * cj[a]
                for b in range(8):
                    a0, a1, a2 = b >> 2, (b >> 1) & 1, b & 1
                    self.gid[ci, b] = S.elem_nodes[e, ((pa + a0) * n1 + pb + a1) * n1 + pc + a2]

    def matrix(self, sig):
        """CSR matrix over the owned nodes for scales sig [ncell][3]."""
        n = self.space.nowned
        D = np.einsum("ca,cqabd->cqbd", sig, self.U)
        K = np.einsum("cq,cqrd,cqsd->crs", self.W, D, D)
        r = np.broadcast_to(self.gid[:, :, None], K.shape)
        c = np.broadcast_to(self.gid[:, None, :], K.shape)
        m = (r < n) & (c < n)
        return sp.csr_matrix((K[m], (r[m], c[m])), shape=(n, n))

    def scale_gradient(self, sig, w, z):
        """d (w^T K(sig) z) / d sig [ncell][3] (w, z over the owned nodes)."""
        n = self.space.nowned
        wl = np.where(self.gid < n, np.concatenate([w, [0.0]])[np.minimum(self.gid, n)], 0.0)
        zl = np.where(self.gid < n, np.concatenate([z, [0.0]])[np.minimum(self.gid, n)], 0.0)
        Uw = np.einsum("cqabd,cb->cqad", self.U, wl)
        Uz = np.einsum("cqabd,cb->cqad", self.U, zl)
        Dw = np.einsum("ca,cqad->cqd", sig, Uw)
        Dz = np.einsum("ca,cqad->cqd", sig, Uz)
        return np.einsum("cq,cqad,cqd->ca", self.W, Uw, Dz) + np.einsum("cq,cqd,cqad->ca", self.W, Dw, Uz)

    def features(self, kvert=None):
        """[ncell][11]: log h of the three sub-intervals, the sub-cell centre, log physical lengths of the reference
        directions at the centre minus their mean, log det J minus their sum, log k at the centre minus the mean log
        k of the element's vertices."""
        S = self.space
        p = S.p
        zg, _ = O.gauss_lobatto(p + 1)
        P3 = p ** 3
        F = np.zeros((S.nelem * P3, 11))
        for e in range(S.nelem):
            X = S.elem_vx[e]
            for pidx in range(P3):
                pa, pb, pc = pidx // (p * p), (pidx // p) % p, pidx % p
                c = [0.5 * (zg[pa] + zg[pa + 1]), 0.5 * (zg[pb] + zg[pb + 1]), 0.5 * (zg[pc] + zg[pc + 1])]
                J = _jac(X, *c)
                ln = 0.5 * np.log((J ** 2).sum(0))
                lk = 0.0
                if kvert is not None:
                    ke = np.asarray(kvert)[e]
                    f = [[0.5 * (1 - c[0]), 0.5 * (1 + c[0])], [0.5 * (1 - c[1]), 0.5 * (1 + c[1])],
                         [0.5 * (1 - c[2]), 0.5 * (1 + c[2])]]
                    kc = sum(ke[t] * f[0][t >> 2] * f[1][(t >> 1) & 1] * f[2][t & 1] for t in range(8))
                    lk = np.log(kc) - np.log(ke).mean()
                F[e * P3 + pidx] = [np.log(self.h[pa]), np.log(self.h[pb]), np.log(self.h[pc]), *c, *(ln - ln.mean()),
                                    np.log(np.linalg.det(J)) - 3 * ln.mean(), lk]
        return F

"""The numpy restatement of the embedded-Q1 matrix with learned sub-cell scales (oracle/q1_scaled.py): at unit
scales it equals the C oracle's reference-arithmetic assembly (pinned by the refout norms), the corrected scales
1/h make it integrate the energy of linear functions exactly on affine elements of degree 1 (where the element-natural
inverse Jacobian equals the pointwise one), its scale gradient matches finite differences, and the network oracle's
kind-1 gradient matches finite differences."""
import numpy as np

from dohp_amd.mesh import block_vertices, element_vertex_index
from oracle import mlp_oracle as L
from oracle import oracle as O
from oracle.q1_scaled import ScaledQ1


def _space(M=(2, 2, 2), p=3, perturb=0.1):
    V = block_vertices(M, box=(0, 1, 0, 1, 0, 1), perturb=perturb, seed=7)
    k = np.exp(0.3 * np.sin(3 * V[:, 0]) * np.cos(2 * V[:, 1]))
    kel = k[element_vertex_index(M)]
    S = O.BlockSpace(M, p, V)
    return S, kel


def test_unit_scales_equal_reference_assembly():
    for p in (2, 3, 5):
        S, kel = _space(p=p)
        Q = ScaledQ1(S, kel)
        E = O.Ellip(S, p=2.0, kvert=kel.reshape(-1))
        ref = E.q1_matrix(np.zeros(S.nclosure)).toarray()
        got = Q.matrix(np.ones((len(Q.W), 3))).toarray()
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


def test_corrected_scales_integrate_linear_energy_exactly_at_degree_one():
    # p = 1: one sub-cell per element, the element-natural points are the sub-cell's own; with sigma = 1/h = 1/2 the
    # sub-cell matrix is the trilinear stiffness (a quarter of the reference's, the refout factor 4 at p = 1)
    S, _ = _space(M=(3, 3, 3), p=1, perturb=0.0)
    Q = ScaledQ1(S)
    A = Q.matrix(np.full((len(Q.W), 3), 0.5)).toarray()
    R = Q.matrix(np.ones((len(Q.W), 3))).toarray()
    np.testing.assert_allclose(A, R / 4, rtol=1e-13, atol=1e-15)


def test_scale_gradient_matches_finite_differences():
    S, kel = _space(p=3)
    Q = ScaledQ1(S, kel)
    rng = np.random.default_rng(2)
    sig = np.exp(0.3 * rng.standard_normal((len(Q.W), 3)))
    w, z = rng.standard_normal(S.nowned), rng.standard_normal(S.nowned)
    g = Q.scale_gradient(sig, w, z)
    for _ in range(3):
        v = rng.standard_normal(sig.shape)
        h = 1e-6
        fd = (w @ (Q.matrix(sig + h * v) @ z) - w @ (Q.matrix(sig - h * v) @ z)) / (2 * h)
        assert abs(fd - (g * v).sum()) <= 1e-6 * abs(fd) + 1e-9


def test_cell_network_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    w = rng.standard_normal(L.nparam(1)) * 0.2
    X = rng.standard_normal((40, 11))
    T = rng.standard_normal((40, 3))
    loss, g = L.loss_and_grad(w, X, T, kind=1)
    for _ in range(4):
        v = rng.standard_normal(w.size)
        h = 1e-6
        fd = (L.loss_and_grad(w + h * v, X, T, kind=1)[0] - L.loss_and_grad(w - h * v, X, T, kind=1)[0]) / (2 * h)
        assert abs(fd - g @ v) <= 1e-6 * abs(fd) + 1e-10
    # zero network: unit scales (the plain matrix)
    np.testing.assert_array_equal(L.cell_scales(np.zeros(L.nparam(1), np.float32), X), np.ones((40, 3)))

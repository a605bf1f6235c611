"""The numpy restatement of the learned node weights, pinned by finite differences and closed forms."""
import numpy as np

from oracle import mlp_oracle as L


def test_bf16_rounding_closed_form():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -9, -2.5, 0.0, np.inf], dtype=np.float32)
    # bf16 keeps 8 mantissa bits (7 stored): 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-8 -> 1 + 2^-6
    np.testing.assert_array_equal(L.bf16(x), np.array([1.0, 1.0, 1 + 2 ** -6, 1.0, -2.5, 0.0, np.inf], np.float32))
    y = np.random.default_rng(0).standard_normal(1000).astype(np.float32)
    b = L.bf16(y)
    assert np.all(np.abs(b - y) <= np.abs(y) * 2 ** -8)
    np.testing.assert_array_equal(L.bf16(b), b)  # idempotent


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    w = rng.standard_normal(L.NPARAM) * 0.2
    X = rng.standard_normal((7, L.NIN))
    T = rng.standard_normal((7, L.NOUT))
    loss, g = L.loss_and_grad(w, X, T)
    h = 1e-6
    for q in list(range(0, L.NPARAM, 97)) + [L.OFF_B1, L.OFF_B2 + 3, L.OFF_W3 + 5, L.NPARAM - 1]:
        wp, wm = w.copy(), w.copy()
        wp[q] += h
        wm[q] -= h
        fd = (L.loss_and_grad(wp, X, T)[0] - L.loss_and_grad(wm, X, T)[0]) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * (1 + abs(fd)), q


def test_features_and_weights_closed_form():
    box = (4, 2, 3)  # patches 2 x 1 x 1; the odd z layer is not covered
    n = 24
    dinv = np.random.default_rng(2).uniform(0.5, 2, n)
    X = L.features(box, 3, (1, 1, 0), dinv)
    I, J, K, idx = L.patches(box)
    assert idx.shape == (2, 8)
    # patch 0 covers box nodes (0..1, 0..1, 0..1): node (i*2 + j)*3 + k
    np.testing.assert_array_equal(idx[0], [0, 1, 3, 4, 6, 7, 9, 10])
    np.testing.assert_allclose(X[:, :8].sum(1), 0, atol=1e-12)
    np.testing.assert_allclose(X[0, :8], -np.log(dinv[idx[0]]) + np.log(dinv[idx[0]]).mean())
    # patch 1 starts at box i = 2, lattice i = 3, position 3 mod 3 = 0; node m=4 has i = 3 -> 4 mod 3 / 3
    np.testing.assert_allclose(X[1, 8:11], [0.0, 1 / 3, 0.0])
    np.testing.assert_allclose(X[1, 8 + 12:8 + 15], [1 / 3, 1 / 3, 0.0])
    # zero network: all weights 1;  b3 = c: weights exp(clamp(c)/2) on the patch nodes
    w = np.zeros(L.NPARAM)
    np.testing.assert_array_equal(L.node_weights(box, w, X), np.ones(n))
    c = np.array([-6.0, -1.0, 0.0, 0.5, 1.0, 2.0, 4.0, 8.0])
    w[L.OFF_B3:] = c
    s = L.node_weights(box, w, X)
    np.testing.assert_allclose(s[idx[1]], np.exp(0.5 * np.clip(c, -4, 4)))
    untouched = np.setdiff1d(np.arange(n), idx.ravel())
    np.testing.assert_array_equal(s[untouched], 1.0)

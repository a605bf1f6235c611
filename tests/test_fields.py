"""Seeded random coefficient fields: determinism and the Gaussian statistics they are built to have."""
import numpy as np

from dohp_amd.fields import gaussian, gaussian_field, lognormal_field


def test_gaussian_moments():
    z = gaussian(11, 200000)
    assert abs(z.mean()) < 0.01
    assert abs(z.var() - 1) < 0.01


def test_field_deterministic_and_unit_variance():
    rng = np.random.default_rng(0)
    X = rng.random((4000, 3))
    g1 = gaussian_field(X, seed=3)
    assert np.array_equal(g1, gaussian_field(X, seed=3))
    assert not np.allclose(g1, gaussian_field(X, seed=4))
    # pointwise variance over independent seeds -> 1 (random Fourier features are exact in law for M -> inf)
    G = np.stack([gaussian_field(X[:50], seed=s, nmodes=256) for s in range(400)])
    assert abs(G.var(axis=0).mean() - 1) < 0.1
    assert abs(G.mean()) < 0.05


def test_field_correlation_length():
    # covariance between points at distance d should be ~exp(-d^2/(2 l^2))
    l = 0.1
    x0 = np.zeros((1, 3))
    for d in (0.05, 0.1, 0.2):
        x1 = np.array([[d, 0, 0]])
        X = np.vstack([x0, x1])
        G = np.stack([gaussian_field(X, seed=s, corr_len=l, nmodes=256) for s in range(2000)])
        c = np.mean(G[:, 0] * G[:, 1])
        assert abs(c - np.exp(-d * d / (2 * l * l))) < 0.08, (d, c)


def test_lognormal_positive():
    X = np.random.default_rng(1).random((1000, 3))
    k = lognormal_field(X, seed=9)
    assert np.all(k > 0)

"""Seeded random coefficient fields for the solver workloads.

log-normal k = exp(g): g is a zero-mean, unit-variance Gaussian field with squared-exponential
covariance exp(-|x-y|^2 / (2 l^2)), sampled by random Fourier features
g(x) = sqrt(2/M) sum_m cos(w_m . x + phi_m), w_m ~ N(0, I/l^2), phi_m ~ U[0, 2pi).
The reference has no random fields; this is the build's own generator (seeded, splitmix64).
"""
from __future__ import annotations

import numpy as np

from .mesh import splitmix64_uniform


def gaussian(seed: int, count: int) -> np.ndarray:
    """Standard normal samples by Box-Muller on the splitmix64 stream of `seed`."""
    u = splitmix64_uniform(seed, 2 * count).reshape(count, 2)
    u1 = np.maximum(u[:, 0], 1e-300)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2 * np.pi * u[:, 1])


def gaussian_field(X: np.ndarray, seed: int, corr_len: float = 0.1, nmodes: int = 512) -> np.ndarray:
    X = np.asarray(X, dtype=np.float64)
    w = gaussian(seed, 3 * nmodes).reshape(nmodes, 3) / corr_len
    phi = 2 * np.pi * splitmix64_uniform(seed + 1, nmodes)
    g = np.zeros(len(X))
    for m0 in range(0, nmodes, 64):
        g += np.cos(X @ w[m0:m0 + 64].T + phi[m0:m0 + 64]).sum(axis=1)
    return np.sqrt(2.0 / nmodes) * g


def lognormal_field(X: np.ndarray, seed: int, corr_len: float = 0.1, nmodes: int = 512) -> np.ndarray:
    return np.exp(gaussian_field(X, seed, corr_len, nmodes))

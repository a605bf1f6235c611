"""Pins of the oracle's element-local mixed-order Poisson apply (orc_expanded_poisson_apply) by
what the mathematics fixes: constants in the kernel of every element stiffness, symmetry, the
exact energy of linear functions on affine elements (Gauss rules integrate the constant
integrand exactly), the integral of a trilinear coefficient, and agreement of each element with
the one-element ellip Jacobian at p = 2 on a block space of its own."""
import numpy as np
import pytest

import oracle.oracle as O
from dohp_amd.mesh import block_vertices, element_vertices, splitmix64_uniform, uniform_integers


def lattice(n):
    z, _ = O.gauss_lobatto(n)
    g = (1 + z) / 2
    A, B, C = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([A.reshape(-1), B.reshape(-1), C.reshape(-1)], axis=1)  # node (i,j,k), k fastest


def affine_element(seed):
    r = splitmix64_uniform(seed, 12)
    A = np.eye(3) + 0.4 * (r[:9].reshape(3, 3) - 0.5)
    b = r[9:] - 0.5
    corners = np.array([(a, bb, c) for a in (0, 1) for bb in (0, 1) for c in (0, 1)], dtype=float)
    return A, b, corners @ A.T + b


def mixed_elements(nelem, seed, lo=1, hi=10):
    V = block_vertices((nelem, 1, 1), box=(0, nelem, 0, 1, 0, 1), perturb=0.15, seed=seed)
    return uniform_integers(seed + 1, nelem, lo, hi), element_vertices((nelem, 1, 1), V)


def test_constants_in_kernel_and_symmetry():
    deg, vx = mixed_elements(10, 3)
    n3 = (deg + 1) ** 3
    y = O.expanded_poisson_apply(deg, vx, np.ones(n3.sum()))
    x1 = 2 * splitmix64_uniform(4, n3.sum()) - 1
    x2 = 2 * splitmix64_uniform(5, n3.sum()) - 1
    K1 = O.expanded_poisson_apply(deg, vx, x1)
    assert np.abs(y).max() < 1e-12 * np.abs(K1).max()
    K2 = O.expanded_poisson_apply(deg, vx, x2)
    off = np.concatenate([[0], np.cumsum(n3)])
    for e in range(len(deg)):
        s = slice(off[e], off[e + 1])
        assert abs(x1[s] @ K2[s] - x2[s] @ K1[s]) < 1e-12 * abs(x1[s] @ K1[s])
        assert x1[s] @ K1[s] > 0


@pytest.mark.parametrize("degree", [1, 2, 4, 7, 10])
def test_linear_energy_on_affine_elements(degree):
    A, b, vx = affine_element(40 + degree)
    g = np.array([0.3, -1.1, 0.7])
    X = lattice(degree + 1) @ A.T + b
    u = X @ g
    c = 2.5
    Ku = O.expanded_poisson_apply([degree], vx[None], u, kvert=np.full((1, 8), c))
    assert abs(u @ Ku - c * (g @ g) * np.linalg.det(A)) < 1e-12 * c * (g @ g)


def test_trilinear_coefficient_integral():
    degree = 3
    A, b, vx = affine_element(7)
    kv = 1.0 + splitmix64_uniform(8, 8)
    g = np.array([1.0, 0.5, -0.25])
    u = (lattice(degree + 1) @ A.T + b) @ g
    Ku = O.expanded_poisson_apply([degree], vx[None], u, kvert=kv[None])
    # the integral of a trilinear function over an affine image of the unit cube is its vertex
    # mean times the volume
    assert abs(u @ Ku - kv.mean() * (g @ g) * np.linalg.det(A)) < 1e-12 * (g @ g)


def test_matches_one_element_spaces():
    deg, vx = mixed_elements(4, 9, lo=2, hi=6)
    n3 = (deg + 1) ** 3
    x = 2 * splitmix64_uniform(10, n3.sum()) - 1
    kv = 1 + splitmix64_uniform(11, 32).reshape(4, 8)
    y = O.expanded_poisson_apply(deg, vx, x, kvert=kv)
    off = np.concatenate([[0], np.cumsum(n3)])
    for e in range(4):
        S = O.BlockSpace((1, 1, 1), int(deg[e]), vx[e].reshape(8, 3), dirichlet_mask=0)
        E = O.Ellip(S, p=2.0, kvert=kv[e][None])
        xe = np.zeros(S.nclosure)
        xe[S.elem_nodes[0]] = x[off[e]:off[e + 1]]
        ye = E.jacobian_apply(np.zeros(S.nclosure), xe)
        assert np.abs(ye[S.elem_nodes[0]] - y[off[e]:off[e + 1]]).max() < 1e-13 * np.abs(y[off[e]:off[e + 1]]).max()

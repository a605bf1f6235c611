"""Synthetic hexahedral block meshes (vertex coordinates only).

Follows the block generator of dohp (src/fs/mesh/interface/genblock.c:145-250):
vertices on an (m, n, p) lattice over a box, vertex (i, j, k) stored at
index (i*n + j)*p + k, hexahedra in (i, j, k) order with k fastest.  Optional
seeded perturbation of interior vertices and the ellip "morph" map
(src/fs/tests/ellip.c:173-181) produce non-affine elements.

The element->node maps and the rest of the function space are built natively
by the library (dFSCreateBlock); this module only produces coordinates.
"""
from __future__ import annotations

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64_uniform(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Uniform doubles in [0, 1) from the splitmix64 stream of `seed` (entries offset..offset+count)."""
    with np.errstate(over="ignore"):
        idx = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def block_vertices(M, box=(-1.0, 1.0, -1.0, 1.0, -1.0, 1.0), perturb: float = 0.0, seed: int = 0,
                   morph: float = 0.0, twist: float = 0.0, stretch: float = 0.0) -> np.ndarray:
    """Vertex coordinates [(M0+1)(M1+1)(M2+1), 3] of an M0 x M1 x M2 hexahedral block.

    perturb: interior vertices move by U(-perturb*h, +perturb*h) per coordinate (seeded).
    morph/twist/stretch: the deformation of ellip.c:173-181 applied after perturbation.
    """
    M = tuple(int(v) for v in M)
    nv = tuple(m + 1 for m in M)
    axes = [np.linspace(box[2 * d], box[2 * d + 1], nv[d]) for d in range(3)]
    X = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, 3).copy()
    if perturb:
        h = np.array([(box[2 * d + 1] - box[2 * d]) / M[d] for d in range(3)])
        I, J, K = np.meshgrid(*[np.arange(n) for n in nv], indexing="ij")
        interior = ((I > 0) & (I < M[0]) & (J > 0) & (J < M[1]) & (K > 0) & (K < M[2])).reshape(-1)
        r = splitmix64_uniform(seed, X.size).reshape(-1, 3)
        X[interior] += (2.0 * r[interior] - 1.0) * perturb * h
    if morph or twist or stretch:
        x, y, z = X[:, 0].copy(), X[:, 1].copy(), X[:, 2].copy()
        X[:, 0] = np.cos(twist * z) * x + np.sin(twist * z) * y
        X[:, 1] = -np.sin(twist * z) * x + np.cos(twist * z) * y
        X[:, 2] = -1 + (z + 1) * (1 + morph * ((x + 1) ** 2 + (y + 1) ** 2)) + stretch * z
    return X


# Boundary sets of the block generator (genblock.c:406-420): 100 = the four walls
# (x and y faces), 200 = top (z max), 300 = bottom (z min).
DIRICHLET_WALLS, DIRICHLET_TOP, DIRICHLET_BOTTOM = 1, 2, 4
DIRICHLET_ALL = 7


def dirichlet_mask_from_sets(sets) -> int:
    mask = 0
    for s in sets:
        mask |= {100: DIRICHLET_WALLS, 200: DIRICHLET_TOP, 300: DIRICHLET_BOTTOM}[int(s)]
    return mask


def element_vertices(M, V) -> np.ndarray:
    """[M0*M1*M2][8][3]: vertex coordinates of each hexahedron of the block (elements in (i, j, k)
    order, k fastest), its vertices in tensor order (a, b, c) with c fastest, the order of
    src/fs/interface/fsgeom.c:151."""
    M = tuple(int(m) for m in M)
    V = np.asarray(V, dtype=np.float64).reshape(M[0] + 1, M[1] + 1, M[2] + 1, 3)
    e0, e1, e2 = (g.reshape(-1) for g in np.meshgrid(*[np.arange(m) for m in M], indexing="ij"))
    t = np.array([(a, b, c) for a in (0, 1) for b in (0, 1) for c in (0, 1)])
    return np.ascontiguousarray(V[e0[:, None] + t[None, :, 0], e1[:, None] + t[None, :, 1], e2[:, None] + t[None, :, 2]])


def element_vertex_index(M) -> np.ndarray:
    """[M0*M1*M2][8] block vertex index of each hexahedron's vertices in tensor order."""
    M = tuple(int(m) for m in M)
    e0, e1, e2 = (g.reshape(-1) for g in np.meshgrid(*[np.arange(m) for m in M], indexing="ij"))
    t = np.array([(a, b, c) for a in (0, 1) for b in (0, 1) for c in (0, 1)])
    return ((e0[:, None] + t[None, :, 0]) * (M[1] + 1) + e1[:, None] + t[None, :, 1]) * (M[2] + 1) + e2[:, None] + t[None, :, 2]


def uniform_integers(seed: int, count: int, lo: int, hi: int) -> np.ndarray:
    """Seeded integers uniform on lo..hi (inclusive) from the splitmix64 stream."""
    return lo + np.minimum((splitmix64_uniform(seed, count) * (hi - lo + 1)).astype(np.int64), hi - lo)


def sine_modes_state(X, seed: int) -> np.ndarray:
    """u0 = sum over 8 modes of 0.1 sin(2 pi k_m . x + phi_m) at the points X [n][3], with k_m in {1..4}^3 and
    phi_m uniform in [0, 2 pi), both drawn from the splitmix64 stream of `seed` (the linearization state of the
    order-6 nonlinear diffusion workload)."""
    r = splitmix64_uniform(seed, 32)
    u = np.zeros(len(X))
    for m in range(8):
        k = 1 + np.floor(r[4 * m:4 * m + 3] * 4)
        u += 0.1 * np.sin(2 * np.pi * (X @ k) + 2 * np.pi * r[4 * m + 3])
    return u


def mixed_order_batch(Me: int = 32, nprob: int = 256):
    """The batched hp workload: an Me^3 perturbed mesh on [0,1]^3 whose elements take degrees uniform on 2..10
    (seeded), plus nprob independent Poisson problems on n^3 perturbed meshes, n uniform on 4..16, degree uniform
    on 2..8.  Returns (degrees [Me^3], element vertices [Me^3][8][3], [(M, degree, V)] of the problems)."""
    M = (Me, Me, Me)
    V = block_vertices(M, box=(0, 1, 0, 1, 0, 1), perturb=0.1, seed=3)
    deg = uniform_integers(4, Me ** 3, 2, 10)
    ns, ps = uniform_integers(6, nprob, 4, 16), uniform_integers(7, nprob, 2, 8)
    probs = []
    for i in range(nprob):
        Mi = (int(ns[i]),) * 3
        probs.append((Mi, int(ps[i]), block_vertices(Mi, box=(0, 1, 0, 1, 0, 1), perturb=0.1, seed=1000 + i)))
    return deg, element_vertices(M, V), probs

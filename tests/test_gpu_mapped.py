"""Spaces built from explicit element maps (dFSCreateFromElements) on the GPU.

The fused kernels read such a space through its element-to-node and element-to-vertex maps instead of
evaluating the block numbering.  Parity: the maps of a block space give the block space's results (and the
oracle's, 1e-12); any renumbering of elements, owned nodes, Dirichlet nodes and vertices gives the same
operator, renumbered (the kernels assume no structure of the maps); the gather/scatter and the deterministic
colour scatter follow the greedy colouring; refusals are argument errors."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from dohp_amd.mesh import block_vertices, element_vertex_index, splitmix64_uniform  # noqa: E402
from oracle import oracle as O  # noqa: E402

TOL64 = 1e-12


@pytest.fixture(scope="module")
def C():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import dohp_amd.capi as capi
    return capi


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def block_vertex_map(M):
    """elem_verts of a block mesh in tensor order."""
    return np.ascontiguousarray(element_vertex_index(M), dtype=np.int32)


def applies(C, fs, op, lin, u, v, shapes):
    """F(u), J(u) v from stored data and recomputed, A v, and their FP32 forms, per pinned shape."""
    out = {}
    for shape in shapes:
        res = {}
        for storage in (0, 1):
            op.set_storage(storage)
            lin.set_storage(storage)
            jmode, lmode = ("stored", "stored") if storage else ("jacobian", "linear")
            for o, mode in ((op, "residual"), (op, jmode), (lin, lmode)):
                o.set_shape(mode, 0, shape)
                o.set_shape(mode, 1, shape)
            f = torch.zeros(fs.nowned, dtype=torch.float64, device="cuda")
            op.function(dev(u), f)
            y = torch.zeros_like(f)
            op.mat_mult(dev(v), y)
            yl = torch.zeros_like(f)
            lin.mat_mult(dev(v), yl)
            y32 = torch.zeros(fs.nowned, dtype=torch.float32, device="cuda")
            op.mat_mult_f32(dev(v, torch.float32), y32)
            res[storage] = [t.cpu().numpy().astype(np.float64) for t in (f, y, yl, y32)]
        out[shape] = res
    return out


@pytest.mark.parametrize("p", [4, 6])
def test_block_maps_give_the_block_space(C, p):
    """A space built from a block space's own maps: every launch shape it allows (four-wavefront,
    one-wavefront, matrix-core) against the block space and the oracle."""
    print(f"mapped = block p={p}", flush=True)
    M = (3, 2, 3)
    V = block_vertices(M, box=(0, 1, 0, 1, 0, 1), perturb=0.1, seed=17)
    fs = C.FS(M, p, V)
    fm = C.FS.from_elements(p, fs.element_nodes(), fs.nowned, fs.nclosure, V, block_vertex_map(M))
    assert (fm.nowned, fm.nclosure, fm.nelem) == (fs.nowned, fs.nclosure, fs.nelem)
    np.testing.assert_allclose(fm.nodal_coordinates(), fs.nodal_coordinates(), rtol=0, atol=1e-15)
    S = O.BlockSpace(M, p, V)
    E = O.Ellip(S, exact=-1, p=1.5, eps=np.sqrt(0.005))  # (no forcing term: C.Ellip's default)
    El = O.Ellip(S, exact=-1, p=2.0)
    uc = np.sin(3 * S.nodal_coordinates()).sum(axis=1)
    v = 2 * splitmix64_uniform(19, fs.nowned) - 1
    vc = np.concatenate([v, np.zeros(fs.nclosure - fs.nowned)])
    refs = [E.function(uc)[:fs.nowned], E.jacobian_apply(uc, vc)[:fs.nowned],
            El.jacobian_apply(np.zeros(fs.nclosure), vc)[:fs.nowned]]
    got = {}
    for name, space in (("block", fs), ("maps", fm)):
        space.dirichlet_commit(uc[fs.nowned:])
        op = C.Ellip(space, p=1.5, eps=np.sqrt(0.005))
        lin = C.Ellip(space, p=2.0)
        got[name] = applies(C, space, op, lin, uc[:fs.nowned], v, (0, 1, 4))
    for shape in (0, 1, 4):
        for storage in (0, 1):
            b, m = got["block"][shape][storage], got["maps"][shape][storage]
            errs = [rel(m[i], refs[i]) for i in range(3)]
            print(f"shape {shape} storage {storage}: oracle {['%.1e' % e for e in errs]} "
                  f"maps vs block {['%.1e' % rel(m[i], b[i]) for i in range(4)]}")
            assert max(errs) < TOL64, (shape, storage, errs)
            assert rel(m[3], refs[1]) < 1e-4
            for i in range(4):  # the same arithmetic in the same order, up to the atomic scatter's order
                assert rel(m[i], b[i]) < (1e-6 if i == 3 else 1e-14), (shape, storage, i)


@pytest.mark.parametrize("p", [1, 2, 3, 5, 7, 8, 10])
def test_block_maps_every_degree(C, p):
    """Every degree the fused kernels compile: the residual, J v (stored and recomputed) and the linear apply on
    a block space's own maps equal the block space's (the tuner's shapes)."""
    print(f"maps every degree p={p}", flush=True)
    M = (2, 3, 2)
    V = block_vertices(M, perturb=0.1, seed=40 + p)
    fs = C.FS(M, p, V)
    fm = C.FS.from_elements(p, fs.element_nodes(), fs.nowned, fs.nclosure, V, block_vertex_map(M))
    X = fs.nodal_coordinates()
    uc = np.sin(2 * X).sum(axis=1)
    v = 2 * splitmix64_uniform(50 + p, fs.nowned) - 1
    res = {}
    for name, space in (("block", fs), ("maps", fm)):
        space.dirichlet_commit(uc[fs.nowned:])
        op = C.Ellip(space, p=1.5, eps=0.1)
        lin = C.Ellip(space, p=2.0)
        out = []
        for storage in (0, 1):
            op.set_storage(storage)
            f = torch.zeros(fs.nowned, dtype=torch.float64, device="cuda")
            op.function(dev(uc[:fs.nowned]), f)
            y = torch.zeros_like(f)
            op.mat_mult(dev(v), y)
            yl = torch.zeros_like(f)
            lin.mat_mult(dev(v), yl)
            out += [f.cpu().numpy(), y.cpu().numpy(), yl.cpu().numpy()]
        res[name] = out
    errs = [rel(m, b) for m, b in zip(res["maps"], res["block"])]
    print("maps vs block", ["%.1e" % e for e in errs])
    assert max(errs) < 1e-14, errs


def test_any_numbering_gives_the_same_operator(C):
    """Random renumbering of elements, owned nodes, Dirichlet nodes and vertices, with a per-vertex
    coefficient, the Bratu term and the exact solution's forcing term: the operator is the block space's,
    renumbered."""
    M, p = (4, 3, 2), 5
    V = block_vertices(M, perturb=0.15, seed=8)
    fs = C.FS(M, p, V)
    rng = np.random.default_rng(4)
    E = fs.element_nodes()
    EV = block_vertex_map(M)
    pe = rng.permutation(fs.nelem)                       # new element k is old element pe[k]
    po = rng.permutation(fs.nowned)                      # new owned node k is old owned node po[k]
    pd = rng.permutation(fs.nclosure - fs.nowned)
    q = np.concatenate([np.argsort(po), fs.nowned + np.argsort(pd)])  # old closure index -> new
    nv = V.shape[0]
    pv = rng.permutation(nv)
    qv = np.argsort(pv)                                  # old vertex -> new
    fm = C.FS.from_elements(p, q[E[pe]], fs.nowned, fs.nclosure, V[pv], qv[EV[pe]])
    X = fs.nodal_coordinates()
    np.testing.assert_allclose(fm.nodal_coordinates()[q], X, rtol=0, atol=1e-14)
    kv = np.exp(np.sin(4 * V[:, 0]) * np.cos(3 * V[:, 2]))
    uc = np.cos(2 * X).sum(axis=1)
    v = 2 * splitmix64_uniform(31, fs.nowned) - 1
    inv = np.argsort(q)                                  # new closure index -> old
    fs.dirichlet_commit(uc[fs.nowned:])
    fm.dirichlet_commit(uc[inv][fs.nowned:])
    out = {}
    for name, space, kvert, perm in (("block", fs, kv, None), ("maps", fm, kv[pv], inv[:fs.nowned])):
        op = C.Ellip(space, p=1.7, eps=0.1, lam=0.3, exact=0, kvert=kvert)  # (with the forcing term)
        lin = C.Ellip(space, p=2.0, kvert=kvert)
        uu = uc[:fs.nowned] if perm is None else uc[perm]
        vv = v if perm is None else v[perm]
        r = applies(C, space, op, lin, uu, vv, (0,))[0]
        if perm is not None:  # back to the block numbering
            back = np.argsort(perm)
            r = {s: [a[back] for a in r[s]] for s in r}
        out[name] = r
    for storage in (0, 1):
        errs = [rel(out["maps"][storage][i], out["block"][storage][i]) for i in range(4)]
        print(f"storage {storage}: renumbered vs block {['%.1e' % e for e in errs]}")
        assert max(errs[:3]) < 1e-13 and errs[3] < 1e-6, (storage, errs)


def test_element_orientations_do_not_change_the_operator(C):
    """Elements seen through a random rotation of the cube (axes permuted and reflected with determinant +1, the node
    and the vertex tensors alike: a reflection would turn the element inside out, which is an error), as neighbouring
    elements of an unstructured mesh are: the Gauss rule is symmetric, so the operator is the block space's to
    rounding."""
    M, p = (3, 3, 3), 4
    V = block_vertices(M, perturb=0.12, seed=21)
    fs = C.FS(M, p, V)
    n = p + 1
    rng = np.random.default_rng(17)
    E = fs.element_nodes().reshape(-1, n, n, n)
    EV = block_vertex_map(M).reshape(-1, 2, 2, 2)
    En, EVn = E.copy(), EV.copy()
    turned = 0
    for e in range(fs.nelem):
        if rng.random() < 0.3:
            continue
        axes, flips = rng.permutation(3), rng.random(3) < 0.5
        odd_perm = sum(axes[i] > axes[j] for i in range(3) for j in range(i + 1, 3)) % 2
        if (odd_perm + flips.sum()) % 2:  # keep the orientation: one more reflection
            flips[2] = not flips[2]
        def orient(T):
            T = np.transpose(T, axes)
            for d in range(3):
                if flips[d]:
                    T = np.flip(T, axis=d)
            return T
        En[e], EVn[e] = orient(E[e]), orient(EV[e])
        turned += 1
    fm = C.FS.from_elements(p, En.reshape(fs.nelem, -1), fs.nowned, fs.nclosure, V, EVn.reshape(fs.nelem, 8))
    X = fs.nodal_coordinates()
    np.testing.assert_allclose(fm.nodal_coordinates(), X, rtol=0, atol=1e-14)
    uc = np.sin(2 * X).sum(axis=1)
    v = 2 * splitmix64_uniform(61, fs.nowned) - 1
    res = {}
    for name, space in (("block", fs), ("turned", fm)):
        space.dirichlet_commit(uc[fs.nowned:])
        op = C.Ellip(space, p=1.5, eps=0.1, exact=0)
        lin = C.Ellip(space, p=2.0)
        res[name] = applies(C, space, op, lin, uc[:fs.nowned], v, (0,))[0]
    for storage in (0, 1):
        errs = [rel(res["turned"][storage][i], res["block"][storage][i]) for i in range(4)]
        print(f"{turned} of {fs.nelem} elements turned, storage {storage}: {['%.1e' % e for e in errs]}")
        assert max(errs[:3]) < 1e-13 and errs[3] < 1e-6, (storage, errs)


def test_gather_scatter_and_colours_on_maps(C):
    """dFSGlobalToExpanded / dFSExpandedToGlobal through renumbered maps (greedy colours), and the
    operator's deterministic colour scatter against its atomic one."""
    M, p = (3, 3, 2), 3
    V = block_vertices(M, perturb=0.1, seed=5)
    fs = C.FS(M, p, V)
    rng = np.random.default_rng(9)
    pe = rng.permutation(fs.nelem)
    q = np.concatenate([rng.permutation(fs.nowned), fs.nowned + rng.permutation(fs.nclosure - fs.nowned)])
    En = q[fs.element_nodes()[pe]]
    fm = C.FS.from_elements(p, En, fs.nowned, fs.nclosure, V, block_vertex_map(M)[pe])
    g = rng.uniform(-1, 1, fs.nowned)
    dv = rng.uniform(-1, 1, fs.nclosure - fs.nowned)
    fm.dirichlet_commit(dv)
    x = torch.zeros(fm.nexpanded, dtype=torch.float64, device="cuda")
    fm.global_to_expanded(dev(g), x, C.INHOMOGENEOUS)
    np.testing.assert_array_equal(x.cpu().numpy(), np.concatenate([g, dv])[En.reshape(-1)])
    xe = rng.uniform(-1, 1, fm.nexpanded)
    gd = dev(np.ones(fm.nowned))
    fm.expanded_to_global(dev(xe), gd, C.INHOMOGENEOUS, C.ADD_VALUES)
    ref = np.ones(fm.nclosure)
    np.add.at(ref, En.reshape(-1), xe)
    np.testing.assert_allclose(gd.cpu().numpy(), ref[:fm.nowned], rtol=1e-15, atol=1e-15)
    op = C.Ellip(fm, p=2.0)
    xv = dev(2 * splitmix64_uniform(4, fm.nowned) - 1)
    ys = []
    for scatter in (0, 1, 0):
        op.set_scatter(scatter)
        y = torch.zeros(fm.nowned, dtype=torch.float64, device="cuda")
        op.mat_mult(xv, y)
        ys.append(y)
    assert rel(ys[1].cpu().numpy(), ys[0].cpu().numpy()) < 1e-14
    assert torch.equal(ys[0], ys[2])


def test_maps_are_checked_and_block_only_features_refused(C):
    M, p = (2, 2, 2), 2
    V = block_vertices(M, perturb=0.0, seed=1)
    fs = C.FS(M, p, V)
    E, EV = fs.element_nodes(), block_vertex_map(M)
    bad = E.copy()
    bad[1, 3] = fs.nclosure
    with pytest.raises(C.DohpError, match="out of range"):
        C.FS.from_elements(p, bad, fs.nowned, fs.nclosure, V, EV)
    with pytest.raises(C.DohpError, match="belongs to no element"):
        C.FS.from_elements(p, E, fs.nowned, fs.nclosure + 1, V, EV)
    badv = EV.copy()
    badv[0, 0] = -1
    with pytest.raises(C.DohpError, match="vertex index out of range"):
        C.FS.from_elements(p, E, fs.nowned, fs.nclosure, V, badv)
    fm = C.FS.from_elements(p, E, fs.nowned, fs.nclosure, V, EV)
    op = C.Ellip(fm, p=2.0)
    with pytest.raises(C.DohpError, match="block space"):
        op.set_shape("linear", 0, 2)
    with pytest.raises(C.DohpError, match="block space"):
        op.assemble_q1()

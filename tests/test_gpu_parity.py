"""Parity of the HIP path (through the C ABI) with the CPU oracle, on the same seeded inputs.

FP64 bar: relative 2-norm difference <= 1e-12.  FP32 results state their error against
the FP64 oracle.
"""
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from dohp_amd.mesh import block_vertices, splitmix64_uniform  # noqa: E402
from oracle import oracle as O  # noqa: E402

G = os.path.join(os.path.dirname(__file__), "golden")
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


def setup(C, M, deg, V, mask=7, **prm):
    fs = C.FS(M, deg, V, dirichlet_mask=mask)
    S = O.BlockSpace(M, deg, V, dirichlet_mask=mask)
    gexact = prm.pop("gexact", -1)
    E = O.Ellip(S, exact=gexact, **prm)
    op = C.Ellip(fs, exact=gexact, **prm)
    return fs, S, E, op


# ---------------------------------------------------------------- dEFSApply
@pytest.mark.parametrize("P,Q,D", [(2, 2, 3), (3, 4, 1), (5, 5, 1), (7, 7, 1), (4, 6, 2), (11, 11, 1)])
def test_efs_apply_batched_all_modes(C, P, Q, D):
    print(f"efs P={P} Q={Q} D={D}", flush=True)
    rng = np.random.default_rng(P * 31 + Q * 7 + D)
    ne = 5
    rc, _ = O.gauss_legendre(Q)
    jinv = rng.uniform(-1, 1, (ne, Q ** 3, 9))
    sizes = {C.APPLY_INTERP: (P ** 3 * D, Q ** 3 * D), C.APPLY_GRAD: (P ** 3 * D, Q ** 3 * D * 3),
             C.APPLY_INTERP_TRANSPOSE: (Q ** 3 * D, P ** 3 * D), C.APPLY_GRAD_TRANSPOSE: (Q ** 3 * D * 3, P ** 3 * D)}
    for amode, (nin, nout) in sizes.items():
        for imode in (C.INSERT_VALUES, C.ADD_VALUES):
            for use_j in ((True, False) if amode == C.APPLY_GRAD else (True,)):
                x = rng.uniform(-1, 1, (ne, nin))
                y0 = rng.uniform(-1, 1, (ne, nout))
                yd = dev(y0)
                jd = dev(jinv) if use_j else None
                C.efs_apply_hex_batched(P, rc, ne, jd, D, dev(x), yd, amode, imode)
                got = yd.cpu().numpy()
                for e in range(ne):
                    ref = O.efs_apply_hex(P, rc, jinv[e] if use_j else None, D, x[e], amode, imode, out=y0[e])
                    assert rel(got[e], ref) < 1e-13, (amode, imode, e)


# ------------------------------------------------------------ gather/scatter
def test_gather_scatter(C):
    M, p = (3, 4, 2), 3
    V = block_vertices(M, perturb=0.1, seed=5)
    fs = C.FS(M, p, V)
    S = O.BlockSpace(M, p, V)
    rng = np.random.default_rng(0)
    g = rng.uniform(-1, 1, fs.nowned)
    dv = rng.uniform(-1, 1, fs.nclosure - fs.nowned)
    fs.dirichlet_commit(dv)
    x = torch.zeros(fs.nexpanded, dtype=torch.float64, device="cuda")
    fs.global_to_expanded(dev(g), x, C.INHOMOGENEOUS)
    np.testing.assert_array_equal(x.cpu().numpy(), np.concatenate([g, dv])[S.elem_nodes.reshape(-1)])
    fs.global_to_expanded(dev(g), x, C.HOMOGENEOUS)
    np.testing.assert_array_equal(x.cpu().numpy(), np.concatenate([g, 0 * dv])[S.elem_nodes.reshape(-1)])
    xe = rng.uniform(-1, 1, fs.nexpanded)
    gd = dev(np.ones(fs.nowned))
    fs.expanded_to_global(dev(xe), gd, C.INHOMOGENEOUS, C.ADD_VALUES)
    ref = np.ones(fs.nclosure)
    np.add.at(ref, S.elem_nodes.reshape(-1), xe)
    np.testing.assert_allclose(gd.cpu().numpy(), ref[:fs.nowned], rtol=1e-15, atol=1e-15)


# ----------------------------------------- config 1: 4^3 p=4 Poisson, perturbed
def test_config1_poisson_apply_and_residual(C):
    print("config1 4^3 p=4 perturbed", flush=True)
    M, p = (4, 4, 4), 4
    V = block_vertices(M, box=(0, 1, 0, 1, 0, 1), perturb=0.1, seed=1)
    fs, S, E, op = setup(C, M, p, V, p=2.0, eps=1.0, gexact=4)
    x = 2 * splitmix64_uniform(7, fs.nowned) - 1
    dv = E.dirichlet_values()
    fs.dirichlet_commit(dv)
    y = torch.zeros(fs.nowned, dtype=torch.float64, device="cuda")
    op.mat_mult(dev(x), y)
    uc0 = E.closure(np.zeros(fs.nowned), np.zeros(len(dv)))
    ref = E.jacobian_apply(uc0, np.concatenate([x, np.zeros(len(dv))]))[:fs.nowned]
    assert rel(y.cpu().numpy(), ref) < TOL64
    f = torch.zeros_like(y)
    op.function(dev(x), f)
    reff = E.function(E.closure(x, dv))[:fs.nowned]
    assert rel(f.cpu().numpy(), reff) < TOL64


# --------------------------------------------------------- ellip refout norms
HIST = json.load(open(os.path.join(G, "ellip_histories.json")))
CASES = {"ellip-e0-b1-p16": (1, 1.6, 0.0, (1.6, 1.7, 1.8)), "ellip-e0-b3-p16": (3, 1.6, 0.0, (1.6, 1.7, 1.8)),
         "ellip-e0-b3-p14-sse": (3, 1.4, 0.0, (1.6, 1.7, 1.8)), "ellip-e0-b3-p17-lam1-sse": (3, 1.7, 1.0, (1, 1, 1))}


@pytest.mark.parametrize("name", list(CASES))
def test_ellip_refout_initial_residual_on_gpu(C, name):
    print(name, flush=True)
    bdeg, ep, lam, (a, b, c) = CASES[name]
    V = block_vertices((4, 4, 4))
    fs, S, E, op = setup(C, (4, 4, 4), bdeg, V, p=ep, eps=1.0, lam=lam, gexact=0, a=a, b=b, c=c)
    dv = E.dirichlet_values()
    fs.dirichlet_commit(dv)
    f = torch.zeros(fs.nowned, dtype=torch.float64, device="cuda")
    op.function(dev(np.zeros(fs.nowned)), f)
    F = f.cpu().numpy()
    assert np.linalg.norm(F) == pytest.approx(HIST[name]["snes"][0], rel=6e-6)
    assert rel(F, E.function(E.closure(np.zeros(fs.nowned), dv))[:fs.nowned]) < TOL64


@pytest.mark.parametrize("deg", range(1, 11))
def test_residual_with_forcing_every_degree(C, deg):
    """The forcing term is assembled on the first residual and reused: two residuals at different states
    of a manufactured solution (exact solution 0, nonlinear with the Bratu term), each against the oracle."""
    print(f"forcing p={deg}", flush=True)
    M = (2, 1, 2) if deg >= 8 else (2, 2, 2)
    V = block_vertices(M, box=(0, 1, 0, 1, 0, 1), perturb=0.1, seed=17)
    fs, S, E, op = setup(C, M, deg, V, p=1.6, eps=0.5, lam=0.5, gexact=0, a=1.2, b=0.8, c=1.1)
    dv = E.dirichlet_values()
    fs.dirichlet_commit(dv)
    f = torch.zeros(fs.nowned, dtype=torch.float64, device="cuda")
    for seed in (31, 32):
        x = 0.3 * (2 * splitmix64_uniform(seed, fs.nowned) - 1)
        op.function(dev(x), f)
        assert rel(f.cpu().numpy(), E.function(E.closure(x, dv))[:fs.nowned]) < TOL64, seed


# ------------------------------------------------- nonlinear residual and Jv
def sine_state(X, seed):
    """Seeded sum of 8 random-phase sine modes of amplitude 0.1."""
    r = splitmix64_uniform(seed, 32)
    u = np.zeros(len(X))
    for m in range(8):
        k = 1 + np.floor(r[4 * m:4 * m + 3] * 4)
        u += 0.1 * np.sin(2 * np.pi * (X @ k) + 2 * np.pi * r[4 * m + 3])
    return u


@pytest.mark.parametrize("M,p,lam", [((3, 3, 3), 6, 0.0), ((2, 3, 2), 3, 0.5), ((2, 2, 2), 8, 0.0)])
def test_nonlinear_function_and_jacobian_action(C, M, p, lam):
    print(f"nonlinear M={M} p={p} lam={lam}", flush=True)
    V = block_vertices(M, box=(0, 1, 0, 1, 0, 1), perturb=0.1, seed=11)
    fs, S, E, op = setup(C, M, p, V, p=1.5, eps=np.sqrt(0.005), lam=lam)
    Xc = S.nodal_coordinates()
    uc = sine_state(Xc, 3)
    fs.dirichlet_commit(uc[fs.nowned:])
    f = torch.zeros(fs.nowned, dtype=torch.float64, device="cuda")
    op.function(dev(uc[:fs.nowned]), f)
    assert rel(f.cpu().numpy(), E.function(uc)[:fs.nowned]) < TOL64
    v = 2 * splitmix64_uniform(9, fs.nowned) - 1
    y = torch.zeros_like(f)
    op.mat_mult(dev(v), y)
    ref = E.jacobian_apply(uc, np.concatenate([v, np.zeros(fs.nclosure - fs.nowned)]))[:fs.nowned]
    assert rel(y.cpu().numpy(), ref) < TOL64
    # FP32 inner apply: error against the FP64 oracle is stated, not zero
    y32 = torch.zeros(fs.nowned, dtype=torch.float32, device="cuda")
    op.mat_mult_f32(dev(v, torch.float32), y32)
    e32 = rel(y32.cpu().numpy(), ref)
    print(f"FP32 Jv relative error vs FP64 oracle: {e32:.3e}")
    assert e32 < 1e-4


@pytest.mark.parametrize("deg", range(1, 11))
def test_stored_quadrature_data_every_degree(C, deg):
    """J v from stored quadrature data (7 values per point with the Bratu term) at every degree, FP64
    against the oracle and against the recomputing path, FP32 with its error stated."""
    print(f"stored p={deg}", flush=True)
    M = (2, 1, 2) if deg >= 8 else (2, 2, 2)
    V = block_vertices(M, box=(0, 1, 0, 1, 0, 1), perturb=0.1, seed=23)
    fs, S, E, op = setup(C, M, deg, V, p=1.6, eps=0.3, lam=0.7)
    uc = sine_state(S.nodal_coordinates(), 7)
    fs.dirichlet_commit(uc[fs.nowned:])
    f = torch.zeros(fs.nowned, dtype=torch.float64, device="cuda")
    op.function(dev(uc[:fs.nowned]), f)
    v = 2 * splitmix64_uniform(13, fs.nowned) - 1
    ref = E.jacobian_apply(uc, np.concatenate([v, np.zeros(fs.nclosure - fs.nowned)]))[:fs.nowned]
    y = torch.zeros_like(f)
    op.mat_mult(dev(v), y)
    assert rel(y.cpu().numpy(), ref) < TOL64
    y32 = torch.zeros(fs.nowned, dtype=torch.float32, device="cuda")
    op.mat_mult_f32(dev(v, torch.float32), y32)
    e32 = rel(y32.cpu().numpy(), ref)
    op.set_storage(0)
    yr = torch.zeros_like(f)
    op.mat_mult(dev(v), yr)
    print(f"FP32 error {e32:.2e}; stored vs recomputed {rel(y.cpu().numpy(), yr.cpu().numpy()):.1e}")
    assert e32 < 1e-4
    assert rel(y.cpu().numpy(), yr.cpu().numpy()) < TOL64


def test_stored_quadrature_data_follows_state_and_coefficient(C):
    """The stored data is formed again after each residual (new state) and after a new coefficient."""
    from dohp_amd.mesh import element_vertex_index
    M, deg = (3, 2, 2), 4
    V = block_vertices(M, box=(0, 1, 0, 1, 0, 1), perturb=0.1, seed=29)
    fs, S, E, op = setup(C, M, deg, V, p=1.7, eps=0.4, lam=0.3)
    X = S.nodal_coordinates()
    v = 2 * splitmix64_uniform(37, fs.nowned) - 1
    vc = np.concatenate([v, np.zeros(fs.nclosure - fs.nowned)])
    f = torch.zeros(fs.nowned, dtype=torch.float64, device="cuda")
    y = torch.zeros_like(f)
    y32 = torch.zeros(fs.nowned, dtype=torch.float32, device="cuda")
    for seed in (3, 4, 5):
        uc = sine_state(X, seed)
        fs.dirichlet_commit(uc[fs.nowned:])
        op.function(dev(uc[:fs.nowned]), f)
        ref = E.jacobian_apply(uc, vc)[:fs.nowned]
        if seed != 4:  # the FP32 copy follows too, also when it was not asked for in between
            op.mat_mult_f32(dev(v, torch.float32), y32)
            assert rel(y32.cpu().numpy(), ref) < 1e-4, seed
        op.mat_mult(dev(v), y)
        assert rel(y.cpu().numpy(), ref) < TOL64, seed
    lin = C.Ellip(fs, p=2.0)
    lin.set_storage(1)  # (by default the faster form is timed and kept: here the stored one is wanted)
    for seed in (41, 42):
        kv = np.exp(splitmix64_uniform(seed, V.shape[0]) - 0.5)
        lin.set_coefficient(kv)
        El = O.Ellip(S, p=2.0, kvert=kv[element_vertex_index(M)])
        lin.mat_mult(dev(v), y)
        assert rel(y.cpu().numpy(), El.jacobian_apply(np.zeros(fs.nclosure), vc)[:fs.nowned]) < TOL64, seed


def test_storage_choice_is_timed_and_kept(C):
    """Storage 2 (default) times both forms on the first application in each precision, keeps the faster,
    and the result matches the oracle whichever it kept; 0 and 1 report themselves."""
    print("storage choice", flush=True)
    M, deg = (4, 3, 3), 5
    V = block_vertices(M, box=(0, 1, 0, 1, 0, 1), perturb=0.1, seed=31)
    fs, S, E, op = setup(C, M, deg, V, p=1.5, eps=0.2)
    uc = sine_state(S.nodal_coordinates(), 8)
    fs.dirichlet_commit(uc[fs.nowned:])
    f = torch.zeros(fs.nowned, dtype=torch.float64, device="cuda")
    op.function(dev(uc[:fs.nowned]), f)
    v = 2 * splitmix64_uniform(43, fs.nowned) - 1
    ref = E.jacobian_apply(uc, np.concatenate([v, np.zeros(fs.nclosure - fs.nowned)]))[:fs.nowned]
    assert op.storage(False)[0] == -1 and op.storage(True)[0] == -1
    y = torch.full((fs.nowned,), 3.0, dtype=torch.float64, device="cuda")  # overwritten
    op.mat_mult(dev(v), y)
    y32 = torch.zeros(fs.nowned, dtype=torch.float32, device="cuda")
    op.mat_mult_f32(dev(v, torch.float32), y32)
    for fp32 in (False, True):
        use, ms = op.storage(fp32)
        print("fp32" if fp32 else "fp64", use, ms)
        assert use in (0, 1) and min(ms) > 0 and use == int(ms[1] < ms[0])
    assert rel(y.cpu().numpy(), ref) < TOL64
    assert rel(y32.cpu().numpy(), ref) < 1e-4
    op.mat_mult(dev(v), y)  # the choice is kept: same result
    assert rel(y.cpu().numpy(), ref) < TOL64
    for mode in (0, 1):
        op.set_storage(mode)
        assert op.storage(False)[0] == mode
        op.mat_mult(dev(v), y)
        assert rel(y.cpu().numpy(), ref) < TOL64, mode


@pytest.mark.parametrize("storage", [0, 1])
@pytest.mark.parametrize("p", [2, 3, 4, 6, 7, 9])
def test_every_launch_shape_matches_oracle(C, p, storage):
    """Each launch shape the tuner chooses from (four-wavefront, one-wavefront, their persistent
    forms and, for degrees 3..7, the four-wavefront shape with the contractions as MFMA products),
    pinned in turn, for the residual, the Jacobian action and the linear apply in FP64 (1e-12) and
    FP32 (stated error), recomputed from the state (storage 0) and from stored quadrature data
    (storage 1).  The mesh's 18 elements leave empty slots in the last workgroup of every shape."""
    print(f"shapes p={p} storage={storage}", flush=True)
    M = (3, 2, 3)
    V = block_vertices(M, box=(0, 1, 0, 1, 0, 1), perturb=0.1, seed=17)
    fs, S, E, op = setup(C, M, p, V, p=1.5, eps=np.sqrt(0.005))
    lin = C.Ellip(fs, p=2.0)
    El = O.Ellip(S, p=2.0)
    uc = sine_state(S.nodal_coordinates(), 5)
    fs.dirichlet_commit(uc[fs.nowned:])
    v = 2 * splitmix64_uniform(19, fs.nowned) - 1
    vc = np.concatenate([v, np.zeros(fs.nclosure - fs.nowned)])
    ref_f = E.function(uc)[:fs.nowned]
    ref_j = E.jacobian_apply(uc, vc)[:fs.nowned]
    ref_l = El.jacobian_apply(np.zeros(fs.nclosure), vc)[:fs.nowned]
    op.set_storage(storage)
    lin.set_storage(storage)
    jmode, lmode = ("stored", "stored") if storage else ("jacobian", "linear")
    for shape in range(5 if 3 <= p <= 7 else 4):
        for mode in ("residual", jmode):
            op.set_shape(mode, 0, shape)
            op.set_shape(mode, 1, shape)
        lin.set_shape(lmode, 0, shape)
        lin.set_shape(lmode, 1, shape)
        f = torch.zeros(fs.nowned, dtype=torch.float64, device="cuda")
        op.function(dev(uc[:fs.nowned]), f)
        y = torch.zeros_like(f)
        op.mat_mult(dev(v), y)
        yl = torch.zeros_like(f)
        lin.mat_mult(dev(v), yl)
        y32 = torch.zeros(fs.nowned, dtype=torch.float32, device="cuda")
        op.mat_mult_f32(dev(v, torch.float32), y32)
        yl32 = torch.zeros_like(y32)
        lin.mat_mult_f32(dev(v, torch.float32), yl32)
        errs = [rel(f.cpu().numpy(), ref_f), rel(y.cpu().numpy(), ref_j), rel(yl.cpu().numpy(), ref_l)]
        e32 = [rel(y32.cpu().numpy(), ref_j), rel(yl32.cpu().numpy(), ref_l)]
        print(f"shape {shape}: fp64 {['%.1e' % e for e in errs]} fp32 {['%.1e' % e for e in e32]}")
        assert max(errs) < TOL64, (shape, errs)
        assert max(e32) < 1e-4, (shape, e32)
        assert op.tuning(jmode, 0)[0] == shape and lin.tuning(lmode, 1)[0] == shape


def test_matrix_core_shape_is_refused_outside_its_degrees(C):
    """The MFMA shape exists for degrees 3..7: pinning it elsewhere is an argument error."""
    for p, ok in ((2, False), (3, True), (7, True), (8, False)):
        M = (2, 2, 2)
        fs = C.FS(M, p, block_vertices(M, perturb=0.0, seed=1))
        op = C.Ellip(fs, p=2.0)
        if ok:
            op.set_shape("linear", 0, 4)
        else:
            with pytest.raises(C.DohpError, match="matrix-core"):
                op.set_shape("linear", 0, 4)


@pytest.mark.parametrize("p", [5, 6])
def test_matrix_core_shape_with_colored_scatter(C, p):
    """The MFMA shape under the deterministic (coloured) scatter: equal to the atomic scatter's result
    to rounding, and run to run bitwise."""
    print(f"mfma colours p={p}", flush=True)
    M = (4, 3, 5)
    V = block_vertices(M, perturb=0.1, seed=3)
    fs, S, E, _ = setup(C, M, p, V, p=2.0)
    v = 2 * splitmix64_uniform(23, fs.nowned) - 1
    vc = np.concatenate([v, np.zeros(fs.nclosure - fs.nowned)])
    ref = E.jacobian_apply(np.zeros(fs.nclosure), vc)[:fs.nowned]
    out = {}
    for scatter in (0, 1):
        op = C.Ellip(fs, p=2.0)
        op.set_scatter(scatter)
        op.set_storage(0)
        op.set_shape("linear", 0, 4)
        ys = []
        for _ in range(2):
            y = torch.zeros(fs.nowned, dtype=torch.float64, device="cuda")
            op.mat_mult(dev(v), y)
            ys.append(y.cpu().numpy())
        assert rel(ys[0], ref) < TOL64
        if scatter == 0:
            assert np.array_equal(ys[0], ys[1])
        out[scatter] = ys[0]
    assert rel(out[0], out[1]) < 1e-14


def test_atomic_and_colored_scatter_agree(C):
    M, p = (4, 4, 4), 5
    V = block_vertices(M, perturb=0.1, seed=2)
    fs = C.FS(M, p, V)
    op = C.Ellip(fs, p=2.0)
    x = dev(2 * splitmix64_uniform(4, fs.nowned) - 1)
    y0 = torch.zeros(fs.nowned, dtype=torch.float64, device="cuda")
    y1 = torch.zeros_like(y0)
    op.set_scatter(0)
    op.mat_mult(x, y0)
    op.set_scatter(1)
    op.mat_mult(x, y1)
    assert rel(y1.cpu().numpy(), y0.cpu().numpy()) < 1e-14
    # the colored scatter is deterministic: bitwise identical on repeat
    op.set_scatter(0)
    y2 = torch.zeros_like(y0)
    op.mat_mult(x, y2)
    assert torch.equal(y0, y2)


@pytest.mark.parametrize("storage", [0, 1])
def test_mat_mult_add_accumulates(C, storage):
    """MatMultAdd (the shell's MATOP_MULT_ADD): y0 + J(u) v against the oracle in every launch shape, from
    stored quadrature data and recomputed, on columns of six elements along z (five-element workgroups fold
    the faces they share); refused under the colour scatter, which does not accumulate."""
    M, p = (2, 2, 6), 6
    print(f"MatMultAdd storage={storage}", flush=True)
    V = block_vertices(M, box=(0, 1, 0, 1, 0, 1), perturb=0.1, seed=31)
    fs, S, E, op = setup(C, M, p, V, p=1.5, eps=np.sqrt(0.005))
    uc = sine_state(S.nodal_coordinates(), 5)
    fs.dirichlet_commit(uc[fs.nowned:])
    f = torch.zeros(fs.nowned, dtype=torch.float64, device="cuda")
    op.function(dev(uc[:fs.nowned]), f)
    op.set_storage(storage)
    v = 2 * splitmix64_uniform(17, fs.nowned) - 1
    y0 = 2 * splitmix64_uniform(19, fs.nowned) - 1
    ref = y0 + E.jacobian_apply(uc, np.concatenate([v, np.zeros(fs.nclosure - fs.nowned)]))[:fs.nowned]
    mode = "stored" if storage else "jacobian"
    for shape in range(5):
        op.set_shape(mode, 0, shape)
        y = dev(y0)
        op.mat_mult_add(dev(v), y)
        assert rel(y.cpu().numpy(), ref) < TOL64, shape
        y = torch.zeros(fs.nowned, dtype=torch.float64, device="cuda")
        op.mat_mult(dev(v), y)
        assert rel(y.cpu().numpy() + y0, ref) < TOL64, shape
    op.set_scatter(0)
    with pytest.raises(C.DohpError):
        op.mat_mult_add(dev(v), dev(y0))


@pytest.mark.parametrize("p", list(range(1, 11)))
def test_poisson_apply_all_orders(C, p):
    M = (2, 2, 2)
    V = block_vertices(M, perturb=0.1, seed=p)
    fs, S, E, op = setup(C, M, p, V, p=2.0)
    x = 2 * splitmix64_uniform(p, fs.nowned) - 1
    y = torch.zeros(fs.nowned, dtype=torch.float64, device="cuda")
    op.mat_mult(dev(x), y)
    uc0 = np.zeros(fs.nclosure)
    ref = E.jacobian_apply(uc0, np.concatenate([x, np.zeros(fs.nclosure - fs.nowned)]))[:fs.nowned]
    assert rel(y.cpu().numpy(), ref) < TOL64


def test_variable_coefficient(C):
    M, p = (3, 3, 3), 4
    V = block_vertices(M, perturb=0.1, seed=8)
    kv = np.exp(splitmix64_uniform(12, V.shape[0]) - 0.5)
    fs = C.FS(M, p, V)
    S = O.BlockSpace(M, p, V)
    everts = np.array([[((e0 + (t >> 2)) * (M[1] + 1) + e1 + ((t >> 1) & 1)) * (M[2] + 1) + e2 + (t & 1)
                        for t in range(8)] for e0 in range(M[0]) for e1 in range(M[1]) for e2 in range(M[2])])
    E = O.Ellip(S, p=2.0, kvert=kv[everts])
    op = C.Ellip(fs, p=2.0, kvert=kv)
    x = 2 * splitmix64_uniform(13, fs.nowned) - 1
    y = torch.zeros(fs.nowned, dtype=torch.float64, device="cuda")
    op.mat_mult(dev(x), y)
    ref = E.jacobian_apply(np.zeros(fs.nclosure), np.concatenate([x, np.zeros(fs.nclosure - fs.nowned)]))
    assert rel(y.cpu().numpy(), ref[:fs.nowned]) < TOL64


def test_negative_jacobian_is_an_error(C):
    M = (2, 2, 2)
    V = block_vertices(M)
    V[:, 0] *= -1  # mirrored block: every element is left-handed, det J < 0 everywhere
    fs = C.FS(M, 2, V)
    with pytest.raises(C.DohpError) as ei:
        C.Ellip(fs, p=2.0)
    assert ei.value.code == C.ERR_NEGATIVE_JACOBIAN


def test_fs_numbering_gpu_matches_oracle(C):
    V = block_vertices((3, 2, 4), perturb=0.1, seed=3)
    fs = C.FS((3, 2, 4), 3, V)
    S = O.BlockSpace((3, 2, 4), 3, V)
    np.testing.assert_array_equal(fs.element_nodes(), S.elem_nodes)


# ---------------------------------------------------------- embedded-Q1 matrix
@pytest.mark.parametrize("M,deg,morph,ep,lam", [((4, 4, 4), 1, 0.0, 1.6, 0.0), ((4, 4, 4), 3, 0.0, 1.6, 0.0),
                                                 ((4, 4, 4), 3, 1.0, 2.0, 0.0), ((3, 2, 3), 4, 0.5, 1.7, 1.0),
                                                 ((2, 2, 2), 6, 0.3, 1.5, 0.0)])
def test_q1_assembly_matches_oracle(C, M, deg, morph, ep, lam):
    """Embedded-Q1 matrix incl. the +-1 derivative and the element-ordered inverse Jacobian."""
    print(f"q1 M={M} p={deg} morph={morph}", flush=True)
    V = block_vertices(M, morph=morph, twist=0.4 * morph, perturb=0.05, seed=4)
    fs, S, E, op = setup(C, M, deg, V, p=ep, eps=1.0, lam=lam, gexact=0)
    dv = E.dirichlet_values()
    fs.dirichlet_commit(dv)
    x = 0.3 * (2 * splitmix64_uniform(21, fs.nowned) - 1)
    A = op.assemble_q1(dev(x)).to_scipy()
    R = E.q1_matrix(E.closure(x, dv))
    assert A.shape == R.shape
    d = (A - R).toarray()
    assert np.abs(d).max() <= 1e-12 * np.abs(R.toarray()).max()
    # SpMV and diagonal on the device
    mat = op.assemble_q1(dev(x))
    z = 2 * splitmix64_uniform(22, fs.nowned) - 1
    yd = torch.zeros(fs.nowned, dtype=torch.float64, device="cuda")
    mat.mult(dev(z), yd)
    assert rel(yd.cpu().numpy(), R @ z) < 1e-13
    dd = torch.zeros_like(yd)
    mat.diagonal(dd)
    np.testing.assert_allclose(dd.cpu().numpy(), R.diagonal(), rtol=1e-14)


def test_q1_refout_preconditioned_norms_on_gpu(C):
    """First preconditioned KSP residuals of the refouts with the GPU-assembled Q1 matrix."""
    import scipy.sparse.linalg as spla
    for name, (bdeg, ep, lam, abc, morph, twist) in {
            "ellip-e0-b3-p16": (3, 1.6, 0.0, (1.6, 1.7, 1.8), 0.0, 0.0),
            "ellip-morph": (3, 2.0, 0.0, (1, 1, 1), 1.0, 0.4)}.items():
        V = block_vertices((4, 4, 4), morph=morph, twist=twist)
        fs, S, E, op = setup(C, (4, 4, 4), bdeg, V, p=ep, eps=1.0, lam=lam, gexact=0, a=abc[0], b=abc[1], c=abc[2])
        dv = E.dirichlet_values()
        fs.dirichlet_commit(dv)
        u0 = dev(np.zeros(fs.nowned))
        f = torch.zeros(fs.nowned, dtype=torch.float64, device="cuda")
        op.function(u0, f)
        A = op.assemble_q1(u0).to_scipy()
        r0 = spla.splu(A.tocsc()).solve(f.cpu().numpy())
        assert np.linalg.norm(r0) == pytest.approx(HIST[name]["ksp"][0][0], rel=6e-6)

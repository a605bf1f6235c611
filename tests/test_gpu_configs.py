"""The workload configurations at their full sizes on the GPU.

- order-6 nonlinear diffusion, 64^3 hexahedra: F(u0) and J(u0) v against the CPU oracle (FP64 <= 1e-12
  relative, stored and recomputed Jacobian), the FP32 Jacobian action's error against the FP64 oracle stated;
- batched hp, a 32^3 mesh of degrees 2..10 plus 256 independent problems: FP64 and FP32 against the oracle;
- log-normal Poisson, 48^3, p = 5: plain FP64 CG, FP32-inner/FP64-outer defect correction and FP64 CG with
  the FP32 preconditioner, each to an FP64 relative residual <= 1e-10, with the final error against the
  manufactured solution stated.
"""
import time

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from dohp_amd.mesh import block_vertices, sine_modes_state, splitmix64_uniform  # noqa: E402
from oracle import oracle as O  # noqa: E402


@pytest.fixture(scope="module")
def C():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import dohp_amd.capi as capi
    return capi


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_order6_nonlinear_diffusion_64cube_matches_oracle(C):
    print("order-6 nonlinear diffusion 64^3: setup", flush=True)
    t0 = time.perf_counter()
    M, p = (64, 64, 64), 6
    V = block_vertices(M, box=(0, 1, 0, 1, 0, 1), perturb=0.1, seed=1)
    fs = C.FS(M, p, V)
    X = fs.nodal_coordinates()
    uc = sine_modes_state(X, 3)
    del X
    fs.dirichlet_commit(uc[fs.nowned:])
    prm = dict(p=1.5, eps=float(np.sqrt(0.005)))
    op = C.Ellip(fs, **prm)
    u = torch.as_tensor(uc[:fs.nowned], dtype=torch.float64, device="cuda")
    v = 2 * splitmix64_uniform(9, fs.nowned) - 1
    vd = torch.as_tensor(v, dtype=torch.float64, device="cuda")
    f = torch.zeros_like(u)
    op.function(u, f)
    y = torch.zeros_like(u)
    op.mat_mult(vd, y)  # the default storage keeps the stored quadrature data at this size
    y32 = torch.zeros(fs.nowned, dtype=torch.float32, device="cuda")
    op.mat_mult_f32(vd.float(), y32)
    op.set_storage(0)
    yr = torch.zeros_like(u)
    op.mat_mult(vd, yr)
    yr32 = torch.zeros_like(y32)
    op.mat_mult_f32(vd.float(), yr32)
    storage = op.storage(False)
    gpu = {k: t.double().cpu().numpy() for k, t in (("f", f), ("y", y), ("y32", y32), ("yr", yr), ("yr32", yr32))}
    del op, f, y, y32, yr, yr32, u, vd
    torch.cuda.empty_cache()
    print(f"gpu done {time.perf_counter() - t0:.1f}s; oracle", flush=True)
    S = O.BlockSpace(M, p, V)
    assert (S.nowned, S.nclosure) == (fs.nowned, fs.nclosure)
    E = O.Ellip(S, exact=-1, **prm)
    fr = E.function(uc)[:fs.nowned]
    yref = E.jacobian_apply(uc, np.concatenate([v, np.zeros(fs.nclosure - fs.nowned)]))[:fs.nowned]
    errs = {"F": rel(gpu["f"], fr), "Jv_stored": rel(gpu["y"], yref), "Jv_recomputed": rel(gpu["yr"], yref),
            "Jv_f32_stored": rel(gpu["y32"], yref), "Jv_f32_recomputed": rel(gpu["yr32"], yref)}
    print(f"64^3 p=6 (N_int {fs.nowned}) errors vs the FP64 oracle: {errs}; storage {storage}; "
          f"{time.perf_counter() - t0:.1f}s", flush=True)
    assert errs["F"] <= 1e-12 and errs["Jv_stored"] <= 1e-12 and errs["Jv_recomputed"] <= 1e-12
    # FP32 applications: stated error (rounding of the stored tensor; FP32 state and contractions recomputing)
    assert errs["Jv_f32_stored"] <= 1e-6 and errs["Jv_f32_recomputed"] <= 1e-4


def test_batched_hp_full_size_matches_oracle(C):
    from dohp_amd.mesh import mixed_order_batch
    print("batched hp 32^3 + 256 problems", flush=True)
    deg, vx, probs = mixed_order_batch(32, 256)
    B = C.Batch()
    seg_mixed = B.add_elements(deg, vx)
    segs = []
    for (Mi, pi, Vi) in probs:
        segs.append(B.add_space(C.FS(Mi, pi, Vi)))
    B.setup()
    n = B.size
    xh = 2 * splitmix64_uniform(9, n) - 1
    x = torch.as_tensor(xh, dtype=torch.float64, device="cuda")
    y = torch.zeros_like(x)
    y32 = torch.zeros(n, dtype=torch.float32, device="cuda")
    B.mult(x, y)
    B.mult_f32(x.float(), y32)
    ref = np.zeros(n)
    o, ln = B.segment(seg_mixed)
    ref[o:o + ln] = O.expanded_poisson_apply(deg, vx, xh[o:o + ln])
    for s, (Mi, pi, Vi) in zip(segs, probs):
        o, ln = B.segment(s)
        S = O.BlockSpace(Mi, pi, Vi)
        E = O.Ellip(S, p=2.0)
        ref[o:o + ln] = E.jacobian_apply(np.zeros(S.nclosure),
                                         np.concatenate([xh[o:o + ln], np.zeros(S.nclosure - S.nowned)]))[:ln]
    e64, e32 = rel(y.cpu().numpy(), ref), rel(y32.double().cpu().numpy(), ref)
    print(f"batch ({n} unknowns) FP64 err {e64:.2e}, FP32 err {e32:.2e} vs the FP64 oracle", flush=True)
    assert e64 <= 1e-12 and e32 <= 1e-5


@pytest.fixture(scope="module")
def lognormal48(C):
    from dohp_amd.fields import lognormal_field
    M, p = (48, 48, 48), 5
    V = block_vertices(M, box=(0, 1, 0, 1, 0, 1), perturb=0.1, seed=1)
    fs = C.FS(M, p, V)
    op = C.Ellip(fs, p=2.0, kvert=lognormal_field(V, seed=11, corr_len=0.1))
    X = fs.nodal_coordinates()[:fs.nowned]
    xs = torch.as_tensor(sine_modes_state(X, 5) + X[:, 0] * (1 - X[:, 0]), dtype=torch.float64, device="cuda")
    b = torch.zeros_like(xs)
    op.mat_mult(xs, b)
    q1 = op.assemble_q1(torch.zeros_like(xs))
    return fs, op, q1, b, xs


@pytest.mark.parametrize("mixed", [0, 1, 2])
def test_lognormal_poisson_48cube_p5_reaches_1e10(C, lognormal48, mixed):
    fs, op, q1, b, xs = lognormal48
    print(f"log-normal 48^3 p=5 (N_int {fs.nowned}) mixed={mixed}", flush=True)
    x = torch.zeros_like(b)
    st = op.solve(q1, b, x, mixed=mixed, rtol=1e-10)
    r = torch.zeros_like(b)
    op.mat_mult(x, r)
    relres = float(torch.linalg.norm(b - r) / torch.linalg.norm(b))
    err = float(torch.linalg.norm(x - xs) / torch.linalg.norm(xs))
    print(f"  its {st['inner_its']} (outer {st['outer_its']}), {st['seconds']:.3f}s, FP64 relres {relres:.2e}, "
          f"error vs manufactured {err:.2e}", flush=True)
    assert relres <= 1e-10
    assert err <= 1e-8


def test_learned_preconditioner_48cube_p6_beats_plain(C):
    """Learned sub-cell scales of the embedded-Q1 matrix, trained on residual/error pairs of plain solves of 12^3
    problems, on the 48^3, p = 6 log-normal problem: fewer CG iterations and less time to 1e-10 than the plain FP64
    solve, with the plain matrix as the fallback after the plain iteration count."""
    import time
    from dohp_amd.learned import ScaledProblem, fit_q1_scales
    t0 = time.perf_counter()
    net, rep = fit_q1_scales(C, torch, train_seeds=[101, 102, 103], val_seeds=[201], M=(12, 12, 12), deg=6, epochs=25,
                             lr=1e-2, npairs=8, every=4, nrandom=2, val_M=(24, 24, 24))
    print(f"training {time.perf_counter() - t0:.1f}s:", {k: v for k, v in rep.items() if k != "history"}, flush=True)
    assert net is not None
    P = ScaledProblem(C, torch, (48, 48, 48), 6, 11)

    def run(q1, **kw):
        x = torch.zeros_like(P.b)
        P.op.solve(q1, P.b, x, rtol=1e-10, **kw)  # warm-up: kernels, hierarchy
        x.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        st = P.op.solve(q1, P.b, x, rtol=1e-10, **kw)
        torch.cuda.synchronize()
        t = time.perf_counter() - t
        r = torch.zeros_like(x)
        P.op.mat_mult(x, r)
        relres = float(torch.linalg.norm(P.b - r) / torch.linalg.norm(P.b))
        err = float(torch.linalg.norm(x - P.xs) / torch.linalg.norm(P.xs))
        return st["inner_its"], st["fallback"], t, relres, err

    plain = run(P.q1)
    q1s = P.assemble(net)
    learned = run(q1s, fallback_q1=P.q1, fallback_its=plain[0])
    learned32 = run(q1s, fallback_q1=P.q1, fallback_its=plain[0], mixed=2)
    for name, r in (("plain FP64", plain), ("learned FP64", learned), ("learned FP64 CG + FP32 V-cycle", learned32)):
        print(f"48^3 p=6 {name}: its {r[0]}, fallback {r[1]}, {r[2]:.3f}s, relres {r[3]:.2e}, error {r[4]:.2e}",
              flush=True)
    for r in (plain, learned, learned32):
        assert r[3] <= 1e-10
    assert learned[0] < plain[0] and learned[1] == 0
    assert learned32[2] < plain[2]

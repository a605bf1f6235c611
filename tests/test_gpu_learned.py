"""Learned node weights of the Q1 multigrid preconditioner: HIP features, weights and training
gradients against the numpy oracle and finite differences; training lowers the energy error of
held-out pairs; learned solves converge to the FP64 tolerance; the fallback engages."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import mlp_oracle as L  # noqa: E402


@pytest.fixture(scope="module")
def C():
    assert torch.cuda.is_available()
    import dohp_amd.capi as capi
    return capi


@pytest.fixture(scope="module")
def small(C):
    from dohp_amd.learned import make_problem
    return make_problem(C, (4, 4, 3), 4, field_seed=3, torch=torch)


def _t(a, dt=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device="cuda")


def test_gradient_matches_oracle(C):
    rng = np.random.default_rng(0)
    w = (rng.standard_normal(L.NPARAM) * 0.2).astype(np.float32)
    B = 300
    X = rng.standard_normal((B, L.NIN)).astype(np.float32)
    T = rng.standard_normal((B, L.NOUT)).astype(np.float32)
    net = C.MLP(w)
    loss, g = net.gradient(_t(X, torch.float32), _t(T, torch.float32))
    lref, gref = L.loss_and_grad(L.bf16(w), X.astype(np.float64), T.astype(np.float64))
    print("loss", loss, lref, "grad rel err", np.linalg.norm(g - gref) / np.linalg.norm(gref))
    assert abs(loss - lref) <= 1e-5 * lref
    assert np.linalg.norm(g - gref) <= 1e-5 * np.linalg.norm(gref)
    m, wb = net.params()
    np.testing.assert_array_equal(m, w)
    np.testing.assert_array_equal(wb, L.bf16(w))


def test_features_and_weights_match_oracle(C, small):
    fs, op, q1, b, xs = small
    box = C.MG(q1).level(0)[0]
    n = fs.nowned
    diag = torch.zeros(n, dtype=torch.float64, device="cuda")
    q1.diagonal(diag)
    dinv = (1.0 / diag).cpu().numpy()
    npatch = C.MLP.num_patches(q1)
    X = torch.zeros((npatch, L.NIN), dtype=torch.float32, device="cuda")
    C.MLP.features(q1, fs, _t(dinv), X)
    Xr = L.features(box, 4, (1, 1, 1), dinv)  # Dirichlet on every face: the owned box starts at lattice 1
    np.testing.assert_allclose(X.cpu().numpy(), Xr, rtol=1e-5, atol=1e-5)
    w = (np.random.default_rng(1).standard_normal(L.NPARAM) * 0.3).astype(np.float32)
    net = C.MLP(w)
    s = torch.zeros(n, dtype=torch.float64, device="cuda")
    net.weights(q1, X, s)
    sr = L.node_weights(box, w, X.cpu().numpy().astype(np.float64))
    np.testing.assert_allclose(s.cpu().numpy(), sr, rtol=2e-5)
    assert float(s.min()) > 0


def test_energy_gradient_matches_finite_differences(C, small):
    fs, op, q1, b, xs = small
    mg = C.MG(q1)
    n = fs.nowned
    rng = np.random.default_rng(4)
    e = _t(rng.standard_normal(n))
    r = torch.zeros_like(e)
    op.mat_mult(e, r)
    s = _t(np.exp(0.3 * rng.standard_normal(n)))
    for alpha in (0.0, 0.05):  # the pair's optimal scale (stationary), and a fixed one
        gs = torch.zeros_like(s)
        E0, terms = C.MLP.energy_weight_gradient(op, mg, s, r, e, alpha, gs)
        assert E0 > 0
        for k in range(3):
            v = _t(rng.standard_normal(n))
            h = 1e-5
            a = alpha if alpha > 0 else 0.0
            Ep, _ = C.MLP.energy_weight_gradient(op, mg, s + h * v, r, e, a)
            Em, _ = C.MLP.energy_weight_gradient(op, mg, s - h * v, r, e, a)
            fd, an = (Ep - Em) / (2 * h), float(torch.dot(gs, v))
            print("alpha", alpha, "directional derivative", fd, an)
            assert abs(fd - an) <= 1e-5 * (abs(fd) + 1e-3 * E0)


def test_training_and_learned_solves(C, small):
    from dohp_amd.learned import fit, init_params
    net, rep = fit(C, torch, train_seeds=[11, 12], val_seeds=[21], M=(6, 6, 6), deg=4, epochs=8, lr=3e-3,
                   lanczos=20)
    print({k: v for k, v in rep.items() if k != "history"})
    for h in rep["history"]:
        print(h)
    # the spectral spread of the training problems' preconditioned operators decreases
    assert rep["history"][-1]["train_log_spread"] < rep["history"][0]["train_log_spread"]
    if net is None:  # no epoch beat the plain preconditioner on the validation problem
        net = C.MLP(init_params(C.MLP.nparam))
    fs, op, q1, b, xs = small
    res = {}
    for name, kw in [("plain", {}), ("learned", {"mlp": net}), ("learned_mixed", {"mlp": net, "mixed": 1}),
                     ("learned_f32_precond", {"mlp": net, "mixed": 2}), ("fallback", {"mlp": net, "fallback_its": 2}),
                     ("fallback_f32_precond", {"mlp": net, "fallback_its": 2, "mixed": 2})]:
        x = torch.zeros_like(b)
        st = op.solve(q1, b, x, rtol=1e-10, pc="mg", **kw)
        r = torch.zeros_like(b)
        op.mat_mult(x, r)
        relres = float(torch.linalg.norm(b - r) / torch.linalg.norm(b))
        print(name, st["inner_its"], st["outer_its"], st["fallback"], relres)
        assert relres <= 1.01e-10
        res[name] = st
    assert res["fallback"]["fallback"] == 1 and res["learned"]["fallback"] == 0
    assert res["fallback_f32_precond"]["fallback"] == 1 and res["learned_f32_precond"]["fallback"] == 0


# ------------------------------------------------- learned sub-cell scales of the embedded-Q1 matrix (kind 1)
@pytest.fixture(scope="module")
def scaled(C):
    from dohp_amd.mesh import block_vertices, element_vertex_index
    from oracle import oracle as O
    from oracle.q1_scaled import ScaledQ1
    M, p = (2, 3, 2), 3
    V = block_vertices(M, box=(0, 1, 0, 1, 0, 1), perturb=0.1, seed=7)
    k = np.exp(0.3 * np.sin(3 * V[:, 0]) * np.cos(2 * V[:, 1]))
    fs = C.FS(M, p, V)
    op = C.Ellip(fs, p=2.0, kvert=k)
    q1 = op.assemble_q1(torch.zeros(fs.nowned, dtype=torch.float64, device="cuda"))
    S = O.BlockSpace(M, p, V)
    kel = k[element_vertex_index(M)]
    return fs, op, q1, ScaledQ1(S, kel), kel


def test_cell_features_match_oracle(C, scaled):
    fs, op, q1, Q, kel = scaled
    X = torch.zeros((op.q1_cell_count(), C.CellMLP.ninputs), dtype=torch.float32, device="cuda")
    op.q1_cell_features(q1, X)
    np.testing.assert_allclose(X.cpu().numpy(), Q.features(kel), rtol=1e-5, atol=2e-6)


def test_scaled_q1_and_gradient_match_oracle(C, scaled):
    fs, op, q1, Q, kel = scaled
    rng = np.random.default_rng(5)
    sig = np.exp(0.4 * rng.standard_normal((op.q1_cell_count(), 3))).astype(np.float32)
    sd = _t(sig, torch.float32)
    op.set_q1_cell_scales(sd)
    qs = op.assemble_q1(torch.zeros(fs.nowned, dtype=torch.float64, device="cuda"))
    ref = Q.matrix(sig.astype(np.float64)).toarray()
    got = qs.to_scipy().toarray()
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    w, z = rng.standard_normal(fs.nowned), rng.standard_normal(fs.nowned)
    g = torch.zeros((op.q1_cell_count(), 3), dtype=torch.float32, device="cuda")
    op.q1_cell_scale_gradient(qs, _t(w), _t(z), 0.5, g)
    op.q1_cell_scale_gradient(qs, _t(w), _t(z), 0.5, g)  # accumulates
    gr = Q.scale_gradient(sig.astype(np.float64), w, z)
    np.testing.assert_allclose(g.cpu().numpy(), gr, rtol=2e-5, atol=1e-6 * np.abs(gr).max())
    op.set_q1_cell_scales(None)
    q0 = op.assemble_q1(torch.zeros(fs.nowned, dtype=torch.float64, device="cuda"))
    np.testing.assert_allclose(q0.to_scipy().toarray(), q1.to_scipy().toarray(), rtol=0, atol=0)


def test_cell_network_matches_oracle(C):
    rng = np.random.default_rng(6)
    w = (rng.standard_normal(L.nparam(1)) * 0.2).astype(np.float32)
    X = rng.standard_normal((500, 11)).astype(np.float32)
    T = rng.standard_normal((500, 3)).astype(np.float32)
    net = C.CellMLP(w)
    loss, g = net.gradient(_t(X, torch.float32), _t(T, torch.float32))
    lref, gref = L.loss_and_grad(L.bf16(w), X.astype(np.float64), T.astype(np.float64), kind=1)
    assert abs(loss - lref) <= 1e-5 * lref
    assert np.linalg.norm(g - gref) <= 1e-5 * np.linalg.norm(gref)
    sig = torch.zeros((500, 3), dtype=torch.float32, device="cuda")
    net.scales(_t(X, torch.float32), sig)
    np.testing.assert_allclose(sig.cpu().numpy(), L.cell_scales(w, X.astype(np.float64)), rtol=2e-5)


def test_learned_q1_scales_train_and_beat_plain(C):
    """Trained on residual/error pairs of plain solves (two 6^3 problems), the learned embedded-Q1 matrix preconditions
    a held-out 8^3 problem in fewer CG iterations than the plain one; every solve reaches 1e-10; the fallback engages
    when its budget is short."""
    from dohp_amd.learned import ScaledProblem, fit_q1_scales
    net, rep = fit_q1_scales(C, torch, train_seeds=[31, 32], val_seeds=[41], M=(6, 6, 6), deg=5, epochs=15, lr=1e-2,
                             npairs=6, every=3, nrandom=2)
    print({k: v for k, v in rep.items() if k != "history"})
    for h in rep["history"]:
        print(h)
    assert net is not None, "no epoch beat the plain matrix on the validation problem"
    P = ScaledProblem(C, torch, (8, 8, 8), 5, 51)
    plain = P.cg_its()
    q1s = P.assemble(net)
    res = {}
    for name, kw in [("learned", {}), ("learned_mixed", {"mixed": 1}), ("learned_f32_precond", {"mixed": 2}),
                     ("fallback", {"fallback_its": 2}), ("fallback_f32_precond", {"fallback_its": 2, "mixed": 2})]:
        x = torch.zeros_like(P.b)
        st = P.op.solve(q1s, P.b, x, rtol=1e-10, pc="mg", fallback_q1=P.q1, **kw)
        r = torch.zeros_like(P.b)
        P.op.mat_mult(x, r)
        relres = float(torch.linalg.norm(P.b - r) / torch.linalg.norm(P.b))
        print(name, st["inner_its"], st["outer_its"], st["fallback"], relres)
        assert relres <= 1.01e-10
        res[name] = st
    print("held-out 8^3 p=5: plain", plain, "learned", res["learned"]["inner_its"])
    assert res["learned"]["inner_its"] < plain
    assert res["fallback"]["fallback"] == 1 and res["learned"]["fallback"] == 0
    assert res["fallback_f32_precond"]["fallback"] == 1

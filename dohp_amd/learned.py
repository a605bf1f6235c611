"""Training driver of the learned node weights of the Q1 multigrid preconditioner (no counterpart
in the reference).

CG converges at a rate set by the spread of the spectrum of the preconditioned operator
P = S M S J (M the multigrid V-cycle on the Q1 matrix, S the learned node weights).  Training
therefore minimizes log(lambda_max / lambda_min) of P: a Lanczos run in the J inner product on
seeded log-normal training problems gives the extreme Ritz pairs (x, theta), and since
theta = (S r, M S r) / (x, r) with r = J x is a Rayleigh quotient of the symmetric pencil, its
gradient in the node weights is  d log(theta)/ds = 2 r .* M(S r) / (S r, M S r)  (Hellmann-Feynman);
the network's HIP kernels backpropagate it and take an Adam step.  The parameters with the fewest
CG iterations on held-out problems are kept.  Residual/error pairs of plain solves (the iterates'
errors) and their energy error remain available for diagnostics.  torch holds device buffers and
does the Lanczos vector algebra of the training loop.
"""
from __future__ import annotations

import numpy as np

from .fields import lognormal_field
from .mesh import block_vertices, splitmix64_uniform


def make_problem(C, M, deg, field_seed, torch, rhs_seed=5, corr_len=0.1, q1_mode=0):
    """Log-normal Poisson problem on an M box of order deg: (fs, op, q1, b, x_exact).  q1_mode 1
    assembles the corrected embedded-Q1 matrix instead of the reference's arithmetic."""
    V = block_vertices(M, box=(0, 1, 0, 1, 0, 1), perturb=0.1, seed=field_seed + 1000)
    fs = C.FS(M, deg, V)
    op = C.Ellip(fs, p=2.0, kvert=lognormal_field(V, seed=field_seed, corr_len=corr_len))
    xs = torch.as_tensor(2 * splitmix64_uniform(rhs_seed + field_seed, fs.nowned) - 1, dtype=torch.float64,
                         device="cuda")
    b = torch.zeros_like(xs)
    op.mat_mult(xs, b)
    op.set_q1_mode(q1_mode)
    q1 = op.assemble_q1(torch.zeros_like(xs))
    return fs, op, q1, b, xs


def problem_features(C, torch, fs, q1):
    diag = torch.zeros(fs.nowned, dtype=torch.float64, device="cuda")
    q1.diagonal(diag)
    npatch = C.MLP.num_patches(q1)
    X = torch.zeros((npatch, C.MLP.ninputs), dtype=torch.float32, device="cuda")
    C.MLP.features(q1, fs, 1.0 / diag, X)
    return X


def collect_pairs(C, torch, seeds, M=(12, 12, 12), deg=6, every=4, nsnap=8, nrandom=4, q1_mode=0):
    """Per problem: op, q1, multigrid, patch features and residual/error pairs [(r, e)]: errors of
    plain-CG iterates (the modes CG converges on last) and seeded random errors (every mode)."""
    probs = []
    for sd in seeds:
        fs, op, q1, b, xs = make_problem(C, M, deg, sd, torch, q1_mode=q1_mode)
        snap = torch.zeros((nsnap, fs.nowned), dtype=torch.float64, device="cuda")
        x = torch.zeros_like(b)
        st = op.solve(q1, b, x, pc="mg", rtol=1e-11, record=(snap, every))
        errs = [x - snap[k] for k in range(st["nsnap"])]
        for k in range(nrandom):
            errs.append(torch.as_tensor(2 * splitmix64_uniform(7919 * sd + k, fs.nowned) - 1, dtype=torch.float64,
                                        device="cuda"))
        pairs = []
        for e in errs:
            r = torch.zeros_like(e)
            op.mat_mult(e, r)  # r = J e (= b - J x_k for the iterates)
            pairs.append((r, e))
        # the V-cycle of the solves (smoothing degree 1 on the corrected Q1 matrix, 2 on the reference's)
        mg = C.MG(q1, smooth_degree=1 if q1_mode else 2)
        probs.append({"fs": fs, "op": op, "q1": q1, "mg": mg, "X": problem_features(C, torch, fs, q1),
                      "pairs": pairs, "b": b, "its": st["inner_its"]})
    return probs


def _weights(torch, net, p):
    s = torch.ones(p["fs"].nowned, dtype=torch.float64, device="cuda")
    if net is not None:
        net.weights(p["q1"], p["X"], s)
    return s


def shared_scale_energy(C, torch, net, probs):
    """(alpha, mean relative energy error) with alpha the optimal scale shared by all pairs."""
    terms = []
    for p in probs:
        s = _weights(torch, net, p)
        for (r, e) in p["pairs"]:
            terms.append((p, s, r, e, C.MLP.energy_weight_gradient(p["op"], p["mg"], s, r, e)[1]))
    alpha = sum(t[4][0] / t[4][2] for t in terms) / sum(t[4][1] / t[4][2] for t in terms)
    E = [C.MLP.energy_weight_gradient(p["op"], p["mg"], s, r, e, alpha)[0] for (p, s, r, e, _) in terms]
    return alpha, float(np.mean(np.sqrt(E)))


def cg_iterations(C, torch, net, probs):
    its = []
    for p in probs:
        x = torch.zeros_like(p["b"])
        st = p["op"].solve(p["q1"], p["b"], x, pc="mg", rtol=1e-10, mlp=net)
        its.append(st["inner_its"])
    return its


def lanczos_extremes(C, torch, p, s, m=30, nvec=3, seed=0):
    """Extreme Ritz pairs of P = S M S J (self-adjoint in the J inner product), m Lanczos steps with
    full reorthogonalization.  Returns ([(theta, x, Jx)] lowest nvec, [...] highest nvec)."""
    op, mg = p["op"], p["mg"]
    n = p["fs"].nowned
    x0 = torch.as_tensor(2 * splitmix64_uniform(104729 + seed, n) - 1, dtype=torch.float64, device="cuda")
    V, JV = [], []
    v = x0
    Jv = torch.zeros_like(v)
    op.mat_mult(v, Jv)
    nrm = float(torch.dot(v, Jv)) ** 0.5
    v, Jv = v / nrm, Jv / nrm
    alphas, betas = [], []
    t = torch.zeros_like(v)
    w = torch.zeros_like(v)
    for j in range(m):
        V.append(v)
        JV.append(Jv)
        mg.apply(s * Jv, t)
        w = s * t  # P v = S M S J v
        for _ in range(2):  # full reorthogonalization in the J inner product
            for vi, Jvi in zip(V, JV):
                w = w - float(torch.dot(w, Jvi)) * vi
        Jw = torch.zeros_like(w)
        op.mat_mult(w, Jw)
        alphas.append(float(torch.dot(s * t, Jv)))
        b = float(torch.dot(w, Jw)) ** 0.5
        if j + 1 == m or b < 1e-12:
            break
        betas.append(b)
        v, Jv = w / b, Jw / b
    k = len(alphas)
    T = np.diag(alphas) + np.diag(betas[:k - 1], 1) + np.diag(betas[:k - 1], -1)
    th, Y = np.linalg.eigh(T)
    Vm = torch.stack(V[:k])
    JVm = torch.stack(JV[:k])
    def ritz(i):
        y = torch.as_tensor(Y[:, i], dtype=torch.float64, device="cuda")
        return (float(th[i]), y @ Vm, y @ JVm)
    return [ritz(i) for i in range(nvec)], [ritz(k - 1 - i) for i in range(nvec)]


def spread_gradient(C, torch, p, s, low, high):
    """log(mean high theta) - log(mean low theta) estimate and its gradient in s."""
    mg = p["mg"]
    gs = torch.zeros_like(s)
    u = torch.zeros_like(s)
    vals = {}
    for sign, group in ((1.0, high), (-1.0, low)):
        acc = []
        for (_, x, r) in group:
            mg.apply(s * r, u)
            q = float(torch.dot(s * r, u))
            acc.append(q / float(torch.dot(x, r)))
            gs += (sign / len(group)) * 2.0 * r * u / q
        vals[sign] = float(np.mean(acc))
    return float(np.log(vals[1.0] / vals[-1.0])), gs


def init_params(nparam, seed=0):
    """Xavier-uniform hidden layers; the output layer starts at zero (all weights 1)."""
    nin, nh, nout = 32, 64, 8
    rng = np.random.default_rng(seed)
    w1 = rng.uniform(-1, 1, (nh, nin)) * np.sqrt(6.0 / (nin + nh))
    w2 = rng.uniform(-1, 1, (nh, nh)) * np.sqrt(6.0 / (2 * nh))
    p = np.concatenate([w1.ravel(), np.zeros(nh), w2.ravel(), np.zeros(nh), np.zeros(nout * nh), np.zeros(nout)])
    assert p.size == nparam
    return p.astype(np.float32)


def fit(C, torch, train_seeds, val_seeds, M=(12, 12, 12), deg=6, epochs=30, lr=2e-3, seed=0, lanczos=30, nvec=3,
        q1_mode=0, val_M=None):
    """Minimize the spectral spread of S M S J on the training problems (one Adam step per problem
    and epoch, fresh Ritz vectors each time); keep the parameters with the fewest CG iterations on
    the validation problems (of size val_M, default M: validating at the size the network will be
    used at guards against weights that help small problems only).  Returns (net or None if no
    epoch beat the plain preconditioner, report)."""
    probs = collect_pairs(C, torch, train_seeds, M, deg, nsnap=1, nrandom=0, q1_mode=q1_mode)
    val = collect_pairs(C, torch, val_seeds, val_M or M, deg, nsnap=1, nrandom=0, q1_mode=q1_mode)
    net = C.MLP(init_params(C.MLP.nparam, seed))
    plain_its = cg_iterations(C, torch, None, val)
    best = (sum(plain_its), None, -1)
    hist = []
    for ep in range(epochs):
        spreads = []
        for i, p in enumerate(probs):
            s = _weights(torch, net, p)
            low, high = lanczos_extremes(C, torch, p, s, lanczos, nvec, seed=ep * 131 + i)
            L, gs = spread_gradient(C, torch, p, s, low, high)
            spreads.append(L)
            net.weight_step(p["q1"], p["X"], gs, lr)
        its = cg_iterations(C, torch, net, val)
        hist.append({"epoch": ep, "train_log_spread": float(np.mean(spreads)), "val_cg_its": its})
        if sum(its) < best[0]:
            best = (sum(its), net.params()[0].copy(), ep)
    final = C.MLP(best[1]) if best[1] is not None else None
    report = {"train_problems": len(probs), "val_cg_its_plain": plain_its, "best_epoch": best[2],
              "val_cg_its_learned": best[0] if final else None, "history": hist, "useful": final is not None}
    return final, report

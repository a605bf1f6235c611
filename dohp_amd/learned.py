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


# ----------------------------------------------------------------------------------------------------------------
# Learned sub-cell scales of the embedded-Q1 matrix (network kind 1).
#
# The embedded-Q1 matrix in the reference's arithmetic takes sub-cell derivatives of +-1 per element-reference unit
# whatever the width of the GLL sub-interval, which makes it a poor spectral match for the high-order operator away
# from the element centres.  A 3-layer perceptron maps the features of each GLL sub-cell (its reference widths,
# position, physical shape and the local coefficient) to one scale per reference direction multiplying those
# derivatives.  It is trained on residual/error pairs (r, e) of plain solves: the preconditioner z = M(sigma) r (the
# multigrid V-cycle on the matrix assembled with the scales) should reproduce e in the energy norm of the high-order
# operator A, up to one scale alpha shared by the pairs:
#     E = mean_pairs || e - alpha z ||_A^2 / || e ||_A^2.
# With dM = -M dK M (M an approximate inverse of the Q1 matrix K), dE/dsigma = mean 2 alpha w^T (dK/dsigma) z /
# ||e||_A^2 with w = M A (e - alpha z); dohpQ1CellScaleGradient evaluates the bilinear form per sub-cell, the network's
# HIP kernels backpropagate it and take an Adam step.  The solve then preconditions with the learned matrix; the plain
# matrix takes over after as many iterations as the plain solve needed (fallback).

def cell_init_params(nparam, seed=0):
    """Xavier-uniform hidden layers, zero output layer: every scale 1 (the plain matrix)."""
    nin, nh, nout = 11, 64, 3
    rng = np.random.default_rng(seed)
    w1 = rng.uniform(-1, 1, (nh, nin)) * np.sqrt(6.0 / (nin + nh))
    w2 = rng.uniform(-1, 1, (nh, nh)) * np.sqrt(6.0 / (2 * nh))
    p = np.concatenate([w1.ravel(), np.zeros(nh), w2.ravel(), np.zeros(nh), np.zeros(nout * nh), np.zeros(nout)])
    assert p.size == nparam
    return p.astype(np.float32)


class ScaledProblem:
    """A log-normal Poisson problem with its plain embedded-Q1 matrix, sub-cell features, the learned matrix
    (refilled per set of scales) and, for training, residual/error pairs of a plain solve."""

    def __init__(self, C, torch, M, deg, seed, q1_mode=0, npairs=0, every=3, nrandom=0):
        self.C, self.torch = C, torch
        self.fs, self.op, self.q1, self.b, self.xs = make_problem(C, M, deg, seed, torch, q1_mode=q1_mode)
        n = self.fs.nowned
        self.ncell = self.op.q1_cell_count()
        self.X = torch.zeros((self.ncell, C.CellMLP.ninputs), dtype=torch.float32, device="cuda")
        self.op.q1_cell_features(self.q1, self.X)
        self.sig = torch.ones((self.ncell, 3), dtype=torch.float32, device="cuda")
        self.q1s = None
        self.pairs = []
        self.plain_its = None
        if npairs:
            snap = torch.zeros((npairs, n), dtype=torch.float64, device="cuda")
            x = torch.zeros_like(self.b)
            st = self.op.solve(self.q1, self.b, x, pc="mg", rtol=1e-11, record=(snap, every))
            self.plain_its = st["inner_its"]
            errs = [x - snap[k] for k in range(st["nsnap"])]
            for k in range(nrandom):
                errs.append(torch.as_tensor(2 * splitmix64_uniform(7919 * seed + k, n) - 1, dtype=torch.float64,
                                            device="cuda"))
            for e in errs:
                r = torch.zeros_like(e)
                self.op.mat_mult(e, r)  # r = A e (= b - A x_k for the iterates)
                self.pairs.append((r, e))

    def assemble(self, net):
        """Learned matrix for the network's scales (net None: unit scales)."""
        if net is None:
            self.sig.fill_(1.0)
        else:
            net.scales(self.X, self.sig)
        self.op.set_q1_cell_scales(self.sig)
        self.q1s = self.op.assemble_q1(torch_zeros(self.torch, self.b), mat=self.q1s)
        self.op.set_q1_cell_scales(None)
        return self.q1s

    def energy_gradient(self, gsig, smooth_degree=2):
        """Mean relative energy error of the learned V-cycle (Chebyshev smoothing of the given degree) over the pairs
        (shared optimal scale) and its gradient in the sub-cell scales added into gsig."""
        torch, C = self.torch, self.C
        if getattr(self, "_mg", None) is None or self._mg_degree != smooth_degree:
            self._mg, self._mg_degree = C.MG(self.q1s, smooth_degree=smooth_degree), smooth_degree
        else:  # the learned matrix was refilled (same box): recompute the hierarchy in place
            self._mg.refresh(self.q1s)
        mg = self._mg
        zs, terms = [], []
        for (r, e) in self.pairs:
            z = torch.zeros_like(r)
            mg.apply(r, z)
            Az = torch.zeros_like(z)
            self.op.mat_mult(z, Az)
            eAe = float(torch.dot(e, r))
            terms.append((float(torch.dot(z, r)) / eAe, float(torch.dot(z, Az)) / eAe, eAe))
            zs.append(z)
        alpha = sum(t[0] for t in terms) / sum(t[1] for t in terms)
        Es = []
        self.op.set_q1_cell_scales(self.sig)  # the gradient is taken at the scales of the assembled matrix
        for (r, e), z, (_, _, eAe) in zip(self.pairs, zs, terms):
            d = e - alpha * z
            Ad = torch.zeros_like(d)
            self.op.mat_mult(d, Ad)
            Es.append(float(torch.dot(d, Ad)) / eAe)
            w = torch.zeros_like(d)
            mg.apply(Ad, w)
            self.op.q1_cell_scale_gradient(self.q1s, w, z, 2.0 * alpha / (eAe * len(self.pairs)), gsig)
        self.op.set_q1_cell_scales(None)
        return float(np.mean(Es))

    def cg_its(self, q1=None, fallback_its=0, **kw):
        x = self.torch.zeros_like(self.b)
        if q1 is None:
            st = self.op.solve(self.q1, self.b, x, pc="mg", rtol=1e-10, **kw)
        else:
            st = self.op.solve(q1, self.b, x, pc="mg", rtol=1e-10, fallback_q1=self.q1, fallback_its=fallback_its, **kw)
        return st["inner_its"]


def torch_zeros(torch, like):
    return torch.zeros_like(like)


def fit_q1_scales(C, torch, train_seeds, val_seeds, M=(12, 12, 12), deg=6, epochs=30, lr=3e-3, seed=0, npairs=8,
                  every=3, nrandom=2, val_M=None, scale_penalty=0.0, train_degree=1, q1_mode=0):
    """Train the sub-cell scale network on residual/error pairs of plain solves of the training problems (one Adam
    step per problem and epoch on the mean energy error of its pairs); keep the parameters with the fewest CG
    iterations on the validation problems (of size val_M, default M).  scale_penalty adds
    0.5 * scale_penalty * (mean log sigma)^2: the loss does not see a common factor of all scales (alpha absorbs it),
    and the penalty keeps that factor near 1 instead of drifting to the clamp.  train_degree: smoothing degree of the
    V-cycle inside the loss (1, the degree the learned matrix is solved with).  q1_mode 1 learns scales on top of the
    corrected embedded-Q1 arithmetic (and compares with the corrected matrix).  Returns (net or None if no epoch beat the plain matrix, report)."""
    train = [ScaledProblem(C, torch, M, deg, sd, q1_mode=q1_mode, npairs=npairs, every=every, nrandom=nrandom)
             for sd in train_seeds]
    val = [ScaledProblem(C, torch, val_M or M, deg, sd, q1_mode=q1_mode) for sd in val_seeds]
    plain_its = [v.cg_its() for v in val]
    net = C.CellMLP(cell_init_params(C.CellMLP.nparam, seed))
    best = (sum(plain_its), None, -1)
    hist = []
    for ep in range(epochs):
        Es = []
        for p in train:
            p.assemble(net)
            gsig = torch.zeros((p.ncell, 3), dtype=torch.float32, device="cuda")
            Es.append(p.energy_gradient(gsig, train_degree))
            if scale_penalty:
                ls = torch.log(p.sig)
                gsig += scale_penalty * float(ls.mean()) / ls.numel() / p.sig
            net.step(p.X, gsig, lr)
        its = []
        for v in val:
            its.append(v.cg_its(v.assemble(net)))
        sig = val[0].sig
        hist.append({"epoch": ep, "train_energy": float(np.mean(Es)), "val_cg_its": its,
                     "sigma_range": [float(sig.min()), float(sig.max())]})
        if sum(its) < best[0]:
            best = (sum(its), net.params()[0].copy(), ep)
    final = C.CellMLP(best[1]) if best[1] is not None else None
    npair_total = sum(len(p.pairs) for p in train)
    report = {"train_problems": len(train), "train_pairs": npair_total,
              "train_pair_unknowns": int(sum(len(p.pairs) * p.fs.nowned for p in train)),
              "val_cg_its_plain": plain_its, "best_epoch": best[2], "val_cg_its_learned": best[0] if final else None,
              "history": hist, "useful": final is not None}
    return final, report

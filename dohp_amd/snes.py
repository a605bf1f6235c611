"""Newton-Krylov solve of ellip's nonlinear problem around the library's GPU operators.

The reference drives its residual (EllipFunction), Jacobian action (the MATSHELL EllipShellMatMult) and
embedded-Q1 preconditioning matrix (EllipJacobian) with PETSc's SNES and KSP (src/fs/tests/ellip.c:654-747:
SNES rtol 1e-10, KSP and line search from the options database).  PETSc is not part of this build, so this
module restates the two roles that driver relies on, with PETSc's default semantics:

- `cg`: preconditioned conjugate gradients (KSPCG defaults: the preconditioned residual norm is monitored and
  tested);
- `gmres`: GMRES(restart) with left preconditioning from a zero initial guess (KSPGMRES defaults: classical
  Gram-Schmidt, restart 30; converged when the preconditioned residual norm falls to
  max(rtol * ||M^-1 b||, atol), diverged past dtol * ||M^-1 b||), returning the monitored norm history
  (what -ksp_monitor prints);
- `newton`: Newton's method with the "basic" (full step) or "bt" (backtracking, cubic) line search, PETSc's
  default convergence test (||F|| < atol, ||F|| <= rtol ||F_0||, or a step shorter than stol ||x||), returning
  the history of ||F|| (what -snes_monitor prints).

The vectors are device tensors; every operator application goes through the HIP library (residual, Jacobian
action, preconditioner).  The Krylov bookkeeping (dot products, the small Hessenberg least-squares problem) is
plumbing around them.
"""
import math

import numpy as np
import torch


def _norm(v):
    return float(torch.linalg.vector_norm(v))


def gmres(A, M, b, rtol=1e-5, atol=1e-50, dtol=1e5, restart=30, max_it=10000):
    """x ~ A^-1 b.  A(v, out) writes A v into out; M(v, out) writes the preconditioner M^-1 v (identity if
    None).  Returns (x, history of preconditioned residual norms, reason): reason > 0 converged (1 relative,
    2 absolute, 3 happy breakdown), < 0 diverged (-1 iteration limit, -2 dtol)."""
    n = b.numel()
    x = torch.zeros_like(b)
    t = torch.empty_like(b)
    w = torch.empty_like(b)

    def prec(v, out):
        if M is None:
            out.copy_(v)
        else:
            M(v, out)

    prec(b, w)  # r_0 = M^-1 (b - A 0)
    beta = _norm(w)
    hist = [beta]
    rnorm0 = beta
    ttol = max(rtol * rnorm0, atol)
    if beta <= ttol:
        return x, hist, 2 if beta <= atol else 1
    V = torch.empty((restart + 1, n), dtype=b.dtype, device=b.device)
    it = 0
    while True:
        V[0] = w / beta
        H = np.zeros((restart + 1, restart))
        g = np.zeros(restart + 1)
        g[0] = beta
        cs, sn = np.zeros(restart), np.zeros(restart)
        k = 0
        reason = 0
        for j in range(restart):
            A(V[j], t)
            prec(t, w)
            h = V[:j + 1] @ w  # classical Gram-Schmidt: every projection against the same vector
            w -= V[:j + 1].T @ h
            # the Hessenberg column and the new vector's norm come back to the host in one read
            col = torch.cat([h, torch.linalg.vector_norm(w).reshape(1)]).cpu().numpy().astype(np.float64)
            hn = float(col[j + 1])
            for i in range(j):  # previous rotations
                a, c = col[i], col[i + 1]
                col[i], col[i + 1] = cs[i] * a + sn[i] * c, -sn[i] * a + cs[i] * c
            den = math.hypot(col[j], col[j + 1])
            cs[j], sn[j] = (1.0, 0.0) if den == 0 else (col[j] / den, col[j + 1] / den)
            col[j], col[j + 1] = den, 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            H[:j + 2, j] = col
            k = j + 1
            it += 1
            rnorm = abs(g[j + 1])
            hist.append(rnorm)
            if rnorm <= ttol:
                reason = 2 if rnorm <= atol else 1
            elif hn == 0.0:
                reason = 3
            elif rnorm > dtol * rnorm0:
                reason = -2
            elif it >= max_it:
                reason = -1
            if reason:
                break
            V[j + 1] = w / hn
        y = np.linalg.solve(np.triu(H[:k, :k]), g[:k]) if k else np.zeros(0)
        x += V[:k].T @ torch.as_tensor(y, dtype=b.dtype, device=b.device)
        if reason:
            return x, hist, reason
        # restart from the true preconditioned residual
        A(x, t)
        prec(b - t, w)
        beta = _norm(w)
        hist.append(beta)
        if beta <= ttol:
            return x, hist, 2 if beta <= atol else 1


def cg(A, M, b, rtol=1e-5, atol=1e-50, dtol=1e5, max_it=10000):
    """Preconditioned conjugate gradients from x0 = 0 with KSPCG's defaults: the monitored and tested norm is
    the preconditioned residual ||M^-1 r||, converged at max(rtol * ||M^-1 b||, atol).  A and M as in gmres
    (both symmetric positive definite).  Returns (x, norm history, reason)."""
    x = torch.zeros_like(b)
    r = b.clone()
    z = torch.empty_like(b)
    q = torch.empty_like(b)
    if M is None:
        z.copy_(r)
    else:
        M(r, z)
    dp = _norm(z)
    hist = [dp]
    ttol = max(rtol * dp, atol)
    if dp <= ttol:
        return x, hist, 2 if dp <= atol else 1
    p = z.clone()
    rz = float(torch.dot(r, z))
    for it in range(1, max_it + 1):
        A(p, q)
        alpha = rz / float(torch.dot(p, q))
        x += alpha * p
        r -= alpha * q
        if M is None:
            z.copy_(r)
        else:
            M(r, z)
        dp = _norm(z)
        hist.append(dp)
        if dp <= ttol:
            return x, hist, 2 if dp <= atol else 1
        if dp > dtol * hist[0]:
            return x, hist, -2
        rz_new = float(torch.dot(r, z))
        p.mul_(rz_new / rz).add_(z)
        rz = rz_new
    return x, hist, -1


def _line_search_bt(op, x, f, fnorm, Y, t, alpha=1e-4, minlambda=1e-12):
    """PETSc's backtracking line search (order 3): the full step if ||F(x - Y)||^2 <= ||F||^2 + 2 alpha
    slope, else a quadratic then cubic model of ||F(x - lambda Y)||^2 / 2 along the step, with lambda kept
    in [0.1, 0.5] of the previous one.  slope = -|(F, J Y)|.  Returns (x_new, F(x_new), ||F||, lambda)."""
    op.mat_mult(Y, t)
    slope = float(torch.dot(f, t))
    slope = -abs(slope) if slope != 0.0 else -1.0
    lam = 1.0
    xt = x - lam * Y
    ft = torch.empty_like(f)
    op.function(xt, ft)
    gnorm = _norm(ft)
    if 0.5 * gnorm ** 2 <= 0.5 * fnorm ** 2 + lam * alpha * slope:
        return xt, ft, gnorm, lam
    lamtemp = -slope / (gnorm ** 2 - fnorm ** 2 - 2.0 * lam * slope)
    lamprev, gprev = lam, gnorm
    lamtemp = min(lamtemp, 0.5 * lam)
    lam = 0.1 * lam if lamtemp <= 0.1 * lam else lamtemp
    while lam > minlambda:
        xt = x - lam * Y
        op.function(xt, ft)
        gnorm = _norm(ft)
        if 0.5 * gnorm ** 2 <= 0.5 * fnorm ** 2 + lam * alpha * slope:
            return xt, ft, gnorm, lam
        t1 = 0.5 * (gnorm ** 2 - fnorm ** 2) - lam * slope
        t2 = 0.5 * (gprev ** 2 - fnorm ** 2) - lamprev * slope
        a = (t1 / lam ** 2 - t2 / lamprev ** 2) / (lam - lamprev)
        bq = (-lamprev * t1 / lam ** 2 + lam * t2 / lamprev ** 2) / (lam - lamprev)
        d = max(bq * bq - 3.0 * a * slope, 0.0)
        lamtemp = -slope / (2.0 * bq) if a == 0.0 else (-bq + math.sqrt(d)) / (3.0 * a)
        lamprev, gprev = lam, gnorm
        lamtemp = min(lamtemp, 0.5 * lam)
        lam = 0.1 * lam if lamtemp <= 0.1 * lam else lamtemp
    return xt, ft, gnorm, lamprev  # no sufficient decrease: the last trial (at lamprev) is kept


def newton(op, x, pc=None, rtol=1e-10, atol=1e-50, stol=1e-8, max_it=50, linesearch="basic", ksp_rtol=1e-5,
           ksp_atol=1e-50, ksp_restart=30, ksp_max_it=10000, ksp_type="gmres", mg_degree=2, mg_fp32=False):
    """Solve F(x) = 0 for an operator with op.function(x, f) (the residual, which also records x as the
    linearization state) and op.mat_mult(v, y) (the Jacobian action at that state).  x: the initial guess
    on the owned nodes (updated in place).  pc: None (unpreconditioned), "mg" (multigrid V-cycle on the
    embedded-Q1 matrix assembled at each iterate; op.assemble_q1; Chebyshev smoothing of degree mg_degree, applied
    in FP32 on FP16 coefficients with mg_fp32) or a callable pc(op, x) returning M(v, out) for the current iterate;
    ksp_type "gmres" or "cg".  Returns (x, ||F|| history, KSP histories, reason)."""
    f = torch.empty_like(x)
    t = torch.empty_like(x)
    op.function(x, f)
    fnorm = _norm(f)
    hist, ksp_hist = [fnorm], []
    ttol = rtol * fnorm
    if fnorm < atol:
        return x, hist, ksp_hist, 2
    q1, mg = None, None
    for it in range(1, max_it + 1):
        if pc is None:
            M = None
        elif pc == "mg":
            from dohp_amd.capi import MG
            q1 = op.assemble_q1(x, mat=q1)
            if mg is None:
                mg = MG(q1, smooth_degree=mg_degree)
            else:  # the same box: the hierarchy is recomputed in its own buffers
                mg.refresh(q1)
            if mg_fp32:
                r32, z32 = torch.empty(x.shape, dtype=torch.float32, device=x.device), \
                    torch.empty(x.shape, dtype=torch.float32, device=x.device)

                def M(v, out, mg=mg, r32=r32, z32=z32):
                    r32.copy_(v)
                    mg.apply_f32(r32, z32)
                    out.copy_(z32)
            else:
                M = mg.apply
        else:
            M = pc(op, x)
        if ksp_type == "cg":
            Y, kh, kreason = cg(op.mat_mult, M, f, rtol=ksp_rtol, atol=ksp_atol, max_it=ksp_max_it)
        else:
            Y, kh, kreason = gmres(op.mat_mult, M, f, rtol=ksp_rtol, atol=ksp_atol, restart=ksp_restart,
                                   max_it=ksp_max_it)
        ksp_hist.append(kh)
        if linesearch == "basic":
            x -= Y
            op.function(x, f)
            fnorm, lam = _norm(f), 1.0
        elif linesearch == "bt":  # (the last residual evaluated is the accepted point's: its linearization)
            xn, f, fnorm, lam = _line_search_bt(op, x, f, fnorm, Y, t)
            x.copy_(xn)
        else:
            raise ValueError("line search must be 'basic' or 'bt'")
        hist.append(fnorm)
        if fnorm != fnorm:
            return x, hist, ksp_hist, -4
        if fnorm < atol:
            return x, hist, ksp_hist, 2
        if fnorm <= ttol:
            return x, hist, ksp_hist, 3
        if lam * _norm(Y) < stol * _norm(x):
            return x, hist, ksp_hist, 4
    return x, hist, ksp_hist, -5

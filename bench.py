"""Benchmark: matrix-free order-6 operator apply (GDOF/s) on one MI355X.

Workload (default): 64^3 hexahedra on [0,1]^3 with interior vertices perturbed by +-10% of h
(seeded), order p=6 (n=7), homogeneous Dirichlet on all faces for the direction vector,
nonlinear diffusion -div(k(|grad u|) grad u) with k = (1e-2 + |grad u|^2)^(-1/4), linearized at
u0 = a seeded sum of 8 random-phase sine modes of amplitude 0.1.  ellip's p-Laplacian with
p = 1.5, eps^2 = 0.005 has eta = 2^(1/4) k, so the operator is ellip's up to that constant.

A step is one Jacobian-vector product y = J(u0) v (the MATSHELL MatMult a Krylov solver calls)
in FP64 with v ~ U[-1,1] (seeded), inputs resident in HBM.  value = N_int / t_step in GDOF/s.
As in the reference, the Jacobian reads per-point data that the linearization leaves behind (here
one symmetric 3x3 tensor per Gauss point, 48 bytes, formed by the first application after the
residual F(u0) and reused by every later one); the time to form it is reported beside the step
("quadrature_data_f64"), and the same applications recomputing everything from the state at every
call ("*_recomputed").  The residual F(u0), the FP32 Jacobian action and the linear (Poisson) apply
(recomputed by default, from stored data as "linear_*_stored") are timed too and reported under
"variants".  roofline.achieved divides the algorithmic bytes of one J v, B_Jv(8) of alg_bytes (x and the state
read, y written, element maps and vertices, geometry recomputed: no stored data), by the element kernel's own duration
(y += J v, timed with HIP events on its stream); roofline.stored_data_stream gives the same duration against the bytes
the stored-data form must stream.  roofline.traffic is the HBM traffic of one launch measured with rocprofv3
FETCH_SIZE/WRITE_SIZE, read from the profile summary committed under profiles/ when it covers this workload.
compute_roofline divides the reference algorithm's flops of one J v (alg_flops) by the same duration, against the FP64
peak; every variant carries both fractions (FP32 ones against the FP32 peak).

--workload solve: variable-coefficient Poisson -div(k grad u) = f with log-normal k (unit
variance, correlation length 0.1, seeded) on 48^3 elements, order 5; CG preconditioned by the
embedded-Q1 matrix (multigrid V-cycle on it with Jacobi-Chebyshev smoothing), plain FP64 versus
FP32-inner/FP64-outer defect correction and FP64 CG with the preconditioner in FP32, all to
relative residual 1e-10, with the embedded-Q1 matrix in the reference's arithmetic and corrected
(runs named corrected_q1_*); prints one JSON line with iterations, times and the final FP64 error
against the manufactured solution.  `value` is the reference-arithmetic mixed solve.

--workload learned: the same problem family on 48^3 elements, order 6.  The embedded-Q1 matrix (reference arithmetic)
is assembled with per-sub-cell derivative scales from a 3-layer MLP (11 sub-cell features -> 64 -> 64 -> 3, bfloat16
weights, FP32 accumulation) trained by the library's HIP kernels on residual/error pairs of plain solves of three
seeded 12^3 problems (selected on a held-out 24^3 one).  Timed to 1e-10: plain FP64 and FP64 CG with the FP32 V-cycle
on the plain matrix, then FP64, FP32 V-cycle and FP32-inner/FP64-outer on the learned matrix with the plain matrix
taking over after the plain solve's iteration count; the corrected matrix is reported beside them.  The default apply
line carries the same summary under "learned_solve" and the mixed-precision 48^3, p = 5 solve under "solve".

--workload batch: one 32^3 hexahedral mesh (perturbed) whose elements have degrees drawn uniformly
from 2..10 (seeded), each an independent element-local Poisson stiffness, plus 256 independent
Poisson problems (homogeneous Dirichlet) on n^3 meshes, n uniform in 4..16, degree uniform in 2..8,
all applied in one persistent launch (elements bucketed by degree, tile per degree autotuned).  A
step is one FP32 application of the whole batch; FP64 is timed too.  With the CPU leg the oracle
applies the same batch and both precisions state their error against it.

--workload newton: the default workload's p-Laplacian with a manufactured solution on 32^3 elements, order 6, solved
from u = 0 by Newton (backtracking) with GMRES preconditioned by the multigrid V-cycle on each iterate's corrected
embedded-Q1 matrix, to ||F|| <= 1e-10 ||F_0||; reports time, Newton and GMRES iterations and the nodal error.

--copy-probe: plain 1 GiB device copies (the calibration kernel for the HBM counters) and the streaming rates.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def sine_state(X, seed):
    from dohp_amd.mesh import sine_modes_state
    return sine_modes_state(X, seed)


def alg_bytes(s, n_tot, n_int, n, nelem, nvert, jv):
    """Algorithmic bytes of one operator application, geometry recomputed from the Q1 vertices (s bytes per
    value, FP64 geometry):
        B_lin(s) = s N_tot (read x) + s N_int (write y) + 4 n^3 E (int32 element -> node map)
                   + 24 N_vert + 32 E (FP64 vertex coordinates + int32 element -> vertex map)
        B_Jv(s)  = B_lin(s) + s N_tot (read the linearization state u)."""
    b = s * n_tot + s * n_int + 4 * n ** 3 * nelem + 24 * nvert + 32 * nelem
    return b + (s * n_tot if jv else 0)


def _contraction_flops(P, Q, D):
    """Flops of one tensor-product contraction of D components from P^3 to Q^3 values in three passes."""
    return 2 * D * (Q * P ** 3 + Q ** 2 * P ** 2 + Q ** 3 * P)


def alg_flops(n, kind):
    """Flops per element of one application in the reference's algorithm at n = p + 1 points per direction (FP64,
    geometry recomputed on every application): a gradient or its transpose is three n^3 -> n^3 contractions (6 n^4
    flops each), a value interpolation one; the geometry is the Q1 coordinates interpolated and differentiated
    (four 2^3 -> n^3 contractions of three components) plus the inverse, determinant and weight (44 per point); the
    push-forward and its transpose 18 per point each; the pointwise step 8 (linear), 20 (Jacobian) or 40 (residual)
    per point.
        F_lin = 36 n^4 + (18 + 8 + 18) n^3 + G,   F_Jv = 36 n^4 + (18 + 20 + 18) n^3 + G,
        F_res = 48 n^4 + (18 + 40 + 18) n^3 + G,  G = 4 C(2 -> n, 3) + 44 n^3.
    kind "jv_ref": the Jacobian action as the reference's iterator performs it, value and gradient of the direction
    and their transposes (eight n^3 -> n^3 contractions, 48 n^4) with the same geometry, push-forwards and pointwise step
    (172 kflop per element at p = 6; F_Jv, 143 kflop, keeps only what the action needs with lambda = 0)."""
    g = 4 * _contraction_flops(2, n, 3) + 44 * n ** 3
    if kind == "jv_ref":
        return 48 * n ** 4 + 56 * n ** 3 + g
    if kind == "lin":
        return 36 * n ** 4 + 44 * n ** 3 + g
    if kind == "jv":
        return 36 * n ** 4 + 56 * n ** 3 + g
    if kind == "res":
        return 48 * n ** 4 + 76 * n ** 3 + g
    raise ValueError(kind)


FP64_PEAK_TFLOPS, FP32_PEAK_TFLOPS = 78.6, 157.3


def stored_bytes(s, n_int, n, nelem, nq=6):
    """Bytes the stored-data form of an application streams: x read, y written, nq values per Gauss point."""
    return 2 * s * n_int + nq * s * n ** 3 * nelem


def triad_gbs():
    """STREAM triad a = b + s c over three 1 GiB FP64 arrays (24 bytes per entry), best of 10, GB/s."""
    import torch
    n = 1 << 27
    b = torch.ones(n, dtype=torch.float64, device="cuda")
    c = torch.ones(n, dtype=torch.float64, device="cuda")
    a = torch.empty_like(b)
    for _ in range(3):
        torch.add(b, c, alpha=3.0, out=a)
    torch.cuda.synchronize()
    best = 1e9
    for _ in range(10):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        torch.add(b, c, alpha=3.0, out=a)
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1) / 1e3)
    del a, b, c
    torch.cuda.empty_cache()
    return 24 * n / best / 1e9


def time_fn(torch, fn, reps, warm=2):
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        fn()
    t1.record()
    torch.cuda.synchronize()
    return t0.elapsed_time(t1) / 1e3 / reps


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(nthreads, reps=10):
    """Oracle J(u) v (the reference's per-element arithmetic in C, OpenMP over elements) on a bounded sample of
    the same workload, 20^3 elements at p = 6: median of `reps` applies after 2 warm-ups, on nthreads threads and
    on one."""
    from dohp_amd.mesh import block_vertices, splitmix64_uniform
    from oracle import oracle as O
    M = (20, 20, 20)
    V = block_vertices(M, box=(0, 1, 0, 1, 0, 1), perturb=0.1, seed=1)
    S = O.BlockSpace(M, 6, V)
    uc = sine_state(S.nodal_coordinates(), 3)
    v = np.concatenate([2 * splitmix64_uniform(9, S.nowned) - 1, np.zeros(S.nclosure - S.nowned)])
    out = {}
    for nt in (nthreads, 1):
        E = O.Ellip(S, p=1.5, eps=np.sqrt(0.005), nthreads=nt)
        ts = []
        for r in range(reps + 2):
            t0 = time.perf_counter()
            E.jacobian_apply(uc, v)
            if r >= 2:
                ts.append(time.perf_counter() - t0)
        out[nt] = S.nowned / float(np.median(ts)) / 1e9
    return out[nthreads], out[1], f"20^3 elements p=6 J(u)v (N_int {S.nowned}), median of {reps} applies after 2 warm-ups"


def measured_traffic(workload):
    """HBM bytes per launch of the element kernel, from the committed rocprofv3 counter summary."""
    path = os.path.join(ROOT, "profiles", "ellip_jv_f64_64cube_p6_hbm_traffic.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                key, val = line.split(":", 1)[1].rsplit("=", 1)
                if key.strip() == workload:
                    return int(float(val))
    except OSError:
        pass
    return None


def copy_probe():
    """Plain 1 GiB device copies (the calibration kernel of the HBM counters), timed, and the practical
    streaming rates of this GPU: copy (read + write) and a read-only reduction over 4 GiB."""
    import torch
    n = 1 << 27
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    t_copy = time_fn(torch, lambda: b.copy_(a), 10)
    big = torch.ones(4 * n, dtype=torch.float64, device="cuda")
    t_read = time_fn(torch, lambda: big.sum(), 10)
    print(json.dumps({"copy_probe_bytes_read": 8 * n, "copy_probe_bytes_written": 8 * n, "launches": 3,
                      "copy_GBs": round(16 * n / t_copy / 1e9, 1), "read_GBs": round(32 * n / t_read / 1e9, 1)}))


def lognormal_problem(C, torch, Me, p):
    """Log-normal Poisson problem of the mixed-precision solve: (fs, op, b, x_exact)."""
    from dohp_amd.fields import lognormal_field
    from dohp_amd.mesh import block_vertices
    M = (Me, Me, Me)
    V = block_vertices(M, box=(0, 1, 0, 1, 0, 1), perturb=0.1, seed=1)
    fs = C.FS(M, p, V)
    op = C.Ellip(fs, p=2.0, kvert=lognormal_field(V, seed=11, corr_len=0.1))
    X = fs.nodal_coordinates()[:fs.nowned]
    xs = torch.as_tensor(sine_state(X, 5) + X[:, 0] * (1 - X[:, 0]), dtype=torch.float64, device="cuda")
    b = torch.zeros_like(xs)
    op.mat_mult(xs, b)
    return fs, op, b, xs


def timed_solve(torch, op, q1, b, xs, **kw):  # noqa: E302
    """One warm-up solve (kernels, allocations, multigrid hierarchy), then one timed solve from zero: iterations,
    seconds, the FP64 relative residual computed apart and the error against the manufactured solution."""
    x = torch.zeros_like(b)
    op.solve(q1, b, x, rtol=1e-10, **kw)
    x.zero_()
    torch.cuda.synchronize()
    t = time.perf_counter()
    st = op.solve(q1, b, x, rtol=1e-10, **kw)
    torch.cuda.synchronize()
    t = time.perf_counter() - t
    r = torch.zeros_like(x)
    op.mat_mult(x, r)
    return {"seconds": round(t, 4), "outer_its": st["outer_its"], "inner_its": st["inner_its"],
            "fallback": st["fallback"], "relres": float(torch.linalg.norm(b - r) / torch.linalg.norm(b)),
            "error_vs_exact": float(torch.linalg.norm(x - xs) / torch.linalg.norm(xs))}


def solve_summary(C, torch, Me=48, p=5):
    """The mixed-precision solve (log-normal k, 48^3, p = 5) for the headline line: plain FP64 against the
    FP32-inner/FP64-outer defect correction and FP64 CG with the FP32 preconditioner, reference-arithmetic
    embedded-Q1 matrix, multigrid V-cycle."""
    fs, op, b, xs = lognormal_problem(C, torch, Me, p)
    q1 = op.assemble_q1(torch.zeros_like(xs))
    runs = {name: timed_solve(torch, op, q1, b, xs, mixed=m) for name, m in
            (("plain_f64", 0), ("mixed_f32_f64", 1), ("f64_cg_f32_precond", 2))}
    out = {"workload": f"log-normal k Poisson, {Me}^3 hex, p={p}, CG + multigrid on the embedded Q1, to 1e-10",
           "n_int": fs.nowned, "runs": runs,
           "speedup_mixed_vs_plain": round(runs["plain_f64"]["seconds"] / runs["mixed_f32_f64"]["seconds"], 3),
           "speedup_f32_precond_vs_plain": round(runs["plain_f64"]["seconds"] /
                                                 runs["f64_cg_f32_precond"]["seconds"], 3)}
    del fs, op, q1, b, xs
    torch.cuda.empty_cache()
    return out


def solve_workload(args):
    import torch
    import dohp_amd.capi as C
    from dohp_amd.fields import lognormal_field
    from dohp_amd.mesh import block_vertices
    Me, p = args.elems or 48, args.degree or 5
    M = (Me, Me, Me)
    t0 = time.perf_counter()
    V = block_vertices(M, box=(0, 1, 0, 1, 0, 1), perturb=0.1, seed=1)
    fs = C.FS(M, p, V)
    k = lognormal_field(V, seed=11, corr_len=0.1)
    op = C.Ellip(fs, p=2.0, kvert=k)
    X = fs.nodal_coordinates()[:fs.nowned]
    xs = torch.as_tensor(sine_state(X, 5) + X[:, 0] * (1 - X[:, 0]), dtype=torch.float64, device="cuda")
    del X
    b = torch.zeros_like(xs)
    op.mat_mult(xs, b)
    torch.cuda.synchronize()
    t_asm = time.perf_counter()
    q1 = op.assemble_q1(torch.zeros_like(xs))
    torch.cuda.synchronize()
    t_asm = time.perf_counter() - t_asm
    print(f"# solve setup {time.perf_counter() - t0:.1f}s (Q1 assembly {t_asm * 1e3:.1f} ms): {fs.nelem} elements "
          f"p={p}, N_int={fs.nowned}, nnz={q1.nnz}", file=sys.stderr, flush=True)
    op.set_q1_mode(1)
    q1c = op.assemble_q1(torch.zeros_like(xs))
    op.set_q1_mode(0)
    res = {}
    # default smoothing degree: 2 on the reference arithmetic's Q1 matrix, 1 on the corrected one; the runs named
    # *_mg1 / *_mg2 take the other degree
    for name, mixed, qq, mgd in (("plain_f64", 0, q1, 0), ("mixed_f32_f64", 1, q1, 0), ("f64_cg_f32_precond", 2, q1, 0),
                                 ("corrected_q1_plain_f64", 0, q1c, 0), ("corrected_q1_mixed_f32_f64", 1, q1c, 0),
                                 ("corrected_q1_f64_cg_f32_precond", 2, q1c, 0), ("f64_cg_f32_precond_mg1", 2, q1, 1),
                                 ("corrected_q1_f64_cg_f32_precond_mg2", 2, q1c, 2)):
        x = torch.zeros_like(xs)
        op.solve(qq, b, x, mixed=mixed, rtol=1e-10, mg_degree=mgd)  # warm-up (kernels, allocations)
        x.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        st = op.solve(qq, b, x, mixed=mixed, rtol=1e-10, mg_degree=mgd)
        torch.cuda.synchronize()
        t = time.perf_counter() - t
        r = torch.zeros_like(x)
        op.mat_mult(x, r)
        res[name] = {"seconds": round(t, 4), "outer_its": st["outer_its"], "inner_its": st["inner_its"],
                     "relres": float(torch.linalg.norm(b - r) / torch.linalg.norm(b)),
                     "error_vs_exact": float(torch.linalg.norm(x - xs) / torch.linalg.norm(xs))}
        print(f"# {name}: {res[name]}", file=sys.stderr, flush=True)
    best = min(("mixed_f32_f64", "f64_cg_f32_precond"), key=lambda k: res[k]["seconds"])
    out = {"metric": "solve time to relative residual 1e-10", "value": res["mixed_f32_f64"]["seconds"], "unit": "s",
           "n_gpus": 1, "higher_is_better": False, "dtype": "f32 inner / f64 outer", "data": "synthetic",
           "config": {"workload": f"log-normal k Poisson, {Me}^3 hex, p={p}, CG + multigrid on the embedded Q1 (Jacobi-Chebyshev smoothing)",
                      "n_int": fs.nowned, "q1_nnz": q1.nnz, "q1_assembly_ms": round(t_asm * 1e3, 2)},
           "speedup_mixed_vs_plain": round(res["plain_f64"]["seconds"] / res["mixed_f32_f64"]["seconds"], 3),
           "fastest_reduced_precision": best,
           "speedup_fastest_vs_plain": round(res["plain_f64"]["seconds"] / res[best]["seconds"], 3),
           "q1_modes": "runs named corrected_q1_* precondition with the corrected embedded-Q1 matrix (sub-cell "
                       "derivatives 1/h, inverse Jacobian at the point); the others with the reference's arithmetic",
           "runs": res}
    print(json.dumps(out))


LEARNED_TRAINING = dict(train_seeds=[101, 102, 103], val_seeds=[201], M=(12, 12, 12), epochs=25, lr=1e-2, npairs=8,
                        every=4, nrandom=2, val_M=(24, 24, 24))


def learned_summary(C, torch, Me=48, p=6, **overrides):
    """The learned-preconditioner solve (log-normal k, 48^3, p = 6): the sub-cell scale network of the embedded-Q1
    matrix is trained by the library's HIP kernels on residual/error pairs of plain solves of three 12^3 problems
    (selected on a held-out 24^3 one), then the 48^3 problem is solved to 1e-10 with the plain matrix (FP64, and FP64
    CG with the FP32 V-cycle) and with the learned one (FP64, FP32 V-cycle, FP32-inner/FP64-outer), the plain matrix
    taking over after as many iterations as the plain FP64 solve needed; the corrected matrix is shown beside them."""
    from dohp_amd import learned as Ld
    kw = dict(LEARNED_TRAINING, deg=p)
    kw.update(overrides)
    t0 = time.perf_counter()
    net, rep = Ld.fit_q1_scales(C, torch, **kw)
    t_train = time.perf_counter() - t0
    print(f"# learned training {t_train:.1f}s: " + json.dumps({k: v for k, v in rep.items() if k != "history"}),
          file=sys.stderr, flush=True)
    P = Ld.ScaledProblem(C, torch, (Me, Me, Me), p, 11)
    runs = {"plain_f64": timed_solve(torch, P.op, P.q1, P.b, P.xs),
            "plain_f64_cg_f32_precond": timed_solve(torch, P.op, P.q1, P.b, P.xs, mixed=2)}
    fb = runs["plain_f64"]["inner_its"]
    if net is not None:
        q1s = P.assemble(net)
        for name, m in (("learned_f64", 0), ("learned_f64_cg_f32_precond", 2), ("learned_mixed_f32_f64", 1)):
            runs[name] = timed_solve(torch, P.op, q1s, P.b, P.xs, mixed=m, fallback_q1=P.q1, fallback_its=fb)
    P.op.set_q1_mode(1)
    qc = P.op.assemble_q1(torch.zeros_like(P.b))
    runs["corrected_q1_f64_cg_f32_precond"] = timed_solve(torch, P.op, qc, P.b, P.xs, mixed=2)
    for k, v in runs.items():
        print(f"# {k}: {v}", file=sys.stderr, flush=True)
    learned = [k for k in runs if k.startswith("learned")]
    best = min(learned, key=lambda k: runs[k]["seconds"]) if learned else None
    out = {"workload": f"log-normal k Poisson, {Me}^3 hex, p={p}, CG + multigrid on the embedded Q1 (reference "
                       "arithmetic) with learned sub-cell scales, to 1e-10", "n_int": P.fs.nowned,
           "training": {"seconds": round(t_train, 2), **{k: v for k, v in rep.items() if k != "history"}},
           "runs": runs, "fastest_learned": best}
    if best:
        out["speedup_vs_plain_f64"] = round(runs["plain_f64"]["seconds"] / runs[best]["seconds"], 3)
        out["iterations_plain_vs_learned"] = [runs["plain_f64"]["inner_its"], runs[best]["inner_its"]]
    del P
    torch.cuda.empty_cache()
    return out


def learned_workload(args):
    import torch
    import dohp_amd.capi as C
    Me, p = args.elems or 48, args.degree or 6
    s = learned_summary(C, torch, Me, p)
    best = s.get("fastest_learned")
    r = s["runs"][best] if best else s["runs"]["plain_f64"]
    print(json.dumps({"metric": "solve time to relative residual 1e-10 (learned preconditioner)",
                      "value": r["seconds"], "unit": "s", "n_gpus": 1, "higher_is_better": False,
                      "dtype": "f32 / f64", "data": "synthetic", "config": {"workload": s["workload"],
                                                                            "n_int": s["n_int"]},
                      "reported_run": best or "plain_f64", **{k: v for k, v in s.items() if k not in ("workload",)}}))


def batch_problem(Me=32, nprob=256):
    from dohp_amd.mesh import mixed_order_batch
    return mixed_order_batch(Me, nprob)


def batch_workload(args):
    import torch
    import dohp_amd.capi as C
    from dohp_amd.mesh import splitmix64_uniform
    t0 = time.perf_counter()
    Me = args.elems or 32
    deg, vx, probs = batch_problem(Me)
    B = C.Batch()
    seg_mixed = B.add_elements(deg, vx)
    segs, nbytes = [], {4: 0, 8: 0}
    for (Mi, pi, Vi) in probs:
        fs = C.FS(Mi, pi, Vi)
        segs.append(B.add_space(fs))
        for s in (4, 8):
            nbytes[s] += alg_bytes(s, fs.nclosure, fs.nowned, pi + 1, fs.nelem, fs.nvert, False)
        del fs
    nexp = int(((deg + 1) ** 3).sum())
    for s in (4, 8):  # element-local: read x, write y, 8 vertex coordinates per element
        nbytes[s] += 2 * s * nexp + 24 * 8 * deg.size
    B.setup()
    n = B.size
    x = torch.as_tensor(2 * splitmix64_uniform(9, n) - 1, dtype=torch.float64, device="cuda")
    x32 = x.float()
    y, y32 = torch.zeros_like(x), torch.zeros_like(x32)
    B.mult(x, y)          # the first application of each precision tunes the tile per degree
    B.mult_f32(x32, y32)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    nel = deg.size + sum(m[0] ** 3 for (m, _, _) in probs)
    print(f"# setup {t_setup:.1f}s: {nel} elements, {n} packed unknowns ({nexp} element-local)",
          file=sys.stderr, flush=True)
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        B.mult_f32(x32, y32, stream)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        B.mult_f32(x32, y32, stream)
    e1.record(stream)
    torch.cuda.synchronize()
    t_step = e0.elapsed_time(e1) / 1e3 / args.steps
    t64 = time_fn(torch, lambda: B.mult(x, y, stream), max(3, args.steps // 2))
    B.mult(x, y, stream)
    B.mult_f32(x32, y32, stream)
    torch.cuda.synchronize()
    err32 = float(torch.linalg.norm(y32.double() - y) / torch.linalg.norm(y))
    print(f"# batch fp32 {t_step * 1e6:.1f} us, fp64 {t64 * 1e6:.1f} us, fp32 vs fp64 {err32:.2e}",
          file=sys.stderr, flush=True)
    names = ["workgroup", "wavefront", "mfma"]
    tuning = {}
    for d in range(1, 11):
        for fp32 in (0, 1):
            tile, ms = B.tuning(d, fp32)
            if tile >= 0:
                tuning[f"p{d}_{'f32' if fp32 else 'f64'}"] = {"tile": names[tile], "ms": ms}
    # FP32 with the MFMA contraction tile offered to the tuner from degree 4 up (BF16 operands, FP32
    # accumulation; 3 = operands split into two BF16 parts), and with it fixed for every degree >= 4
    mfma, ymf = {}, {}
    for passes in (3, 1):
        for forced in (False, True):
            B.set_mfma(passes)
            for d in range(4, 11):
                B.set_tile(d, True, 2 if forced else -1)
            ym = torch.zeros_like(y32)
            B.mult_f32(x32, ym, stream)
            tm = time_fn(torch, lambda: B.mult_f32(x32, ym, stream), max(3, args.steps // 2))
            B.mult_f32(x32, ym, stream)
            torch.cuda.synchronize()
            key = f"bf16x{passes}_{'all' if forced else 'tuned'}"
            ymf[key] = ym.double().cpu().numpy()
            mfma[key] = {"us": round(tm * 1e6, 1), "gdofs": round(n / tm / 1e9, 3),
                         "rel_err_vs_f64": float(torch.linalg.norm(ym.double() - y) / torch.linalg.norm(y)),
                         "tiles": {f"p{d}": names[B.tuning(d, True)[0]] for d in range(4, 11)
                                   if B.tuning(d, True)[0] >= 0}}
            print(f"# batch fp32 {key}: {mfma[key]}", file=sys.stderr, flush=True)
    B.set_mfma(0)
    for d in range(4, 11):
        B.set_tile(d, True, -1)
    cpu, oracle_err = None, None
    if not args.no_cpu:
        from oracle import oracle as O
        xh, yh, y32h = x.cpu().numpy(), y.cpu().numpy(), y32.double().cpu().numpy()
        ncores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        ref = np.zeros(n)
        tc = time.perf_counter()
        o, ln = B.segment(seg_mixed)
        ref[o:o + ln] = O.expanded_poisson_apply(deg, vx, xh[o:o + ln], nthreads=ncores)
        for s, (Mi, pi, Vi) in zip(segs, probs):
            o, ln = B.segment(s)
            S = O.BlockSpace(Mi, pi, Vi)
            E = O.Ellip(S, p=2.0, nthreads=ncores)
            ref[o:o + ln] = E.jacobian_apply(np.zeros(S.nclosure),
                                             np.concatenate([xh[o:o + ln], np.zeros(S.nclosure - S.nowned)]))[:ln]
        tc = time.perf_counter() - tc
        nr = np.linalg.norm(ref)
        oracle_err = {"f64": float(np.linalg.norm(yh - ref) / nr), "f32": float(np.linalg.norm(y32h - ref) / nr)}
        for key, ym in ymf.items():
            oracle_err[f"f32_{key}"] = float(np.linalg.norm(ym - ref) / nr)
        cpu = {"value": round(n / tc / 1e9, 4), "unit": "GDOF/s", "cores": ncores, "kind": "port",
               "sample": "the whole batch, one oracle application (element loop over OpenMP threads)"}
    achieved = nbytes[4] / t_step / 1e9
    out = {"metric": "GDOF/s batched mixed-order apply (FP32)", "value": round(n / t_step / 1e9, 3), "unit": "GDOF/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t_step * 1e3, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"{Me}^3 mesh of degree U{{2..10}} (element-local) + {len(probs)} Poisson problems "
                                  "n^3, n U{4..16}, degree U{2..8}, one persistent launch",
                      "elements": int(nel), "unknowns": int(n), "element_local_unknowns": nexp},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel_us": round(t_step * 1e6, 1),
                        "algorithmic_bytes": int(nbytes[4])},
           "cpu_baseline": cpu,
           "f64": {"us": round(t64 * 1e6, 1), "gdofs": round(n / t64 / 1e9, 3),
                   "hbm_frac": round(nbytes[8] / t64 / 1e9 / HBM_PEAK_GBS, 4)},
           "f32_rel_err_vs_f64": err32, "rel_err_vs_oracle": oracle_err, "tuning": tuning, "f32_mfma": mfma}
    print(json.dumps(out))


def newton_workload(args):
    """Nonlinear solve of the default workload's problem class with the library's Newton-GMRES layer: the
    p-Laplacian (p = 1.5, eps^2 = 0.005) with manufactured solution 0 (cos(x) e^y sin(z) type, Dirichlet data
    from it), perturbed mesh, from u = 0 with the backtracking line search; GMRES preconditioned by the
    multigrid V-cycle on the (corrected) embedded-Q1 matrix of each iterate (smoothing degree 1, applied in FP32 on
    FP16 coefficients); SNES rtol 1e-10, KSP rtol 1e-5."""
    import torch
    import dohp_amd.capi as C
    from dohp_amd import snes
    from dohp_amd.mesh import block_vertices
    Me, p = args.elems or 32, args.degree or 6
    M = (Me, Me, Me)
    V = block_vertices(M, box=(0, 1, 0, 1, 0, 1), perturb=0.1, seed=1)
    fs = C.FS(M, p, V)
    prm = dict(p=1.5, eps=float(np.sqrt(0.005)), lam=0.0, exact=0, a=1.2, b=0.9, c=1.1)
    op = C.Ellip(fs, **prm)
    op.set_q1_mode(1)
    X = fs.nodal_coordinates()
    # manufactured solution 0 of ellip (ellip.c:42-117): u = cos(a x) exp(b y) sin(c z)
    exact = np.cos(prm["a"] * X[:, 0]) * np.exp(prm["b"] * X[:, 1]) * np.sin(prm["c"] * X[:, 2])
    fs.dirichlet_commit(exact[fs.nowned:])
    runs = {}
    for rep in range(2):  # the first run also builds the Q1 pattern and tunes the kernels
        x = torch.zeros(fs.nowned, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        t = time.perf_counter()
        x, hist, ksp, reason = snes.newton(op, x, pc="mg", linesearch="bt", mg_degree=1, mg_fp32=True)
        torch.cuda.synchronize()
        t = time.perf_counter() - t
        err = float(np.abs(x.cpu().numpy() - exact[:fs.nowned]).max() / np.abs(exact[:fs.nowned]).max())
        runs[rep] = {"seconds": round(t, 3), "newton_its": len(hist) - 1, "gmres_its": sum(len(k) - 1 for k in ksp),
                     "reason": reason, "fnorm_rel": hist[-1] / hist[0], "max_nodal_error_rel": err}
        print(f"# newton run {rep}: {runs[rep]}", file=sys.stderr, flush=True)
    r = runs[1]
    print(json.dumps({"metric": "Newton-GMRES solve time to ||F|| <= 1e-10 ||F0||", "value": r["seconds"], "unit": "s",
                      "n_gpus": 1, "higher_is_better": False, "dtype": "f64", "data": "synthetic",
                      "config": {"workload": f"p-Laplacian p=1.5 (manufactured), {Me}^3 hex, p={p}, Newton + GMRES + "
                                             "FP32 multigrid on the corrected embedded Q1", "n_int": fs.nowned},
                      "runs": runs}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--elems", type=int, default=None)
    ap.add_argument("--degree", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true", help="leave the mixed-precision solve out of the apply line")
    ap.add_argument("--scatter", type=int, default=1, help="0 = eight-colour deterministic, 1 = atomics")
    ap.add_argument("--workload", choices=["apply", "solve", "learned", "batch", "newton"], default="apply")
    ap.add_argument("--copy-probe", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this operator has no multi-GPU path (its element-partitioned "
              "sharding is a design, not code).", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        return copy_probe()
    if args.workload == "solve":
        return solve_workload(args)
    if args.workload == "learned":
        return learned_workload(args)
    if args.workload == "newton":
        return newton_workload(args)
    if args.workload == "batch":
        return batch_workload(args)

    import torch
    import dohp_amd.capi as C
    from dohp_amd.mesh import block_vertices, element_vertex_index, splitmix64_uniform

    dev = torch.device("cuda:0")
    Me, p = args.elems or 64, args.degree or 6
    M = (Me, Me, Me)
    n = p + 1
    t_setup = time.perf_counter()
    V = block_vertices(M, box=(0, 1, 0, 1, 0, 1), perturb=0.1, seed=1)
    fs = C.FS(M, p, V)
    X = fs.nodal_coordinates()
    u0 = sine_state(X, 3)
    del X
    fs.dirichlet_commit(u0[fs.nowned:])
    op = C.Ellip(fs, p=1.5, eps=np.sqrt(0.005))
    op.set_scatter(args.scatter)
    lin = C.Ellip(fs, p=2.0)
    lin.set_scatter(args.scatter)
    u = torch.as_tensor(u0[:fs.nowned], dtype=torch.float64, device=dev)
    v = torch.as_tensor(2 * splitmix64_uniform(9, fs.nowned) - 1, dtype=torch.float64, device=dev)
    v32 = v.float()
    f = torch.zeros_like(u)
    y = torch.zeros_like(u)
    y32 = torch.zeros_like(v32)
    stream = torch.cuda.current_stream()
    op.function(u, f, stream)
    op.mat_mult(v, y, stream)  # the storage choice is timed and the quadrature data formed here (reported below)
    op.mat_mult_f32(v32, y32, stream)
    torch.cuda.synchronize()
    stored64, stored32 = op.storage(False)[0], op.storage(True)[0]
    t_setup = time.perf_counter() - t_setup
    print(f"# setup {t_setup:.1f}s: {fs.nelem} elements p={p}, N_int={fs.nowned}, N_tot={fs.nclosure}",
          file=sys.stderr, flush=True)

    # timed region: exactly `steps` FP64 Jacobian actions, synchronized on both sides
    for _ in range(args.warmup):
        op.mat_mult(v, y, stream)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        op.mat_mult(v, y, stream)
    e1.record(stream)
    torch.cuda.synchronize()
    t_step = e0.elapsed_time(e1) / 1e3 / args.steps
    print(f"# Jv fp64 {t_step * 1e6:.1f} us", file=sys.stderr, flush=True)
    # the element kernel alone (y += J v), same stream
    t_kernel = None
    if args.scatter == 1:
        t_kernel = time_fn(torch, lambda: op.mat_mult_add(v, y, stream), args.steps, warm=1)
        print(f"# Jv fp64 element kernel {t_kernel * 1e6:.1f} us", file=sys.stderr, flush=True)

    variants = {}
    reps = max(3, args.steps // 2)

    def timed(name, fn, s, jv, nq, kind=None):
        t = time_fn(torch, fn, reps)
        b = alg_bytes(s, fs.nclosure, fs.nowned, n, fs.nelem, fs.nvert, jv)
        kind = kind or ("jv" if jv else "lin")
        peak = FP64_PEAK_TFLOPS if s == 8 else FP32_PEAK_TFLOPS
        variants[name] = {"us": round(t * 1e6, 1), "gdofs": round(fs.nowned / t / 1e9, 2),
                          "hbm_frac": round(b / t / 1e9 / HBM_PEAK_GBS, 4),
                          "flop_frac": round(alg_flops(n, kind) * fs.nelem / t / 1e12 / peak, 4)}
        if nq:  # what the stored-data form streams, as a fraction of the same peak
            variants[name]["stored_stream_frac"] = round(stored_bytes(s, fs.nowned, n, fs.nelem, nq) / t / 1e9 /
                                                         HBM_PEAK_GBS, 4)
        print(f"# {name} {t * 1e6:.1f} us", file=sys.stderr, flush=True)
        return t

    # the same J v steps with the device NaN check's wait switched off (dohpEllipSetChecks 0, how the library's own
    # Krylov loops apply the operator): the applications queue back to back
    op.set_checks(0)
    t_async = time_fn(torch, lambda: op.mat_mult(v, y, stream), args.steps)
    op.set_checks(1)
    variants["jv_f64_unchecked"] = {"us": round(t_async * 1e6, 1), "gdofs": round(fs.nowned / t_async / 1e9, 2)}
    print(f"# Jv fp64 unchecked {t_async * 1e6:.1f} us", file=sys.stderr, flush=True)
    t_res = timed("residual_f64", lambda: op.function(u, f, stream), 8, True, 0, "res")
    # forming the quadrature data: the first application after a residual pays it
    t_pair = time_fn(torch, lambda: (op.function(u, f, stream), op.mat_mult(v, y, stream)), reps)
    variants["quadrature_data_f64"] = {"us": round((t_pair - t_res - t_step) * 1e6, 1),
                                       "bytes": 6 * 8 * n ** 3 * fs.nelem}
    print(f"# quadrature data {(t_pair - t_res - t_step) * 1e6:.1f} us", file=sys.stderr, flush=True)
    timed("jv_f32", lambda: op.mat_mult_f32(v32, y32, stream), 4, True, 6 if stored32 == 1 else 0)
    timed("linear_f64", lambda: lin.mat_mult(v, y, stream), 8, False, 0)
    timed("linear_f32", lambda: lin.mat_mult_f32(v32, y32, stream), 4, False, 0)
    # FP32 error against the FP64 apply (same inputs)
    op.mat_mult(v, y, stream)
    op.mat_mult_f32(v32, y32, stream)
    torch.cuda.synchronize()
    variants["jv_f32"]["rel_err_vs_f64"] = float(torch.linalg.norm(y32.double() - y) / torch.linalg.norm(y))
    tuning = {f"stored_{'f32' if f else 'f64'}": op.tuning("stored", f) for f in (0, 1)}
    tuning["residual_f64"] = op.tuning("residual", 0)
    tuning.update({f"linear_{'f32' if f else 'f64'}": lin.tuning("linear", f) for f in (0, 1)})
    # form chosen by the default storage mode (1 stored, 0 recomputed) and its timing [recomputed, stored] ms
    variants["storage_choice"] = {f"{name}_{'f32' if f else 'f64'}": o.storage(f)
                                  for name, o in (("jacobian", op), ("linear", lin)) for f in (0, 1)}
    # the linear operator from stored data (the default keeps the faster form: recomputing, at this order)
    lin.set_storage(1)
    timed("linear_f64_stored", lambda: lin.mat_mult(v, y, stream), 8, False, 6)
    timed("linear_f32_stored", lambda: lin.mat_mult_f32(v32, y32, stream), 4, False, 6)
    tuning.update({f"linear_stored_{'f32' if f else 'f64'}": lin.tuning("stored", f) for f in (0, 1)})
    lin.set_storage(2)
    # the Jacobian recomputing the state's gradient and the geometry at every application (no storage)
    ys = torch.zeros_like(y)
    op.mat_mult(v, ys, stream)
    op.set_storage(0)
    yr = torch.zeros_like(y)
    op.mat_mult(v, yr, stream)
    torch.cuda.synchronize()
    variants["stored_vs_recomputed_rel_diff"] = float(torch.linalg.norm(yr - ys) / torch.linalg.norm(ys))
    del ys, yr
    timed("jv_f64_recomputed", lambda: op.mat_mult(v, y, stream), 8, True, 0)
    timed("jv_f32_recomputed", lambda: op.mat_mult_f32(v32, y32, stream), 4, True, 0)
    tuning.update({f"jacobian_{'f32' if f else 'f64'}": op.tuning("jacobian", f) for f in (0, 1)})
    variants["tuning"] = tuning
    # the matrix-core shape (the one-dimensional contractions as 16x16x4 MFMA products in the kernel's
    # precision), pinned, against the shape the tuner kept: time and relative difference of the results (the
    # residual first: the stored entries then form their quadrature data again, outside the timed calls)
    mfma = {}
    y_ref = torch.zeros_like(y)
    lin.set_storage(0)
    for label, o, mode, fp32, storage, apply in (
            ("residual_f64", op, "residual", 0, 0, lambda o, out: o.function(u, out, stream)),
            ("jv_f64_stored", op, "stored", 0, 1, lambda o, out: o.mat_mult(v, out, stream)),
            ("jv_f32_stored", op, "stored", 1, 1, lambda o, out: o.mat_mult_f32(v32, out, stream)),
            ("jv_f64_recomputed", op, "jacobian", 0, 0, lambda o, out: o.mat_mult(v, out, stream)),
            ("jv_f32_recomputed", op, "jacobian", 1, 0, lambda o, out: o.mat_mult_f32(v32, out, stream)),
            ("linear_f64", lin, "linear", 0, 0, lambda o, out: o.mat_mult(v, out, stream)),
            ("linear_f32", lin, "linear", 1, 0, lambda o, out: o.mat_mult_f32(v32, out, stream))) \
            if 3 <= p <= 7 else ():  # (the shape's degrees)
        op.set_storage(storage)
        out = y if fp32 == 0 else y32
        ref = y_ref if fp32 == 0 else torch.zeros_like(y32)
        apply(o, ref)
        t_chosen = time_fn(torch, lambda: apply(o, ref), 4)
        chosen = o.tuning(mode, fp32)[0]
        o.set_shape(mode, fp32, 4)
        apply(o, out)
        t = time_fn(torch, lambda: apply(o, out), 4)
        torch.cuda.synchronize()
        o.set_shape(mode, fp32, chosen)
        mfma[label] = {"us": round(t * 1e6, 1), "tuned_shape_us": round(t_chosen * 1e6, 1),
                       "rel_diff": float(torch.linalg.norm((out - ref).double()) / torch.linalg.norm(ref.double()))}
    variants["mfma_shape"] = mfma or "the matrix-core shape covers degrees 3..7"
    lin.set_storage(2)
    del y_ref
    op.set_storage(2)
    # the same J(u0) v on a space built from the block's element maps (dFSCreateFromElements): the kernels read
    # the element-to-node and element-to-vertex maps (the 4 n^3 E bytes of the byte model) instead of
    # evaluating the block numbering
    fm = C.FS.from_elements(p, fs.element_nodes(), fs.nowned, fs.nclosure, V, element_vertex_index(M))
    fm.dirichlet_commit(u0[fs.nowned:])
    opm = C.Ellip(fm, p=1.5, eps=np.sqrt(0.005))
    opm.set_scatter(args.scatter)
    if stored64 in (0, 1):
        opm.set_storage(stored64)
    ym = torch.zeros_like(y)
    opm.function(u, f, stream)
    opm.mat_mult(v, ym, stream)
    op.function(u, f, stream)
    op.mat_mult(v, y, stream)
    t_m = time_fn(torch, lambda: opm.mat_mult(v, ym, stream), args.steps)
    torch.cuda.synchronize()
    B_m = alg_bytes(8, fs.nclosure, fs.nowned, n, fs.nelem, fs.nvert, True)
    variants["jv_f64_element_maps"] = {
        "us": round(t_m * 1e6, 1), "gdofs": round(fs.nowned / t_m / 1e9, 2),
        "hbm_frac": round(B_m / t_m / 1e9 / HBM_PEAK_GBS, 4),
        "rel_diff_vs_block": float(torch.linalg.norm(ym - y) / torch.linalg.norm(y))}
    print(f"# jv_f64 on element maps {t_m * 1e6:.1f} us", file=sys.stderr, flush=True)
    del opm, fm, ym

    B = alg_bytes(8, fs.nclosure, fs.nowned, n, fs.nelem, fs.nvert, True)
    t_k = t_kernel or t_step
    achieved = B / t_k / 1e9
    form = "from stored quadrature data, " if stored64 == 1 else ""
    workload = f"nonlinear diffusion J(u0)v {form}{Me}^3 hex, p={p}, FP64, perturbed mesh"
    triad = triad_gbs()
    traffic = measured_traffic(workload)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel_us": round(t_k * 1e6, 1),
                "algorithmic_bytes": int(B),
                "byte_model": "B_Jv(8) = 8 N_tot + 8 N_int + 4 n^3 E + 24 N_vert + 32 E + 8 N_tot (geometry "
                              "recomputed from the Q1 vertices)",
                "measured_triad_GBs": round(triad, 1), "frac_vs_measured_triad": round(achieved / triad, 4)}
    if traffic:
        roofline["traffic_over_algorithmic"] = round(traffic / B, 3)
    F = alg_flops(n, "jv") * fs.nelem
    F_ref = alg_flops(n, "jv_ref") * fs.nelem
    compute = {"bound": "fp64", "achieved": round(F / t_k / 1e12, 2), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
               "frac": round(F / t_k / 1e12 / FP64_PEAK_TFLOPS, 4), "algorithmic_flops": int(F),
               "frac_reference_count": round(F_ref / t_k / 1e12 / FP64_PEAK_TFLOPS, 4),
               "reference_count_flops": int(F_ref),
               "flop_model": "F_Jv = 36 n^4 + 56 n^3 + 4 C(2->n, 3) + 44 n^3 per element, C(P->Q, D) = "
                             "2 D (Q P^3 + Q^2 P^2 + Q^3 P): the reference's gradient, push-forward, pointwise "
                             "Jacobian, transposes and recomputed geometry"}
    if stored64 == 1:  # the stored form streams its per-point tensor: what it reaches on its own bytes
        sb = stored_bytes(8, fs.nowned, n, fs.nelem)
        roofline["stored_data_stream"] = {"bytes": int(sb), "GBs": round(sb / t_k / 1e9, 1),
                                          "frac": round(sb / t_k / 1e9 / HBM_PEAK_GBS, 4)}
    cpu = None
    if not args.no_cpu:
        ncores = os.cpu_count() or 1
        try:
            ncores = len(os.sched_getaffinity(0))
        except AttributeError:
            pass
        val, val1, sample = cpu_baseline(ncores)
        cpu = {"value": round(val, 4), "unit": "GDOF/s", "cores": ncores, "kind": "port", "sample": sample,
               "value_1core": round(val1, 5), "cpu_model": cpu_model(), "nproc": os.cpu_count()}
    out = {
        "metric": "GDOF/s operator apply at order 6",
        "value": round(fs.nowned / t_step / 1e9, 3),
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_step * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload, "elements": fs.nelem, "degree": p, "n_int": fs.nowned,
                   "n_tot": fs.nclosure, "scatter": "colored" if args.scatter == 0 else "atomic",
                   "jacobian_data": "stored per Gauss point (6 FP64 values)" if stored64 == 1 else "recomputed"},
        "roofline": roofline,
        "compute_roofline": compute,
        "cpu_baseline": cpu,
        "variants": variants,
    }
    if not args.no_solve:
        del op, lin, u, v, v32, f, y, y32, fs
        torch.cuda.empty_cache()
        out["solve"] = solve_summary(C, torch)
        print(f"# solve: {json.dumps(out['solve'])}", file=sys.stderr, flush=True)
        out["learned_solve"] = learned_summary(C, torch)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

// Krylov drivers above the operator: the role PETSc's KSP plays for ellip
// (src/fs/tests/ellip.c:701-747), on device vectors.
//
//   * PCG with the embedded-Q1 matrix as preconditioner, applied approximately either by a
//     Jacobi-preconditioned Chebyshev iteration of fixed degree (a fixed polynomial in
//     D^{-1} A_q1 times D^{-1}, hence symmetric, so CG stays valid) or by a multigrid V-cycle
//     on the Q1 matrix with Jacobi-Chebyshev smoothing (symmetric as well).
//   * Plain path: everything in FP64.
//   * Mixed path: FP64 outer defect correction x += A32^{-1}(b - A64 x) with the inner PCG in
//     FP32 (FP32 operator and Q1 applies); residuals of the outer loop and every Krylov dot
//     product accumulate in FP64.
//   * FP32-preconditioner path: FP64 PCG whose preconditioner (V-cycle or Chebyshev, learned
//     weights included) runs in FP32.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "common.h"

using namespace dohp;

extern "C" dErr dohpEllipGetQ1Context_Private(dohpEllip op, dohpFS *fs, dohpEllipParam *prm, const double **kvert,
                                              const double **state, int *q1mode);

namespace {

constexpr int kRedBlocks = 1024;

template <typename T>
__global__ void dot_partial_kernel(long n, const T *__restrict__ a, const T *__restrict__ b, double *__restrict__ part)
{
  __shared__ double sh[256];
  double s = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    s += (double)a[i] * (double)b[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void dot_final_kernel(int nb, const double *__restrict__ part, double *__restrict__ out)
{
  __shared__ double sh[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += part[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

// y = a*x + b*y
template <typename T>
__global__ void axpby_kernel(long n, T a, const T *__restrict__ x, T b, T *__restrict__ y)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}

// z = s * d * r  (diagonal scaling)
template <typename T>
__global__ void dscale_kernel(long n, T s, const T *__restrict__ dinv, const T *__restrict__ r, T *__restrict__ z)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = s * dinv[i] * r[i];
}

__global__ void inv_kernel(long n, const double *__restrict__ d, double *__restrict__ di, float *__restrict__ di32)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    di[i] = 1.0 / d[i];
    di32[i] = (float)di[i];
  }
}

template <typename T>
__global__ void fill_kernel(long n, T v, T *__restrict__ y)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = v;
}

// y32 = s * x64 ; x64 += s * y32
__global__ void demote_kernel(long n, double s, const double *__restrict__ x, float *__restrict__ y)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)(s * x[i]);
}
__global__ void promote_add_kernel(long n, double s, const float *__restrict__ y, double *__restrict__ x)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += s * (double)y[i];
}
__global__ void promote_kernel(long n, const float *__restrict__ y, double *__restrict__ x)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = (double)y[i];
}

inline dim3 grid(long n) { return dim3((n + 255) / 256); }

struct Ctx {
  hipStream_t st;
  long n;
  Pool *pool;  // the operator's: every work vector of a solve comes from it
  PoolBuf<double> part, res;
};

template <typename T>
dErr dot(Ctx &c, const T *a, const T *b, double *out)
{
  hipLaunchKernelGGL(dot_partial_kernel<T>, dim3(kRedBlocks), dim3(256), 0, c.st, c.n, a, b, c.part.ptr);
  hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(256), 0, c.st, kRedBlocks, c.part.ptr, c.res.ptr);
  DOHP_HIP(hipMemcpyAsync(out, c.res.ptr, sizeof(double), hipMemcpyDeviceToHost, c.st));
  DOHP_HIP(hipStreamSynchronize(c.st));
  return 0;
}

template <typename T>
dErr axpby(Ctx &c, T a, const T *x, T b, T *y)
{
  hipLaunchKernelGGL(axpby_kernel<T>, grid(c.n), dim3(256), 0, c.st, c.n, a, x, b, y);
  DOHP_HIP(hipGetLastError());
  return 0;
}

// Operators of one precision.
template <typename T>
struct Ops;
template <>
struct Ops<double> {
  static dErr A(dohpEllip op, const double *x, double *y, hipStream_t s) { return dohpEllipMatMult_Private(op, x, y, s); }
  static dErr Q(dohpMat m, const double *x, double *y, hipStream_t s) { return dohpMatMult(m, x, y, s); }
};
template <>
struct Ops<float> {
  static dErr A(dohpEllip op, const float *x, float *y, hipStream_t s) { return dohpEllipMatMult_Private(op, x, y, s); }
  static dErr Q(dohpMat m, const float *x, float *y, hipStream_t s) { return dohpMatMultF32(m, x, y, s); }
};

// Preconditioner: degree-k Chebyshev iteration on A_q1 z = r with Jacobi scaling, z0 = 0.
template <typename T>
struct Cheb {
  dohpMat q1;
  const T *dinv;
  double lmin, lmax;
  int degree;
  PoolBuf<T> res, d, w;
  dErr apply(Ctx &c, const T *r, T *z)
  {
    const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin), sigma = theta / delta;
    double rho = 1.0 / sigma;
    hipLaunchKernelGGL(dscale_kernel<T>, grid(c.n), dim3(256), 0, c.st, c.n, (T)(1.0 / theta), dinv, r, d.ptr);
    DOHP_HIP(hipMemcpyAsync(z, d.ptr, sizeof(T) * c.n, hipMemcpyDeviceToDevice, c.st));
    DOHP_HIP(hipMemcpyAsync(res.ptr, r, sizeof(T) * c.n, hipMemcpyDeviceToDevice, c.st));
    for (int k = 1; k < degree; k++) {
      DOHP_CALL(Ops<T>::Q(q1, d.ptr, w.ptr, c.st));
      DOHP_CALL(axpby<T>(c, T(-1), w.ptr, T(1), res.ptr));  // res = r - A z
      const double rho_new = 1.0 / (2.0 * sigma - rho);
      hipLaunchKernelGGL(dscale_kernel<T>, grid(c.n), dim3(256), 0, c.st, c.n, (T)(2.0 * rho_new / delta), dinv,
                         res.ptr, w.ptr);
      DOHP_CALL(axpby<T>(c, T(1), w.ptr, (T)(rho_new * rho), d.ptr));  // d = rho' rho d + 2 rho'/delta D^-1 res
      DOHP_CALL(axpby<T>(c, T(1), d.ptr, T(1), z));
      rho = rho_new;
    }
    return 0;
  }
};

__global__ void scale2_kernel(long n, const double *__restrict__ s, const void *__restrict__ x, void *__restrict__ y,
                              int f32)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (f32) ((float *)y)[i] = (float)s[i] * ((const float *)x)[i];
  else ((double *)y)[i] = s[i] * ((const double *)x)[i];
}

// Preconditioner of one precision: Jacobi-Chebyshev on the Q1 matrix, or a multigrid V-cycle on
// it, optionally corrected by learned node weights, z = S M(S r).  An FP64 preconditioner may hand
// the work to an FP32 one (low): r is rounded to FP32, z comes back exactly.
template <typename T>
struct Prec {
  Cheb<T> *cheb = nullptr;
  dohpMG mg = nullptr;
  dohpMG mg_fb = nullptr;  // plain multigrid taking over from mg (a learned embedded-Q1 matrix's) when learned_on drops
  const double *s = nullptr;  // learned node weights, or null
  T *tmp = nullptr;
  bool learned_on = false;
  Prec<float> *low = nullptr;
  float *lr = nullptr, *lz = nullptr;
  dErr apply(Ctx &c, const T *r, T *z);
};
template <typename T>
dErr mg_apply(dohpMG mg, const T *r, T *z, hipStream_t st);
template <>
dErr mg_apply<double>(dohpMG mg, const double *r, double *z, hipStream_t st) { return dohpMGApply(mg, r, z, st); }
template <>
dErr mg_apply<float>(dohpMG mg, const float *r, float *z, hipStream_t st) { return dohpMGApplyF32(mg, r, z, st); }
template <typename T>
dErr Prec<T>::apply(Ctx &c, const T *r, T *z)
{
  if constexpr (sizeof(T) == 8) {
    if (low) {
      hipLaunchKernelGGL(demote_kernel, grid(c.n), dim3(256), 0, c.st, c.n, 1.0, r, lr);
      DOHP_CALL(low->apply(c, lr, lz));
      hipLaunchKernelGGL(promote_kernel, grid(c.n), dim3(256), 0, c.st, c.n, lz, z);
      DOHP_HIP(hipGetLastError());
      return 0;
    }
  }
  if (!mg) return cheb->apply(c, r, z);
  if (mg_fb && !learned_on) return mg_apply<T>(mg_fb, r, z, c.st);
  if (!(s && learned_on)) return mg_apply<T>(mg, r, z, c.st);
  const int f32 = sizeof(T) == 4;
  hipLaunchKernelGGL(scale2_kernel, grid(c.n), dim3(256), 0, c.st, c.n, s, (const void *)r, (void *)tmp, f32);
  DOHP_CALL(mg_apply<T>(mg, tmp, z, c.st));
  hipLaunchKernelGGL(scale2_kernel, grid(c.n), dim3(256), 0, c.st, c.n, s, (const void *)z, (void *)z, f32);
  DOHP_HIP(hipGetLastError());
  return 0;
}

// Iteration budget shared by the sweeps of one solve: after `fallback_its` iterations with the
// learned correction, the plain preconditioner takes over.
struct Budget {
  int used = 0;
  int fallback_its = 0;
  int fell_back = 0;
};

// Largest eigenvalue of D^{-1} A_q1 by power iteration (FP64).
dErr estimate_lmax(Ctx &c, dohpMat q1, const double *dinv, int its, double *lmax)
{
  PoolBuf<double> v, w;
  DOHP_CALL(v.alloc(*c.pool, c.n));
  DOHP_CALL(w.alloc(*c.pool, c.n));
  std::vector<double> h(c.n);
  uint64_t s = 0x9E3779B97F4A7C15ULL;
  for (long i = 0; i < c.n; i++) {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    h[i] = (double)(s >> 11) / 9007199254740992.0 - 0.5;
  }
  DOHP_HIP(hipMemcpy(v.ptr, h.data(), sizeof(double) * c.n, hipMemcpyHostToDevice));
  double lam = 0, nv;
  DOHP_CALL(dot(c, v.ptr, v.ptr, &nv));
  DOHP_CALL(axpby<double>(c, 0.0, v.ptr, 1.0 / std::sqrt(nv), v.ptr));
  for (int k = 0; k < its; k++) {
    DOHP_CALL(dohpMatMult(q1, v.ptr, w.ptr, c.st));
    hipLaunchKernelGGL(dscale_kernel<double>, grid(c.n), dim3(256), 0, c.st, c.n, 1.0, dinv, w.ptr, w.ptr);
    double vw, ww;
    DOHP_CALL(dot(c, v.ptr, w.ptr, &vw));
    DOHP_CALL(dot(c, w.ptr, w.ptr, &ww));
    lam = vw;
    DOHP_CALL(axpby<double>(c, 1.0 / std::sqrt(ww), w.ptr, 0.0, v.ptr));
  }
  *lmax = lam;
  return 0;
}

// PCG on A x = b (x holds the initial guess).  Stops at ||r|| <= rtol ||b|| or max_it.  When the
// budget switches the learned correction off, CG restarts with the plain preconditioner from the
// current iterate.  snap (FP64 only): x after every `every`-th iteration, at most nsnap copies.
// flexible: beta in the Polak-Ribiere form z_{k+1}.(r_{k+1} - r_k) / z_k.r_k = -alpha z_{k+1}.q_k / z_k.r_k (one more
// dot product), equal to the usual z_{k+1}.r_{k+1} / z_k.r_k when the preconditioner is symmetric and exact, and
// robust when it is not: an FP32 V-cycle inside FP64 CG is symmetric only to FP32 rounding, which on the reference
// arithmetic's matrix at 48^3, p = 6 sometimes doubled the iteration count (91 -> 193).
template <typename T>
dErr pcg(Ctx &c, dohpEllip op, Prec<T> &M, Budget &bud, const T *b, T *x, double rtol, int max_it, int *its,
         double *rnorm, double *bnorm, T *snap = nullptr, int nsnap = 0, int every = 0, int *nsnapped = nullptr,
         bool flexible = false)
{
  PoolBuf<T> r, z, p, q;
  DOHP_CALL(r.alloc(*c.pool, c.n));
  DOHP_CALL(z.alloc(*c.pool, c.n));
  DOHP_CALL(p.alloc(*c.pool, c.n));
  DOHP_CALL(q.alloc(*c.pool, c.n));
  DOHP_CALL(Ops<T>::A(op, x, q.ptr, c.st));
  DOHP_HIP(hipMemcpyAsync(r.ptr, b, sizeof(T) * c.n, hipMemcpyDeviceToDevice, c.st));
  DOHP_CALL(axpby<T>(c, T(-1), q.ptr, T(1), r.ptr));
  double bb, rr;
  DOHP_CALL(dot(c, b, b, &bb));
  DOHP_CALL(dot(c, r.ptr, r.ptr, &rr));
  *bnorm = std::sqrt(bb);
  DOHP_CHECK(std::isfinite(bb) && std::isfinite(rr), dohpERR_NAN, "PCG: non-finite right-hand side or initial residual");
  const double tol = rtol * std::sqrt(bb);
  int k = 0, ns = 0;
  double rz = 0, alpha = 0;
  bool restart = false;
  auto precondition = [&]() -> dErr {
    bool &learned_on = M.low ? M.low->learned_on : M.learned_on;
    if (learned_on && bud.fallback_its > 0 && bud.used >= bud.fallback_its) {
      learned_on = false;
      bud.fell_back = 1;
      restart = true;
    }
    return M.apply(c, r.ptr, z.ptr);
  };
  if (std::sqrt(rr) > tol) {
    DOHP_CALL(precondition());
    restart = false;
    DOHP_HIP(hipMemcpyAsync(p.ptr, z.ptr, sizeof(T) * c.n, hipMemcpyDeviceToDevice, c.st));
    DOHP_CALL(dot(c, r.ptr, z.ptr, &rz));
  }
  while (std::sqrt(rr) > tol && k < max_it) {
    DOHP_CALL(Ops<T>::A(op, p.ptr, q.ptr, c.st));
    double pq;
    DOHP_CALL(dot(c, p.ptr, q.ptr, &pq));
    alpha = rz / pq;
    DOHP_CALL(axpby<T>(c, (T)alpha, p.ptr, T(1), x));
    DOHP_CALL(axpby<T>(c, (T)-alpha, q.ptr, T(1), r.ptr));
    DOHP_CALL(dot(c, r.ptr, r.ptr, &rr));
    DOHP_CHECK(std::isfinite(rr), dohpERR_NAN, "PCG: non-finite residual at iteration " + std::to_string(k + 1));
    k++;
    bud.used++;
    if (snap && every > 0 && k % every == 0 && ns < nsnap) {
      DOHP_HIP(hipMemcpyAsync(snap + (long)ns * c.n, x, sizeof(T) * c.n, hipMemcpyDeviceToDevice, c.st));
      ns++;
    }
    if (std::sqrt(rr) <= tol) break;
    DOHP_CALL(precondition());
    double rz_new;
    DOHP_CALL(dot(c, r.ptr, z.ptr, &rz_new));
    if (restart) {  // the preconditioner changed: new Krylov sequence from here
      DOHP_HIP(hipMemcpyAsync(p.ptr, z.ptr, sizeof(T) * c.n, hipMemcpyDeviceToDevice, c.st));
      restart = false;
    } else {
      double beta = rz_new / rz;
      if (flexible) {
        double zq;
        DOHP_CALL(dot(c, z.ptr, q.ptr, &zq));
        beta = -alpha * zq / rz;
      }
      DOHP_CALL(axpby<T>(c, T(1), z.ptr, (T)beta, p.ptr));
    }
    rz = rz_new;
  }
  *its = k;
  *rnorm = std::sqrt(rr);
  if (nsnapped) *nsnapped = ns;
  return 0;
}

}  // namespace

namespace {

// Multigrid hierarchy of q1 at smoothing degree `deg` (0 = the default for its arithmetic), kept with the matrix.
dErr hierarchy(dohpMat q1, int deg, hipStream_t st, dohpMG *mg)
{
  deg = deg > 0 ? deg : (q1->corrected || q1->scaled ? 1 : 2);
  if (q1->mg && q1->mg_degree == deg) {
    if (q1->mg_stale) DOHP_CALL(dohpMGRefresh(q1->mg, q1, st));
  } else {
    dohpMGDestroy(&q1->mg);
    DOHP_CALL(dohpMGCreate(q1, deg, 0, &q1->mg, st));
    q1->mg_degree = deg;
  }
  q1->mg_stale = false;
  *mg = q1->mg;
  return 0;
}

dErr solve_impl(dohpEllip op, dohpMat q1, dohpMLP mlp, const double *b, double *x, const dohpSolveOpts *opts,
                dohpSolveStats *stats, double *snap, int nsnap, int every, int *nsnapped, hipStream_t st,
                dohpMat q1fb = nullptr)
{
  DOHP_CHECK(op && q1 && b && x && opts && stats, dohpERR_ARG, "null argument");
  DOHP_CHECK(opts->pc >= 0 && opts->pc <= 1, dohpERR_ARG, "pc must be 0 (Jacobi-Chebyshev) or 1 (multigrid)");
  DOHP_CHECK(!mlp || opts->pc == 1, dohpERR_ARG, "the learned correction applies to the multigrid preconditioner");
  DOHP_CHECK(!q1fb || opts->pc == 1, dohpERR_ARG, "the fallback matrix applies to the multigrid preconditioner");
  const auto t0 = std::chrono::steady_clock::now();
  Ctx c;
  c.st = st;
  long nrows;
  DOHP_CALL(dohpMatGetSizes(q1, &nrows, nullptr));
  c.n = nrows;
  c.pool = dohpEllipPool_Private(op);
  DOHP_CALL(c.part.alloc(*c.pool, kRedBlocks));
  DOHP_CALL(c.res.alloc(*c.pool, 1));
  *stats = dohpSolveStats{};
  // Jacobi scaling and, for the one-level preconditioner, the Chebyshev interval
  // [lmax/ratio, 1.05 lmax] of D^{-1} A_q1
  PoolBuf<double> diag, dinv;
  PoolBuf<float> dinv32;
  DOHP_CALL(diag.alloc(*c.pool, c.n));
  DOHP_CALL(dinv.alloc(*c.pool, c.n));
  DOHP_CALL(dinv32.alloc(*c.pool, c.n));
  DOHP_CALL(dohpMatGetDiagonal(q1, diag.ptr, c.st));
  hipLaunchKernelGGL(inv_kernel, grid(c.n), dim3(256), 0, c.st, c.n, diag.ptr, dinv.ptr, dinv32.ptr);
  double lmax = 1.0;
  if (opts->pc == 0) {
    DOHP_CALL(estimate_lmax(c, q1, dinv.ptr, opts->eig_its > 0 ? opts->eig_its : 20, &lmax));
    lmax *= 1.05;
  }
  const double lmin = lmax / (opts->cheb_ratio > 1 ? opts->cheb_ratio : 30.0);
  stats->lmax = lmax;
  const int degree = opts->cheb_degree > 0 ? opts->cheb_degree : 4;
  struct {
    dohpMG mg = nullptr, mg_fb = nullptr;
  } mgh;
  if (opts->pc == 1) {  // the hierarchy is kept with the matrix until it is refilled
    // default smoothing degree: 1 on the corrected Q1 matrix and on one with learned sub-cell scales (fastest,
    // measured), 2 on the reference arithmetic's (where degree 1 with the FP32 V-cycle can stall); a fallback
    // matrix gets its own default
    DOHP_CALL(hierarchy(q1, opts->mg_degree, c.st, &mgh.mg));
    if (q1fb) DOHP_CALL(hierarchy(q1fb, opts->mg_degree, c.st, &mgh.mg_fb));
    DOHP_CHECK(opts->mg_coeffs == 0 || opts->mg_coeffs == 1, dohpERR_ARG, "mg_coeffs must be 0 (FP16) or 1 (FP32)");
    DOHP_CALL(dohpMGSetCoefficients(mgh.mg, opts->mg_coeffs == 0));
    if (mgh.mg_fb) DOHP_CALL(dohpMGSetCoefficients(mgh.mg_fb, opts->mg_coeffs == 0));
  }
  Budget bud;
  bud.fallback_its = mlp || q1fb ? opts->fallback_its : 0;
  // learned node weights (computed once: the features do not depend on the residual)
  PoolBuf<double> wts;
  PoolBuf<double> tmp64;
  PoolBuf<float> tmp32;
  if (mlp) {
    dohpFS fs;
    dohpEllipParam prm;
    const double *kv, *stt;
    DOHP_CALL(dohpEllipGetQ1Context_Private(op, &fs, &prm, &kv, &stt, nullptr));
    long np;
    DOHP_CALL(dohpMLPNumPatches(q1, &np));
    PoolBuf<float> X;
    DOHP_CALL(X.alloc(*c.pool, (size_t)(np > 0 ? np : 1) * dohpMLPNumInputs()));
    DOHP_CALL(wts.alloc(*c.pool, c.n));
    DOHP_CALL(dohpMLPFeatures(q1, fs, dinv.ptr, X.ptr, c.st));
    DOHP_CALL(dohpMLPWeights(mlp, q1, X.ptr, wts.ptr, nullptr, c.st));
    DOHP_CALL(tmp64.alloc(*c.pool, c.n));
    DOHP_CALL(tmp32.alloc(*c.pool, c.n));
    DOHP_HIP(hipStreamSynchronize(c.st));
  }
  const auto t_setup = std::chrono::steady_clock::now();
  stats->setup_seconds = std::chrono::duration<double>(t_setup - t0).count();
  if (opts->mixed == 2) {  // FP64 PCG, preconditioner in FP32
    Cheb<float> C32{q1, dinv32.ptr, lmin, lmax, degree};
    DOHP_CALL(C32.res.alloc(*c.pool, c.n));
    DOHP_CALL(C32.d.alloc(*c.pool, c.n));
    DOHP_CALL(C32.w.alloc(*c.pool, c.n));
    Prec<float> M32;
    M32.cheb = &C32;
    M32.mg = mgh.mg;
    M32.mg_fb = mgh.mg_fb;
    M32.s = wts.ptr;
    M32.tmp = tmp32.ptr;
    M32.learned_on = mlp != nullptr || q1fb != nullptr;
    PoolBuf<float> lr, lz;
    DOHP_CALL(lr.alloc(*c.pool, c.n));
    DOHP_CALL(lz.alloc(*c.pool, c.n));
    Prec<double> M;
    M.low = &M32;
    M.lr = lr.ptr;
    M.lz = lz.ptr;
    int its;
    double rn, bn;
    DOHP_CALL(pcg<double>(c, op, M, bud, b, x, opts->rtol, opts->max_it, &its, &rn, &bn, snap, nsnap, every, nsnapped,
                          true));
    stats->inner_its = its;
    stats->outer_its = 1;
    stats->rnorm = rn;
    stats->bnorm = bn;
  } else if (!opts->mixed) {
    Cheb<double> C64{q1, dinv.ptr, lmin, lmax, degree};
    DOHP_CALL(C64.res.alloc(*c.pool, c.n));
    DOHP_CALL(C64.d.alloc(*c.pool, c.n));
    DOHP_CALL(C64.w.alloc(*c.pool, c.n));
    Prec<double> M;
    M.cheb = &C64;
    M.mg = mgh.mg;
    M.mg_fb = mgh.mg_fb;
    M.s = wts.ptr;
    M.tmp = tmp64.ptr;
    M.learned_on = mlp != nullptr || q1fb != nullptr;
    int its;
    double rn, bn;
    DOHP_CALL(pcg<double>(c, op, M, bud, b, x, opts->rtol, opts->max_it, &its, &rn, &bn, snap, nsnap, every, nsnapped));
    stats->inner_its = its;
    stats->outer_its = 1;
    stats->rnorm = rn;
    stats->bnorm = bn;
  } else {
    Cheb<float> C32{q1, dinv32.ptr, lmin, lmax, degree};
    DOHP_CALL(C32.res.alloc(*c.pool, c.n));
    DOHP_CALL(C32.d.alloc(*c.pool, c.n));
    DOHP_CALL(C32.w.alloc(*c.pool, c.n));
    Prec<float> M;
    M.cheb = &C32;
    M.mg = mgh.mg;
    M.mg_fb = mgh.mg_fb;
    M.s = wts.ptr;
    M.tmp = tmp32.ptr;
    M.learned_on = mlp != nullptr || q1fb != nullptr;
    PoolBuf<double> r, t;
    PoolBuf<float> r32, e32;
    DOHP_CALL(r.alloc(*c.pool, c.n));
    DOHP_CALL(t.alloc(*c.pool, c.n));
    DOHP_CALL(r32.alloc(*c.pool, c.n));
    DOHP_CALL(e32.alloc(*c.pool, c.n));
    double bb;
    DOHP_CALL(dot(c, b, b, &bb));
    stats->bnorm = std::sqrt(bb);
    int outer = 0, inner_total = 0;
    for (;;) {
      // FP64 defect r = b - A64 x
      DOHP_CALL(dohpEllipMatMult_Private(op, x, t.ptr, c.st));
      DOHP_HIP(hipMemcpyAsync(r.ptr, b, sizeof(double) * c.n, hipMemcpyDeviceToDevice, c.st));
      DOHP_CALL(axpby<double>(c, -1.0, t.ptr, 1.0, r.ptr));
      double rr;
      DOHP_CALL(dot(c, r.ptr, r.ptr, &rr));
      stats->rnorm = std::sqrt(rr);
      DOHP_CHECK(std::isfinite(rr), dohpERR_NAN, "defect correction: non-finite FP64 residual");
      if (stats->rnorm <= opts->rtol * stats->bnorm || outer >= opts->max_outer) break;
      // FP32 correction, normalized so the inner problem is O(1)
      const double s = 1.0 / stats->rnorm;
      hipLaunchKernelGGL(demote_kernel, grid(c.n), dim3(256), 0, c.st, c.n, s, r.ptr, r32.ptr);
      hipLaunchKernelGGL(fill_kernel<float>, grid(c.n), dim3(256), 0, c.st, c.n, 0.0f, e32.ptr);
      int its;
      double rn, bn;
      // the last sweep needs only the reduction that is left to reach rtol ||b||
      const double need = 0.5 * opts->rtol * stats->bnorm / stats->rnorm;
      const double irtol = std::max(opts->inner_rtol, need);
      DOHP_CALL(pcg<float>(c, op, M, bud, r32.ptr, e32.ptr, irtol, opts->max_it, &its, &rn, &bn));
      inner_total += its;
      hipLaunchKernelGGL(promote_add_kernel, grid(c.n), dim3(256), 0, c.st, c.n, stats->rnorm, e32.ptr, x);
      outer++;
    }
    stats->outer_its = outer;
    stats->inner_its = inner_total;
  }
  stats->fallback = bud.fell_back;
  DOHP_HIP(hipStreamSynchronize(c.st));
  stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

}  // namespace

extern "C" dErr dohpEllipSolve(dohpEllip op, dohpMat q1, const double *b, double *x, const dohpSolveOpts *opts,
                               dohpSolveStats *stats, void *stream)
{
  return solve_impl(op, q1, nullptr, b, x, opts, stats, nullptr, 0, 0, nullptr, (hipStream_t)stream);
}

extern "C" dErr dohpEllipSolveLearned(dohpEllip op, dohpMat q1, dohpMLP mlp, const double *b, double *x,
                                      const dohpSolveOpts *opts, dohpSolveStats *stats, void *stream)
{
  DOHP_CHECK(mlp, dohpERR_ARG, "null network");
  return solve_impl(op, q1, mlp, b, x, opts, stats, nullptr, 0, 0, nullptr, (hipStream_t)stream);
}

extern "C" dErr dohpEllipSolveRecord(dohpEllip op, dohpMat q1, const double *b, double *x, const dohpSolveOpts *opts,
                                     dohpSolveStats *stats, double *snap_dev, int nsnap, int every, int *nsnapped,
                                     void *stream)
{
  DOHP_CHECK(opts && !opts->mixed, dohpERR_ARG, "iterates are recorded by the plain FP64 solve");
  DOHP_CHECK(snap_dev && nsnap > 0 && every > 0 && nsnapped, dohpERR_ARG, "bad snapshot arguments");
  return solve_impl(op, q1, nullptr, b, x, opts, stats, snap_dev, nsnap, every, nsnapped, (hipStream_t)stream);
}

extern "C" dErr dohpEllipSolveFallback(dohpEllip op, dohpMat q1, dohpMat q1_fallback, const double *b, double *x,
                                       const dohpSolveOpts *opts, dohpSolveStats *stats, void *stream)
{
  DOHP_CHECK(q1_fallback, dohpERR_ARG, "null fallback matrix");
  return solve_impl(op, q1, nullptr, b, x, opts, stats, nullptr, 0, 0, nullptr, (hipStream_t)stream, q1_fallback);
}

// Batched dEFSApply for isotropic tensor-product hexahedra (dEFSApply_Tensor_Hex,
// src/jacobi/impls/tensor/efstopo.c:368-423, with TensorMult_Hex_nounroll,
// src/jacobi/impls/tensor/inlinetmulthex.h:19-100, and the geometry mappings of
// efstopo.c:518-623).  One workgroup per element; each 1-D contraction is one pass over
// an LDS tile, components of a block size D > 1 are processed one at a time.
#include <algorithm>
#include <mutex>
#include <vector>

#include "common.h"

using namespace dohp;

namespace {

// dst (dims nout) = A applied along `dim` of src (dims nin).  A is [nout][nin] row-major
// (leading dimension lda) or, when trans, [nin][nout] read transposed.
__device__ void contract(const double *A, bool trans, int lda, int dim, const int nin[3], const int nout[3],
                         const double *src, double *dst)
{
  const int total = nout[0] * nout[1] * nout[2];
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    int o[3] = {t / (nout[1] * nout[2]), (t / nout[2]) % nout[1], t % nout[2]};
    const int od = o[dim];
    double s = 0.0;
    for (int i = 0; i < nin[dim]; i++) {
      o[dim] = i;
      const double a = trans ? A[i * lda + od] : A[od * lda + i];
      s += a * src[(o[0] * nin[1] + o[1]) * nin[2] + o[2]];
    }
    dst[t] = s;
  }
  __syncthreads();
}

// Three passes (dim 0, 1, 2) from an (a,a,a) tile to a (b,b,b) tile; result lands in w0.
__device__ void tensor3(const double *A0, const double *A1, const double *A2, bool trans, int lda, int a, int b,
                        double *w0, double *w1)
{
  const int s0[3] = {a, a, a}, s1[3] = {b, a, a}, s2[3] = {b, b, a}, s3[3] = {b, b, b};
  contract(A0, trans, lda, 0, s0, s1, w0, w1);
  contract(A1, trans, lda, 1, s1, s2, w1, w0);
  contract(A2, trans, lda, 2, s2, s3, w0, w1);
  const int n = b * b * b;
  for (int t = threadIdx.x; t < n; t += blockDim.x) w0[t] = w1[t];
  __syncthreads();
}

__global__ void efs_apply_kernel(int P, int Q, const double *__restrict__ tab, const double *__restrict__ jinv, int D,
                                 const double *__restrict__ in, double *__restrict__ out, int amode, int imode)
{
  extern __shared__ double sm[];
  const int M = P > Q ? P : Q, M3 = M * M * M;
  double *sB = sm, *sD = sm + Q * P, *w0 = sD + Q * P, *w1 = w0 + M3, *acc = w1 + M3;
  for (int t = threadIdx.x; t < 2 * Q * P; t += blockDim.x) sm[t] = tab[t];
  __syncthreads();
  const long e = blockIdx.x;
  const int P3 = P * P * P, Q3 = Q * Q * Q;
  const bool add = imode == dADD_VALUES;
  switch (amode) {
  case dAPPLY_INTERP:
  case dAPPLY_INTERP_TRANSPOSE: {
    const bool tr = amode == dAPPLY_INTERP_TRANSPOSE;
    const int a = tr ? Q : P, b = tr ? P : Q, a3 = a * a * a, b3 = b * b * b;
    const double *x = in + e * a3 * D;
    double *y = out + e * b3 * D;
    for (int d = 0; d < D; d++) {
      for (int t = threadIdx.x; t < a3; t += blockDim.x) w0[t] = x[t * D + d];
      __syncthreads();
      tensor3(sB, sB, sB, tr, P, a, b, w0, w1);
      for (int t = threadIdx.x; t < b3; t += blockDim.x) y[t * D + d] = add ? y[t * D + d] + w0[t] : w0[t];
      __syncthreads();
    }
  } break;
  case dAPPLY_GRAD: {
    const double *x = in + e * P3 * D;
    double *y = out + e * Q3 * D * 3;
    const double *J = jinv ? jinv + e * Q3 * 9 : nullptr;
    for (int d = 0; d < D; d++) {
      for (int c = 0; c < 3; c++) {  // df_c = (D (x) B (x) B) x, etc.: three full contractions
        for (int t = threadIdx.x; t < P3; t += blockDim.x) w0[t] = x[t * D + d];
        __syncthreads();
        tensor3(c == 0 ? sD : sB, c == 1 ? sD : sB, c == 2 ? sD : sB, false, P, P, Q, w0, w1);
        for (int q = threadIdx.x; q < Q3; q += blockDim.x) {
          const double df = w0[q];
          for (int ee = 0; ee < 3; ee++) {
            const double v = J ? df * J[q * 9 + c * 3 + ee] : (ee == c ? df : 0.0);
            double *dst = &y[(q * D + d) * 3 + ee];
            if (c == 0 && !add) *dst = v;
            else if (J || ee == c) *dst += v;
          }
        }
        __syncthreads();
      }
    }
  } break;
  case dAPPLY_GRAD_TRANSPOSE: {
    const double *x = in + e * Q3 * D * 3;
    double *y = out + e * P3 * D;
    const double *J = jinv + e * Q3 * 9;
    for (int d = 0; d < D; d++) {
      for (int t = threadIdx.x; t < P3; t += blockDim.x) acc[t] = add ? y[t * D + d] : 0.0;
      for (int c = 0; c < 3; c++) {  // df_c[q] = sum_e jinv[q][c][e] x[q][d][e], then (D^T|B^T)^3
        for (int q = threadIdx.x; q < Q3; q += blockDim.x) {
          const double *u = &x[(q * D + d) * 3];
          w0[q] = u[0] * J[q * 9 + c * 3 + 0] + u[1] * J[q * 9 + c * 3 + 1] + u[2] * J[q * 9 + c * 3 + 2];
        }
        __syncthreads();
        tensor3(c == 0 ? sD : sB, c == 1 ? sD : sB, c == 2 ? sD : sB, true, P, Q, P, w0, w1);
        for (int t = threadIdx.x; t < P3; t += blockDim.x) acc[t] += w0[t];
        __syncthreads();
      }
      for (int t = threadIdx.x; t < P3; t += blockDim.x) y[t * D + d] = acc[t];
      __syncthreads();
    }
  } break;
  }
}

// 1-D weights of a tensor rule, passed by value.
struct RuleW {
  double w[kMaxN];
};

// dRuleComputePhysical for every point of nelem elements (one thread per point): the cofactor inverse of
// dGeomInvert3 (include/dohpgeom.h:57-75, same operation order), jw = w_i w_j w_k det.  flag bit 0: a
// negative determinant (jacobi.c:397); bit 1: a non-finite one.
__global__ void rule_physical_kernel(int Q, RuleW rw, long npts, const double *__restrict__ jac,
                                     double *__restrict__ jinv, double *__restrict__ jw, int *__restrict__ flag)
{
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npts) return;
  const int Q3 = Q * Q * Q, q = (int)(t % Q3), i = q / (Q * Q), j = (q / Q) % Q, k = q % Q;
  double a[9];
#pragma unroll
  for (int r = 0; r < 9; r++) a[r] = jac[t * 9 + r];
  const double b0 = (a[1 * 3 + 1] * a[2 * 3 + 2] - a[2 * 3 + 1] * a[1 * 3 + 2]);
  const double b3 = -(a[1 * 3 + 0] * a[2 * 3 + 2] - a[2 * 3 + 0] * a[1 * 3 + 2]);
  const double b6 = (a[1 * 3 + 0] * a[2 * 3 + 1] - a[2 * 3 + 0] * a[1 * 3 + 1]);
  const double det = a[0] * b0 + a[1] * b3 + a[2] * b6;
  const double id = 1.0 / det;
  double *b = jinv + t * 9;
  b[0] = id * b0;
  b[1] = -id * (a[0 * 3 + 1] * a[2 * 3 + 2] - a[2 * 3 + 1] * a[0 * 3 + 2]);
  b[2] = id * (a[0 * 3 + 1] * a[1 * 3 + 2] - a[1 * 3 + 1] * a[0 * 3 + 2]);
  b[3] = id * b3;
  b[4] = id * (a[0 * 3 + 0] * a[2 * 3 + 2] - a[2 * 3 + 0] * a[0 * 3 + 2]);
  b[5] = -id * (a[0 * 3 + 0] * a[1 * 3 + 2] - a[1 * 3 + 0] * a[0 * 3 + 2]);
  b[6] = id * b6;
  b[7] = -id * (a[0 * 3 + 0] * a[2 * 3 + 1] - a[2 * 3 + 0] * a[0 * 3 + 1]);
  b[8] = id * (a[0 * 3 + 0] * a[1 * 3 + 1] - a[1 * 3 + 0] * a[0 * 3 + 1]);
  jw[t] = rw.w[i] * rw.w[j] * rw.w[k] * det;
  const int f = (det < 0.0 ? 1 : 0) | (__builtin_isfinite(det) ? 0 : 2);
  if (f) atomicOr(flag, f);
}

}  // namespace

extern "C" dErr dRuleComputePhysicalBatched(int Q, const double *weight, int nelem, const double *jac, double *jinv,
                                            double *jw, void *stream)
{
  DOHP_CHECK(Q >= 1 && Q <= kMaxN, dohpERR_ARG, "dRuleComputePhysicalBatched supports 1 <= Q <= 16");
  DOHP_CHECK(nelem >= 0, dohpERR_ARG, "negative element count");
  DOHP_CHECK(weight && jac && jinv && jw, dohpERR_ARG, "null argument");  // jacobi.c:389-390
  if (nelem == 0) return 0;
  RuleW rw{};
  for (int q = 0; q < Q; q++) rw.w[q] = weight[q];
  hipStream_t st = (hipStream_t)stream;
  DevBuf<int> flag;
  DOHP_CALL(flag.alloc(1));
  DOHP_HIP(hipMemsetAsync(flag.ptr, 0, sizeof(int), st));
  const long npts = (long)nelem * Q * Q * Q;
  hipLaunchKernelGGL(rule_physical_kernel, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, st, Q, rw, npts, jac,
                     jinv, jw, flag.ptr);
  DOHP_HIP(hipGetLastError());
  int h = 0;
  DOHP_HIP(hipMemcpyAsync(&h, flag.ptr, sizeof(int), hipMemcpyDeviceToHost, st));
  DOHP_HIP(hipStreamSynchronize(st));
  DOHP_CHECK(!(h & 1), dohpERR_NEGATIVE_JACOBIAN, "Negative Jacobian at a quadrature point");
  DOHP_CHECK(!(h & 2), dohpERR_NAN, "Non-finite Jacobian determinant at a quadrature point");
  return 0;
}

namespace {

constexpr int kEfsMaxP = 12;
constexpr int kEfsTab = 2 * kEfsMaxP * kEfsMaxP;  // B and D, [Q][P] each

dErr efs_check(int P, int Q, int nelem, const double *rcoord, const double *jinv, int D, const double *in,
               double *out, int amode, int imode)
{
  DOHP_CHECK(P >= 2 && P <= kEfsMaxP && Q >= 1 && Q <= kEfsMaxP, dohpERR_ARG,
             "dEFSApplyHexBatched supports 2<=P<=12, 1<=Q<=12");
  DOHP_CHECK(nelem >= 0, dohpERR_ARG, "negative element count");
  DOHP_CHECK(D >= 1, dohpERR_ARG, "dofs must be positive");  // jacobi.c:495
  DOHP_CHECK(imode == dINSERT_VALUES || imode == dADD_VALUES, dohpERR_ARG, "InsertMode not supported");
  DOHP_CHECK(amode >= dAPPLY_INTERP && amode <= dAPPLY_GRAD_TRANSPOSE, dohpERR_ARG, "invalid dApplyMode");
  DOHP_CHECK(amode != dAPPLY_GRAD_TRANSPOSE || jinv, dohpERR_ARG, "No Jinv, need a mapping");
  DOHP_CHECK(rcoord && in && out, dohpERR_ARG, "null argument");
  return 0;
}

dErr efs_launch(int P, int Q, const double *dtab, int nelem, const double *jinv, int D, const double *in, double *out,
                int amode, int imode, hipStream_t st)
{
  const int M = std::max(P, Q), M3 = M * M * M;
  const size_t shmem = sizeof(double) * (2 * Q * P + 3 * M3);
  hipLaunchKernelGGL(efs_apply_kernel, dim3(nelem), dim3(256), shmem, st, P, Q, dtab, jinv, D, in, out, amode,
                     imode);
  DOHP_HIP(hipGetLastError());
  return 0;
}

// The 1-D tables of the allocation-free entry point: one copy per GPU, written by a kernel that takes them by
// value (no copy from host memory), read by the apply kernel that follows on the same stream.
static __device__ __align__(256) double scratch[kEfsTab];

struct EfsTab {
  double v[kEfsTab];
};

__global__ void efs_store_tab_kernel(EfsTab t, int n, double *__restrict__ dst)
{
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = t.v[i];
}

// Device address of `scratch` on the current device, looked up once per device (the host-side name of a
// __device__ array is not its device address).
dErr efs_scratch(double **ptr)
{
  static std::mutex mu;
  static double *addr[64] = {};
  int dev = 0;
  DOHP_HIP(hipGetDevice(&dev));
  DOHP_CHECK(dev >= 0 && dev < 64, dohpERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void *p = nullptr;
    DOHP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    addr[dev] = (double *)p;
  }
  *ptr = addr[dev];
  return 0;
}

}  // namespace

extern "C" dErr dEFSApplyHexBatched(int P, int Q, const double *rcoord, int nelem, const double *jinv, int D,
                                    const double *in, double *out, int amode, int imode, void *stream)
{
  DOHP_CALL(efs_check(P, Q, nelem, rcoord, jinv, D, in, out, amode, imode));
  if (nelem == 0) return 0;
  std::vector<double> tab(2 * Q * P);
  DOHP_CALL(dohpTensorBasis(P, Q, rcoord, tab.data(), tab.data() + Q * P, nullptr, nullptr));
  DevBuf<double> dtab;
  DOHP_CALL(dtab.upload(tab.data(), tab.size()));
  hipStream_t st = (hipStream_t)stream;
  DOHP_CALL(efs_launch(P, Q, dtab.ptr, nelem, jinv, D, in, out, amode, imode, st));
  DOHP_HIP(hipStreamSynchronize(st));  // dtab is freed on return
  return 0;
}

extern "C" dErr dEFSApplyHexBatchedNoAlloc(int P, int Q, const double *rcoord, int nelem, const double *jinv, int D,
                                           const double *in, double *out, int amode, int imode, void *stream)
{
  DOHP_CALL(efs_check(P, Q, nelem, rcoord, jinv, D, in, out, amode, imode));
  DOHP_CHECK(2 * Q * P <= kEfsTab, dohpERR_ARG, "dEFSApplyHexBatchedNoAlloc: tables exceed the 2*12*12 staging area");
  if (nelem == 0) return 0;
  double *dtab = nullptr;
  DOHP_CALL(efs_scratch(&dtab));
  EfsTab t;  // host stack: B and D of the rule, passed to the device by value
  DOHP_CALL(dohpTensorBasis(P, Q, rcoord, t.v, t.v + Q * P, nullptr, nullptr));
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(efs_store_tab_kernel, dim3(1), dim3(256), 0, st, t, 2 * Q * P, dtab);
  DOHP_HIP(hipGetLastError());
  return efs_launch(P, Q, dtab, nelem, jinv, D, in, out, amode, imode, st);
}

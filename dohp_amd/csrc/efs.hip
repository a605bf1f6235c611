// Batched dEFSApply for isotropic tensor-product hexahedra (dEFSApply_Tensor_Hex,
// src/jacobi/impls/tensor/efstopo.c:368-423, with TensorMult_Hex_nounroll,
// src/jacobi/impls/tensor/inlinetmulthex.h:19-100, and the geometry mappings of
// efstopo.c:518-623).  One workgroup per element; each 1-D contraction is one pass over
// an LDS tile, components of a block size D > 1 are processed one at a time.
#include <algorithm>
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

}  // namespace

extern "C" dErr dEFSApplyHexBatched(int P, int Q, const double *rcoord, int nelem, const double *jinv, int D,
                                    const double *in, double *out, int amode, int imode, void *stream)
{
  DOHP_CHECK(P >= 2 && P <= 12 && Q >= 1 && Q <= 12, dohpERR_ARG, "dEFSApplyHexBatched supports 2<=P<=12, 1<=Q<=12");
  DOHP_CHECK(D >= 1, dohpERR_ARG, "dofs must be positive");  // jacobi.c:495
  DOHP_CHECK(imode == dINSERT_VALUES || imode == dADD_VALUES, dohpERR_ARG, "InsertMode not supported");
  DOHP_CHECK(amode >= dAPPLY_INTERP && amode <= dAPPLY_GRAD_TRANSPOSE, dohpERR_ARG, "invalid dApplyMode");
  DOHP_CHECK(amode != dAPPLY_GRAD_TRANSPOSE || jinv, dohpERR_ARG, "No Jinv, need a mapping");
  DOHP_CHECK(rcoord && in && out, dohpERR_ARG, "null argument");
  if (nelem == 0) return 0;
  std::vector<double> tab(2 * Q * P);
  DOHP_CALL(dohpTensorBasis(P, Q, rcoord, tab.data(), tab.data() + Q * P, nullptr, nullptr));
  DevBuf<double> dtab;
  DOHP_CALL(dtab.upload(tab.data(), tab.size()));
  const int M = std::max(P, Q), M3 = M * M * M;
  const size_t shmem = sizeof(double) * (2 * Q * P + 3 * M3);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(efs_apply_kernel, dim3(nelem), dim3(256), shmem, st, P, Q, dtab.ptr, jinv, D, in, out, amode,
                     imode);
  DOHP_HIP(hipGetLastError());
  DOHP_HIP(hipStreamSynchronize(st));  // dtab is freed on return
  return 0;
}

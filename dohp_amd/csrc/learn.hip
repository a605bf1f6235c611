// Learned correction of the embedded-Q1 preconditioner (no counterpart in the reference).
//
// The owned nodes of a block space form a box of the node lattice; it is tiled by disjoint
// 2x2x2 patches (nodes of an odd last layer are not covered).  A 3-layer perceptron
//     h1 = silu(W1 f + b1) (64),  h2 = silu(W2 h1 + b2) (64),  o = W3 h2 + b3 (8)
// maps the features f of a patch (32 inputs: log D_m - mean_m log D_m of the Q1 diagonal D, and
// each node's position inside its element, (lattice index mod p) / p per direction) to
// log-weights of its nodes, s_m = exp(clamp(o_m) / 2) (s = 1 off the patches).  The
// preconditioner's output for a residual r is corrected to
//     z = S M(S r),   S = diag(s),  M the multigrid V-cycle on the Q1 matrix,
// which stays linear, symmetric and positive definite, so plain CG remains valid.  Training
// minimizes the relative energy error of the preconditioned residual over residual/error pairs
// (r, e) of the operator J:
//     E = || e - a z ||_J^2 / || e ||_J^2
// with one scale a shared by a set of pairs (its optimum over the set; CG does not see a global
// factor, but the spread of M J's spectrum across the pairs' modes is what it converges on),
// whose gradient in s is  dE/ds = -2a [ (J d) .* M(S r) + r .* M(S J d) ] / ||e||_J^2, d = e - a z
// (M symmetric; a stationary for the set).  Weights are bfloat16 values (round to nearest even of an FP32
// master copy); every product accumulates in FP32.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "common.h"

using namespace dohp;

namespace {

constexpr int NH = 64;
// Parameter layout of a network with NIN_ inputs and NOUT_ outputs: [W1 NH x NIN | b1 | W2 NH x NH | b2 | W3 NOUT x NH | b3].
template <int NIN_, int NOUT_>
struct Shape {
  static constexpr int NIN = NIN_, NOUT = NOUT_;
  static constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NH * NIN, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
                       OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NOUT * NH, NPARAM = OFF_B3 + NOUT;
};
// Network kinds: 0 = node weights around the V-cycle (32 patch features -> 8 log-weights of the patch nodes);
// 1 = sub-cell derivative scales of the embedded-Q1 matrix (11 sub-cell features -> 3 log-scales).
using PatchNet = Shape<32, 8>;
using CellNet = Shape<kQ1CellInputs, 3>;
constexpr int NIN = PatchNet::NIN, NOUT = PatchNet::NOUT;
constexpr int NPARAM_MAX = PatchNet::NPARAM > CellNet::NPARAM ? PatchNet::NPARAM : CellNet::NPARAM;
inline int nparam_of(int kind) { return kind == 1 ? CellNet::NPARAM : PatchNet::NPARAM; }
inline int nin_of(int kind) { return kind == 1 ? CellNet::NIN : PatchNet::NIN; }
inline int nout_of(int kind) { return kind == 1 ? CellNet::NOUT : PatchNet::NOUT; }
constexpr float kClamp = 4.0f;  // |log weight| and |log scale| limit

struct PBox {
  int nx, ny, nz;  // node box
  __host__ __device__ int px() const { return nx / 2; }
  __host__ __device__ int py() const { return ny / 2; }
  __host__ __device__ int pz() const { return nz / 2; }
  __host__ __device__ long npatch() const { return (long)px() * py() * pz(); }
};

__device__ __forceinline__ void patch_coords(const PBox &b, long p, int &P, int &Q, int &R)
{
  const long pyz = (long)b.py() * b.pz();
  P = (int)(p / pyz);
  Q = (int)((p / b.pz()) % b.py());
  R = (int)(p % b.pz());
}

__device__ __forceinline__ long patch_node(const PBox &b, int P, int Q, int R, int m)
{
  const int I = 2 * P + (m >> 2), J = 2 * Q + ((m >> 1) & 1), K = 2 * R + (m & 1);
  return ((long)I * b.ny + J) * b.nz + K;
}

__device__ __forceinline__ float silu(float x) { return x / (1.0f + __expf(-x)); }
__device__ __forceinline__ float dsilu(float x)
{
  const float s = 1.0f / (1.0f + __expf(-x));
  return s * (1.0f + x * (1.0f - s));
}

// Weights are read as scalar operands (uniform over the wavefront) from the constant space.
typedef const __attribute__((address_space(4))) float CW;

template <class S>
__device__ __forceinline__ void mlp_forward(CW *w, const float (&f)[S::NIN], float (&z1)[NH], float (&z2)[NH],
                                            float (&o)[S::NOUT])
{
  float h1[NH];
#pragma unroll
  for (int i = 0; i < NH; i++) {
    float a = w[S::OFF_B1 + i];
#pragma unroll
    for (int j = 0; j < S::NIN; j++) a = fmaf(w[S::OFF_W1 + i * S::NIN + j], f[j], a);
    z1[i] = a;
    h1[i] = silu(a);
  }
  float h2[NH];
#pragma unroll
  for (int i = 0; i < NH; i++) {
    float a = w[S::OFF_B2 + i];
#pragma unroll
    for (int j = 0; j < NH; j++) a = fmaf(w[S::OFF_W2 + i * NH + j], h1[j], a);
    z2[i] = a;
    h2[i] = silu(a);
  }
#pragma unroll
  for (int i = 0; i < S::NOUT; i++) {
    float a = w[S::OFF_B3 + i];
#pragma unroll
    for (int j = 0; j < NH; j++) a = fmaf(w[S::OFF_W3 + i * NH + j], h2[j], a);
    o[i] = a;
  }
}

// Features of every patch (independent of the residual).  lo0..lo2: lattice index of the box's
// first node per direction; p: element degree.
__global__ void features_kernel(PBox b, int p, int lo0, int lo1, int lo2, const double *__restrict__ dinv,
                                float *__restrict__ X)
{
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= b.npatch()) return;
  int P, Q, R;
  patch_coords(b, q, P, Q, R);
  float ld[8], mean = 0.f;
  for (int m = 0; m < 8; m++) {
    ld[m] = -logf((float)dinv[patch_node(b, P, Q, R, m)]);
    mean += ld[m];
  }
  mean *= 0.125f;
  float *x = X + q * NIN;
  for (int m = 0; m < 8; m++) {
    x[m] = ld[m] - mean;
    const int I = lo0 + 2 * P + (m >> 2), J = lo1 + 2 * Q + ((m >> 1) & 1), K = lo2 + 2 * R + (m & 1);
    x[8 + 3 * m + 0] = (float)(I % p) / (float)p;
    x[8 + 3 * m + 1] = (float)(J % p) / (float)p;
    x[8 + 3 * m + 2] = (float)(K % p) / (float)p;
  }
}

// Forward over B samples, keeping the pre-activations for training (Z1, Z2 may be null).
template <class S>
__global__ void forward_kernel(int B, const float *__restrict__ wts, const float *__restrict__ X,
                               float *__restrict__ Z1, float *__restrict__ Z2, float *__restrict__ O)
{
  constexpr int NIN = S::NIN, NOUT = S::NOUT;
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B) return;
  float f[NIN], z1[NH], z2[NH], o[NOUT];
#pragma unroll
  for (int j = 0; j < NIN; j++) f[j] = X[s * NIN + j];
  mlp_forward<S>((CW *)wts, f, z1, z2, o);
  if (Z1) {
#pragma unroll
    for (int i = 0; i < NH; i++) {
      Z1[s * NH + i] = z1[i];
      Z2[s * NH + i] = z2[i];
    }
  }
#pragma unroll
  for (int i = 0; i < NOUT; i++) O[s * NOUT + i] = o[i];
}

// Node weights from the patch outputs: s = exp(clamp(o)/2) on patches, 1 elsewhere.
__global__ void weights_kernel(PBox b, const float *__restrict__ O, double *__restrict__ s64, float *__restrict__ s32)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = (long)b.nx * b.ny * b.nz;
  if (i >= n) return;
  const long yz = (long)b.ny * b.nz;
  const int I = (int)(i / yz), J = (int)((i / b.nz) % b.ny), K = (int)(i % b.nz);
  double w = 1.0;
  if (I < 2 * b.px() && J < 2 * b.py() && K < 2 * b.pz()) {
    const long q = ((long)(I >> 1) * b.py() + (J >> 1)) * b.pz() + (K >> 1);
    const int m = ((I & 1) << 2) | ((J & 1) << 1) | (K & 1);
    const float o = fminf(fmaxf(O[q * NOUT + m], -kClamp), kClamp);
    w = exp(0.5 * (double)o);
  }
  s64[i] = w;
  if (s32) s32[i] = (float)w;
}

// dE/do for the patch outputs from dE/ds at the nodes: ds/do = s/2 inside the clamp, 0 outside.
__global__ void dweights_kernel(PBox b, const float *__restrict__ O, const double *__restrict__ s,
                                const double *__restrict__ gs, float *__restrict__ dO)
{
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= b.npatch()) return;
  int P, Q, R;
  patch_coords(b, q, P, Q, R);
  for (int m = 0; m < 8; m++) {
    const long i = patch_node(b, P, Q, R, m);
    const float o = O[q * NOUT + m];
    dO[q * NOUT + m] = (o > -kClamp && o < kClamp) ? (float)(0.5 * s[i] * gs[i]) : 0.f;
  }
}

// Squared-error loss against targets (the gradient check of the training kernels).
template <class S>
__global__ void mse_grad_kernel(int B, const float *__restrict__ O, const float *__restrict__ Tg, float *__restrict__ dO,
                                float *__restrict__ sqerr)
{
  constexpr int NOUT = S::NOUT;
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B) return;
  float e2 = 0.f;
  for (int i = 0; i < NOUT; i++) {
    const float d = O[s * NOUT + i] - Tg[s * NOUT + i];
    e2 += d * d;
    dO[s * NOUT + i] = 2.0f * d / (float)(B * NOUT);
  }
  sqerr[s] = e2;
}

// Backward over B samples: dZ2, dZ1 from dO; Z1 and Z2 are overwritten by the activations silu(Z1), silu(Z2), which
// the parameter gradient reads.
template <class S>
__global__ void backward_kernel(int B, const float *__restrict__ wts, float *__restrict__ Z1, float *__restrict__ Z2,
                                const float *__restrict__ dO, float *__restrict__ dZ1, float *__restrict__ dZ2)
{
  constexpr int NOUT = S::NOUT, OFF_W3 = S::OFF_W3, OFF_W2 = S::OFF_W2;
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B) return;
  CW *w = (CW *)wts;
  float g[NOUT], d2[NH];
  for (int i = 0; i < NOUT; i++) g[i] = dO[s * NOUT + i];
  for (int j = 0; j < NH; j++) {
    float a = 0.f;
    for (int i = 0; i < NOUT; i++) a = fmaf(w[OFF_W3 + i * NH + j], g[i], a);
    const float z = Z2[s * NH + j];
    d2[j] = a * dsilu(z);
    dZ2[s * NH + j] = d2[j];
    Z2[s * NH + j] = silu(z);
  }
  for (int j = 0; j < NH; j++) {
    float a = 0.f;
    for (int i = 0; i < NH; i++) a = fmaf(w[OFF_W2 + i * NH + j], d2[i], a);
    const float z = Z1[s * NH + j];
    dZ1[s * NH + j] = a * dsilu(z);
    Z1[s * NH + j] = silu(z);
  }
}

// Parameter gradient in two deterministic stages: a (parameter block, sample chunk) grid where each thread sums one
// parameter's products over its chunk's samples in order into part[chunk][q], then grad_reduce_kernel sums the chunks
// in order.  A1, A2 are the activations left by backward_kernel.
constexpr int kGradChunk = 1024;  // samples per chunk
template <class S>
__global__ void __launch_bounds__(256) grad_kernel(int B, const float *__restrict__ X, const float *__restrict__ A1,
                                                   const float *__restrict__ A2, const float *__restrict__ dO,
                                                   const float *__restrict__ dZ1, const float *__restrict__ dZ2,
                                                   float *__restrict__ part)
{
  constexpr int NIN = S::NIN, NOUT = S::NOUT, NPARAM = S::NPARAM, OFF_B1 = S::OFF_B1, OFF_W2 = S::OFF_W2,
                OFF_B2 = S::OFF_B2, OFF_W3 = S::OFF_W3, OFF_B3 = S::OFF_B3;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= NPARAM) return;
  const long s0 = (long)blockIdx.y * kGradChunk, s1 = s0 + kGradChunk < B ? s0 + kGradChunk : B;
  float acc = 0.f;
  if (q < OFF_B1) {
    const int i = q / NIN, j = q % NIN;
    for (long s = s0; s < s1; s++) acc = fmaf(dZ1[s * NH + i], X[s * NIN + j], acc);
  } else if (q < OFF_W2) {
    const int i = q - OFF_B1;
    for (long s = s0; s < s1; s++) acc += dZ1[s * NH + i];
  } else if (q < OFF_B2) {
    const int i = (q - OFF_W2) / NH, j = (q - OFF_W2) % NH;
    for (long s = s0; s < s1; s++) acc = fmaf(dZ2[s * NH + i], A1[s * NH + j], acc);
  } else if (q < OFF_W3) {
    const int i = q - OFF_B2;
    for (long s = s0; s < s1; s++) acc += dZ2[s * NH + i];
  } else if (q < OFF_B3) {
    const int i = (q - OFF_W3) / NH, j = (q - OFF_W3) % NH;
    for (long s = s0; s < s1; s++) acc = fmaf(dO[s * NOUT + i], A2[s * NH + j], acc);
  } else {
    const int i = q - OFF_B3;
    for (long s = s0; s < s1; s++) acc += dO[s * NOUT + i];
  }
  part[(long)blockIdx.y * NPARAM + q] = acc;
}

__global__ void grad_reduce_kernel(int nparam, int nchunk, const float *__restrict__ part, float *__restrict__ grad)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nparam) return;
  float acc = 0.f;
  for (int c = 0; c < nchunk; c++) acc += part[(long)c * nparam + q];
  grad[q] = acc;
}

__device__ __forceinline__ float to_bf16_value(float x)
{
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return __uint_as_float(u & 0xffff0000u);
}

// Adam on the FP32 master weights; the network uses their bf16 values.
__global__ void adam_kernel(int nparam, int t, float lr, float b1, float b2, float eps, const float *__restrict__ grad,
                            float *__restrict__ master, float *__restrict__ m, float *__restrict__ v,
                            float *__restrict__ wts)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nparam) return;
  const float g = grad[q];
  const float mq = b1 * m[q] + (1.f - b1) * g, vq = b2 * v[q] + (1.f - b2) * g * g;
  m[q] = mq;
  v[q] = vq;
  const float mh = mq / (1.f - powf(b1, (float)t)), vh = vq / (1.f - powf(b2, (float)t));
  const float w = master[q] - lr * mh / (sqrtf(vh) + eps);
  master[q] = w;
  wts[q] = to_bf16_value(w);
}

// elementwise helpers of the energy gradient
__global__ void mul_kernel(long n, const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ c)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c[i] = a[i] * b[i];
}
__global__ void axpy_kernel(long n, double a, const double *__restrict__ x, const double *__restrict__ y,
                            double *__restrict__ out)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = y[i] + a * x[i];
}
// gs = c (Jd .* u + r .* v)
__global__ void energy_grad_kernel(long n, double c, const double *__restrict__ Jd, const double *__restrict__ u,
                                   const double *__restrict__ r, const double *__restrict__ v, double *__restrict__ gs)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) gs[i] = c * (Jd[i] * u[i] + r[i] * v[i]);
}
__global__ void dot_kernel(long n, const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ part)
{
  __shared__ double sh[256];
  double s = 0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) s += a[i] * b[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Sub-cell derivative scales sigma = exp(clamp(o)) from the cell network's outputs.
__global__ void cell_scales_kernel(long ncell, const float *__restrict__ O, float *__restrict__ sig)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 3 * ncell) sig[i] = __expf(fminf(fmaxf(O[i], -kClamp), kClamp));
}
// dL/do = dL/dsigma * sigma inside the clamp, 0 outside.
__global__ void dcell_scales_kernel(long ncell, const float *__restrict__ O, const float *__restrict__ gsig,
                                    float *__restrict__ dO)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * ncell) return;
  const float o = O[i];
  dO[i] = (o > -kClamp && o < kClamp) ? gsig[i] * __expf(o) : 0.f;
}

inline dim3 grid_of(long n, int bs = 256) { return dim3((unsigned)((n + bs - 1) / bs)); }

float host_bf16(float x)
{
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float y;
  std::memcpy(&y, &u, 4);
  return y;
}

}  // namespace

struct _p_dohpMLP {
  int kind = 0;  // 0: patch node weights, 1: embedded-Q1 sub-cell scales
  DevBuf<float> wts, master, m, v, grad;
  int step = 0;
  long B = 0;  // capacity of the per-sample buffers
  DevBuf<float> Z1, Z2, O, dO, dZ1, dZ2, sqerr;
  DevBuf<float> part;  // per-chunk partial parameter gradients
};

namespace {

dErr ensure(_p_dohpMLP *n, long B)
{
  if (n->B >= B) return 0;
  n->B = B;
  DOHP_CALL(n->Z1.alloc((size_t)B * NH));
  DOHP_CALL(n->Z2.alloc((size_t)B * NH));
  DOHP_CALL(n->dZ1.alloc((size_t)B * NH));
  DOHP_CALL(n->dZ2.alloc((size_t)B * NH));
  DOHP_CALL(n->O.alloc((size_t)B * nout_of(n->kind)));
  DOHP_CALL(n->dO.alloc((size_t)B * nout_of(n->kind)));
  DOHP_CALL(n->sqerr.alloc(B));
  DOHP_CALL(n->part.alloc((size_t)((B + kGradChunk - 1) / kGradChunk) * NPARAM_MAX));
  return 0;
}

template <class S>
dErr backward_and_grad_s(_p_dohpMLP *n, long B, const float *X, hipStream_t st)
{
  hipLaunchKernelGGL(backward_kernel<S>, grid_of(B, 128), dim3(128), 0, st, (int)B, n->wts.ptr, n->Z1.ptr, n->Z2.ptr,
                     n->dO.ptr, n->dZ1.ptr, n->dZ2.ptr);
  const int nchunk = (int)((B + kGradChunk - 1) / kGradChunk);
  DOHP_CHECK(nchunk <= 65535, dohpERR_ARG, "too many samples for one gradient");
  hipLaunchKernelGGL(grad_kernel<S>, dim3((S::NPARAM + 255) / 256, nchunk), dim3(256), 0, st, (int)B, X, n->Z1.ptr,
                     n->Z2.ptr, n->dO.ptr, n->dZ1.ptr, n->dZ2.ptr, n->part.ptr);
  hipLaunchKernelGGL(grad_reduce_kernel, grid_of(S::NPARAM), dim3(256), 0, st, S::NPARAM, nchunk, n->part.ptr,
                     n->grad.ptr);
  DOHP_HIP(hipGetLastError());
  return 0;
}
dErr backward_and_grad(_p_dohpMLP *n, long B, const float *X, hipStream_t st)
{
  return n->kind == 1 ? backward_and_grad_s<CellNet>(n, B, X, st) : backward_and_grad_s<PatchNet>(n, B, X, st);
}
dErr forward(_p_dohpMLP *n, long B, const float *X, hipStream_t st)
{
  if (n->kind == 1)
    hipLaunchKernelGGL(forward_kernel<CellNet>, grid_of(B, 128), dim3(128), 0, st, (int)B, n->wts.ptr, X, n->Z1.ptr,
                       n->Z2.ptr, n->O.ptr);
  else
    hipLaunchKernelGGL(forward_kernel<PatchNet>, grid_of(B, 128), dim3(128), 0, st, (int)B, n->wts.ptr, X, n->Z1.ptr,
                       n->Z2.ptr, n->O.ptr);
  DOHP_HIP(hipGetLastError());
  return 0;
}
dErr adam_step(_p_dohpMLP *n, float lr, hipStream_t st)
{
  n->step++;
  const int np = nparam_of(n->kind);
  hipLaunchKernelGGL(adam_kernel, grid_of(np), dim3(256), 0, st, np, n->step, lr, 0.9f, 0.999f, 1e-8f, n->grad.ptr,
                     n->master.ptr, n->m.ptr, n->v.ptr, n->wts.ptr);
  DOHP_HIP(hipGetLastError());
  return 0;
}

dErr box_of(dohpMat q1, PBox *b)
{
  DOHP_CHECK(q1 && q1->box[0] > 0, dohpERR_ARG, "learned correction needs the stencil form of a block-space Q1 matrix");
  *b = PBox{q1->box[0], q1->box[1], q1->box[2]};
  return 0;
}

dErr host_dot(const double *a, const double *b, long n, double *out, DevBuf<double> &part, hipStream_t st)
{
  hipLaunchKernelGGL(dot_kernel, dim3(256), dim3(256), 0, st, n, a, b, part.ptr);
  double h[256];
  DOHP_HIP(hipMemcpyAsync(h, part.ptr, sizeof(h), hipMemcpyDeviceToHost, st));
  DOHP_HIP(hipStreamSynchronize(st));
  double t = 0;
  for (int i = 0; i < 256; i++) t += h[i];
  *out = t;
  return 0;
}

// Energy error of the scaled preconditioner for (r, e) and, if gs is given, its gradient in the
// node weights s.  alpha > 0 fixes the scale a (a scale shared by many pairs), otherwise a is the
// pair's optimal one.  terms (optional) receives (z, r), (z, J z), (e, J e).
dErr energy_and_grad(dohpEllip op, dohpMG mg, long nn, const double *s, const double *r, const double *e,
                     double alpha, double *gs, double *energy, double *terms, hipStream_t st)
{
  DevBuf<double> t, u, z, d, Jd, v, part;
  for (DevBuf<double> *p : {&t, &u, &z, &d, &Jd, &v}) DOHP_CALL(p->alloc(nn));
  DOHP_CALL(part.alloc(256));
  hipLaunchKernelGGL(mul_kernel, grid_of(nn), dim3(256), 0, st, nn, s, r, t.ptr);      // t = S r
  DOHP_CALL(dohpMGApply(mg, t.ptr, u.ptr, st));                                        // u = M S r
  hipLaunchKernelGGL(mul_kernel, grid_of(nn), dim3(256), 0, st, nn, s, u.ptr, z.ptr);  // z = S u
  DOHP_CALL(dohpEllipMatMult_Private(op, z.ptr, Jd.ptr, st));                             // J z
  double zr, zJz, eJe;
  DOHP_CALL(host_dot(z.ptr, r, nn, &zr, part, st));
  DOHP_CALL(host_dot(z.ptr, Jd.ptr, nn, &zJz, part, st));
  DOHP_CALL(host_dot(e, r, nn, &eJe, part, st));  // (e, J e) = (e, r)
  DOHP_CHECK(zJz > 0 && eJe > 0, dohpERR_GENERIC, "energy error: nonpositive energy");
  if (terms) {
    terms[0] = zr;
    terms[1] = zJz;
    terms[2] = eJe;
  }
  const double a = alpha > 0 ? alpha : zr / zJz;
  hipLaunchKernelGGL(axpy_kernel, grid_of(nn), dim3(256), 0, st, nn, -a, z.ptr, e, d.ptr);  // d = e - a z
  DOHP_CALL(dohpEllipMatMult_Private(op, d.ptr, Jd.ptr, st));                                   // J d
  double dJd;
  DOHP_CALL(host_dot(d.ptr, Jd.ptr, nn, &dJd, part, st));
  if (energy) *energy = std::max(dJd, 0.0) / eJe;
  if (gs) {
    hipLaunchKernelGGL(mul_kernel, grid_of(nn), dim3(256), 0, st, nn, s, Jd.ptr, t.ptr);  // t = S J d
    DOHP_CALL(dohpMGApply(mg, t.ptr, v.ptr, st));                                         // v = M S J d
    hipLaunchKernelGGL(energy_grad_kernel, grid_of(nn), dim3(256), 0, st, nn, -2.0 * a / eJe, Jd.ptr, u.ptr, r, v.ptr,
                       gs);
  }
  DOHP_HIP(hipGetLastError());
  DOHP_HIP(hipStreamSynchronize(st));
  return 0;
}

}  // namespace

extern "C" int dohpMLPNumParams(void) { return PatchNet::NPARAM; }
extern "C" int dohpMLPNumInputs(void) { return PatchNet::NIN; }
extern "C" int dohpMLPNumParamsKind(int kind) { return kind == 0 || kind == 1 ? nparam_of(kind) : -1; }
extern "C" int dohpMLPNumInputsKind(int kind) { return kind == 0 || kind == 1 ? nin_of(kind) : -1; }
extern "C" int dohpMLPNumOutputsKind(int kind) { return kind == 0 || kind == 1 ? nout_of(kind) : -1; }

extern "C" dErr dohpMLPCreate(const float *params_host, dohpMLP *out) { return dohpMLPCreateKind(0, params_host, out); }

extern "C" dErr dohpMLPCreateKind(int kind, const float *params_host, dohpMLP *out)
{
  DOHP_CHECK(out, dohpERR_ARG, "null output");
  DOHP_CHECK(kind == 0 || kind == 1, dohpERR_ARG, "network kind must be 0 (patch node weights) or 1 (sub-cell scales)");
  *out = nullptr;
  _p_dohpMLP *n = new _p_dohpMLP;
  n->kind = kind;
  const int NPARAM = nparam_of(kind);
  std::vector<float> w(NPARAM, 0.0f), wb(NPARAM);
  if (params_host) std::memcpy(w.data(), params_host, sizeof(float) * NPARAM);
  for (int q = 0; q < NPARAM; q++) wb[q] = host_bf16(w[q]);
  std::vector<float> zero(NPARAM, 0.0f);
  dErr err = n->wts.upload(wb.data(), NPARAM);
  if (!err) err = n->master.upload(w.data(), NPARAM);
  if (!err) err = n->m.upload(zero.data(), NPARAM);
  if (!err) err = n->v.upload(zero.data(), NPARAM);
  if (!err) err = n->grad.alloc(NPARAM);
  if (err) {
    delete n;
    return err;
  }
  *out = n;
  return 0;
}

extern "C" dErr dohpMLPGetParams(dohpMLP n, float *master_host, float *bf16_values_host)
{
  DOHP_CHECK(n, dohpERR_ARG, "null network");
  const int NPARAM = nparam_of(n->kind);
  DOHP_HIP(hipDeviceSynchronize());
  if (master_host) DOHP_HIP(hipMemcpy(master_host, n->master.ptr, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
  if (bf16_values_host) DOHP_HIP(hipMemcpy(bf16_values_host, n->wts.ptr, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" dErr dohpMLPDestroy(dohpMLP *n)
{
  if (n && *n) {
    delete *n;
    *n = nullptr;
  }
  return 0;
}

extern "C" dErr dohpMLPNumPatches(dohpMat q1, long *npatch)
{
  PBox b;
  DOHP_CALL(box_of(q1, &b));
  *npatch = b.npatch();
  return 0;
}

extern "C" dErr dohpMLPFeatures(dohpMat q1, dohpFS fs, const double *dinv, float *X_dev, void *stream)
{
  PBox b;
  DOHP_CALL(box_of(q1, &b));
  DOHP_CHECK(fs && dinv && X_dev, dohpERR_ARG, "null argument");
  const int m = fs->dirichlet_mask;
  const int lo0 = (m & 1) ? 1 : 0, lo1 = lo0, lo2 = (m & 4) ? 1 : 0;
  const long np = b.npatch();
  if (np > 0)
    hipLaunchKernelGGL(features_kernel, grid_of(np), dim3(256), 0, (hipStream_t)stream, b, fs->p, lo0, lo1, lo2, dinv,
                       X_dev);
  DOHP_HIP(hipGetLastError());
  return 0;
}

extern "C" dErr dohpMLPWeights(dohpMLP n, dohpMat q1, const float *X_dev, double *s_dev, float *s32_dev, void *stream)
{
  PBox b;
  DOHP_CALL(box_of(q1, &b));
  DOHP_CHECK(n && X_dev && s_dev, dohpERR_ARG, "null argument");
  DOHP_CHECK(n->kind == 0, dohpERR_ARG, "node weights need a patch network (kind 0)");
  hipStream_t st = (hipStream_t)stream;
  const long np = b.npatch(), nn = (long)b.nx * b.ny * b.nz;
  DOHP_CALL(ensure(n, np > 0 ? np : 1));
  if (np > 0) DOHP_CALL(forward(n, np, X_dev, st));
  hipLaunchKernelGGL(weights_kernel, grid_of(nn), dim3(256), 0, st, b, n->O.ptr, s_dev, s32_dev);
  DOHP_HIP(hipGetLastError());
  return 0;
}

extern "C" dErr dohpMLPGradient(dohpMLP n, int B, const float *X_dev, const float *T_dev, float *grad_host,
                                double *loss, void *stream)
{
  DOHP_CHECK(n && B > 0 && X_dev && T_dev, dohpERR_ARG, "bad argument");
  hipStream_t st = (hipStream_t)stream;
  DOHP_CALL(ensure(n, B));
  DOHP_CALL(forward(n, B, X_dev, st));
  if (n->kind == 1)
    hipLaunchKernelGGL(mse_grad_kernel<CellNet>, grid_of(B, 128), dim3(128), 0, st, B, n->O.ptr, T_dev, n->dO.ptr,
                       n->sqerr.ptr);
  else
    hipLaunchKernelGGL(mse_grad_kernel<PatchNet>, grid_of(B, 128), dim3(128), 0, st, B, n->O.ptr, T_dev, n->dO.ptr,
                       n->sqerr.ptr);
  DOHP_CALL(backward_and_grad(n, B, X_dev, st));
  std::vector<float> e(B);
  DOHP_HIP(hipMemcpyAsync(e.data(), n->sqerr.ptr, sizeof(float) * B, hipMemcpyDeviceToHost, st));
  if (grad_host)
    DOHP_HIP(hipMemcpyAsync(grad_host, n->grad.ptr, sizeof(float) * nparam_of(n->kind), hipMemcpyDeviceToHost, st));
  DOHP_HIP(hipStreamSynchronize(st));
  if (loss) {
    double t = 0;
    for (int s = 0; s < B; s++) t += e[s];
    *loss = t / ((double)B * nout_of(n->kind));
  }
  return 0;
}

extern "C" dErr dohpMLPEnergyWeightGradient(dohpEllip op, dohpMG mg, long nn, const double *s, const double *r,
                                            const double *e, double alpha, double *gs_dev, double *energy,
                                            double *terms, void *stream)
{
  DOHP_CHECK(op && mg && s && r && e, dohpERR_ARG, "null argument");
  return energy_and_grad(op, mg, nn, s, r, e, alpha, gs_dev, energy, terms, (hipStream_t)stream);
}

// One Adam step given dL/ds of some loss L at the node weights of the current parameters.
extern "C" dErr dohpMLPWeightStep(dohpMLP n, dohpMat q1, const float *X_dev, const double *gs, float lr, void *stream)
{
  PBox b;
  DOHP_CALL(box_of(q1, &b));
  DOHP_CHECK(n && X_dev && gs, dohpERR_ARG, "null argument");
  DOHP_CHECK(n->kind == 0, dohpERR_ARG, "node-weight step needs a patch network (kind 0)");
  hipStream_t st = (hipStream_t)stream;
  const long nn = (long)b.nx * b.ny * b.nz, np = b.npatch();
  DOHP_CHECK(np > 0, dohpERR_ARG, "no patches");
  DevBuf<double> s;
  DOHP_CALL(s.alloc(nn));
  DOHP_CALL(dohpMLPWeights(n, q1, X_dev, s.ptr, nullptr, st));
  hipLaunchKernelGGL(dweights_kernel, grid_of(np), dim3(256), 0, st, b, n->O.ptr, s.ptr, gs, n->dO.ptr);
  DOHP_CALL(backward_and_grad(n, np, X_dev, st));
  DOHP_CALL(adam_step(n, lr, st));
  DOHP_HIP(hipStreamSynchronize(st));
  return 0;
}

extern "C" dErr dohpMLPEnergyStep(dohpMLP n, dohpEllip op, dohpMat q1, dohpMG mg, const float *X_dev, const double *r,
                                  const double *e, double alpha, float lr, double *energy, void *stream)
{
  PBox b;
  DOHP_CALL(box_of(q1, &b));
  DOHP_CHECK(n && op && mg && X_dev && r && e, dohpERR_ARG, "null argument");
  DOHP_CHECK(n->kind == 0, dohpERR_ARG, "node-weight step needs a patch network (kind 0)");
  hipStream_t st = (hipStream_t)stream;
  const long nn = (long)b.nx * b.ny * b.nz, np = b.npatch();
  DOHP_CHECK(np > 0, dohpERR_ARG, "no patches");
  DevBuf<double> s, gs;
  DOHP_CALL(s.alloc(nn));
  DOHP_CALL(gs.alloc(nn));
  DOHP_CALL(dohpMLPWeights(n, q1, X_dev, s.ptr, nullptr, st));
  DOHP_CALL(energy_and_grad(op, mg, nn, s.ptr, r, e, alpha, gs.ptr, energy, nullptr, st));
  hipLaunchKernelGGL(dweights_kernel, grid_of(np), dim3(256), 0, st, b, n->O.ptr, s.ptr, gs.ptr, n->dO.ptr);
  DOHP_CALL(backward_and_grad(n, np, X_dev, st));
  DOHP_CALL(adam_step(n, lr, st));
  DOHP_HIP(hipStreamSynchronize(st));
  return 0;
}

extern "C" dErr dohpMLPCellScales(dohpMLP n, long ncell, const float *X_dev, float *scales_dev, void *stream)
{
  DOHP_CHECK(n && X_dev && scales_dev && ncell > 0, dohpERR_ARG, "bad argument");
  DOHP_CHECK(n->kind == 1, dohpERR_ARG, "sub-cell scales need a cell network (kind 1)");
  hipStream_t st = (hipStream_t)stream;
  DOHP_CALL(ensure(n, ncell));
  DOHP_CALL(forward(n, ncell, X_dev, st));
  hipLaunchKernelGGL(cell_scales_kernel, grid_of(3 * ncell), dim3(256), 0, st, ncell, n->O.ptr, scales_dev);
  DOHP_HIP(hipGetLastError());
  return 0;
}

extern "C" dErr dohpMLPCellStep(dohpMLP n, long ncell, const float *X_dev, const float *gsig_dev, float lr, void *stream)
{
  DOHP_CHECK(n && X_dev && gsig_dev && ncell > 0, dohpERR_ARG, "bad argument");
  DOHP_CHECK(n->kind == 1, dohpERR_ARG, "sub-cell step needs a cell network (kind 1)");
  hipStream_t st = (hipStream_t)stream;
  DOHP_CALL(ensure(n, ncell));
  DOHP_CALL(forward(n, ncell, X_dev, st));
  hipLaunchKernelGGL(dcell_scales_kernel, grid_of(3 * ncell), dim3(256), 0, st, ncell, n->O.ptr, gsig_dev, n->dO.ptr);
  DOHP_CALL(backward_and_grad(n, ncell, X_dev, st));
  DOHP_CALL(adam_step(n, lr, st));
  DOHP_HIP(hipStreamSynchronize(st));
  return 0;
}

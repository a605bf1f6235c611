// Fused matrix-free element kernels of the ellip operator (src/fs/tests/ellip.c:333-472).
//
// One launch evaluates, per element: the gather x_e = E x (fs.c:543-586), the Q1 geometry
// at the Gauss points (fsrulesetit.c:549-561, jacobi.c:383-414), interpolation and gradient
// (dEFSApply INTERP/GRAD, efstopo.c:368-402), the pointwise residual or Jacobian action
// (ellip.c:355-377), the transposed applies (efstopo.c:404-418) and the scatter E^T
// (fs.c:592-640).  Layout: one thread per (i, j) column of an element holds the k column
// in registers; z contractions stay in registers, x and y contractions go through LDS.
// Several elements share a 256-thread workgroup.  Gradients use the collocated derivative
// at the Gauss points (the rule has as many points as the basis has nodes), which equals
// the reference's B * D_gll contraction exactly in exact arithmetic.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "element.h"

using namespace dohp;

namespace {



template <int N, typename T, int MODE, int EPB>
__global__ void __launch_bounds__(wg_threads(N, EPB)) ellip_kernel(const int *__restrict__ elist, long ne,
                                                    const int *__restrict__ enodes, const int *__restrict__ everts,
                                                    const double *__restrict__ vxyz, const double *__restrict__ kv,
                                                    const T *__restrict__ ustate, const T *__restrict__ xdir,
                                                    T *__restrict__ y, long nowned, int atomic,
                                                    const T *__restrict__ gtab, const double *__restrict__ gwphi, Phys ph)
{
  constexpr int NN = N * N, N3 = NN * N, NT = wg_threads(N, EPB), NS = (EPB * N3 + NT - 1) / NT;
  __shared__ T TA[EPB * N3], TGX[EPB * N3], TGY[EPB * N3];
  __shared__ int IDX[EPB * N3];
  __shared__ double VX[EPB * 8 * 4];

  const long lb = xcd_block(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x;
  const int line = tid % (EPB * NN), le = line / NN, ab = line - le * NN, a = ab / N, b = ab - a * N;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int t = s * NT + tid, l = t / N3;
    if (t < EPB * N3) {
      const long es = lb * EPB + l;
      int v = -1;
      if (es < ne) {
        const long e = elist ? elist[es] : es;
        v = enodes[e * N3 + (t - l * N3)];
      }
      IDX[t] = v;
    }
  }
  for (int t = tid; t < EPB * 8; t += NT) {
    const int l = t >> 3, vt = t & 7;
    const long es = lb * EPB + l;
    double *dst = &VX[(l * 8 + vt) * 4];
    if (es < ne) {
      const long e = elist ? elist[es] : es;
      const long vi = everts[e * 8 + vt];
      dst[0] = vxyz[3 * vi];
      dst[1] = vxyz[3 * vi + 1];
      dst[2] = vxyz[3 * vi + 2];
      dst[3] = ph.has_k ? kv[vi] : 1.0;
    } else {  // unit cube for empty slots: finite geometry, nothing is scattered
      dst[0] = vt >> 2;
      dst[1] = (vt >> 1) & 1;
      dst[2] = vt & 1;
      dst[3] = 1.0;
    }
  }
  __syncthreads();
  // gather of the first field, coalesced over the element node lists (tile order = node order);
  // the Jacobian action prefetches the state into registers behind the direction's contractions
  T upre[MODE == MODE_JAC ? NS : 1];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int t = s * NT + tid;
    if (t < EPB * N3) {
      const int c = IDX[t];
      T v = 0;
      if (MODE == MODE_FUN) v = c >= 0 ? ustate[c] : T(0);
      else v = (c >= 0 && c < nowned) ? xdir[c] : T(0);
      TA[t] = v;
      if (MODE == MODE_JAC) upre[s] = c >= 0 ? ustate[c] : T(0);
    }
  }
  __syncthreads();
  int tid_body = tid;  // derived values are formed here, not hoisted above the gather
  asm volatile("" : "+v"(tid_body));
#define DOHP_BATCH_HOOK() ((void)0)
#define DOHP_SCATTER(c, val) \
  do {                       \
    if (atomic)              \
      atomicAdd(&y[c], val); \
    else                     \
      y[c] += val;           \
  } while (0)
#include "ellip_batch.inc"
#undef DOHP_SCATTER
#undef DOHP_BATCH_HOOK
}

// Persistent form of the kernel (one wavefront per workgroup, atomic scatter): each workgroup
// loops over element batches of a contiguous range of its XCD and loads the next batch's node
// indices and vertices (and, for the linear and residual modes, its gathered values) into
// registers while computing the current one.
template <int N, typename T, int MODE, int EPB>
__global__ void __launch_bounds__(wg_threads(N, EPB)) ellip_persistent_kernel(
    long nbatch, long ne, const int *__restrict__ enodes, const int *__restrict__ everts,
    const double *__restrict__ vxyz, const double *__restrict__ kv, const T *__restrict__ ustate,
    const T *__restrict__ xdir, T *__restrict__ y, long nowned, const T *__restrict__ gtab, const double *__restrict__ gwphi, Phys ph)
{
  constexpr int NN = N * N, N3 = NN * N, NT = wg_threads(N, EPB), NS = (EPB * N3 + NT - 1) / NT;
  constexpr bool PREFETCH_VALUES = MODE != MODE_JAC;
  __shared__ T TA[EPB * N3], TGX[EPB * N3], TGY[EPB * N3];
  __shared__ int IDX[EPB * N3];
  __shared__ double VX[EPB * 8 * 4];
  const int tid = threadIdx.x;
  // contiguous batch range of this workgroup's XCD (workgroups are dealt to the 8 XCDs in turn)
  const long per = gridDim.x / 8, xcd = blockIdx.x % 8, w = blockIdx.x / 8;
  const long hi = nbatch * (xcd + 1) / 8;
  long bt = nbatch * xcd / 8 + w;
  constexpr int NV = (EPB * 32 + NT - 1) / NT;  // vertex values (x, y, z, k of 8 vertices) per lane
  int pidx[NS];
  T pval[PREFETCH_VALUES ? NS : 1];
  double pvx[NV];
  auto load_idx = [&](long bb) {
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const int t = s * NT + tid, l = t / N3;
      const long es = bb * EPB + l;
      pidx[s] = (t < EPB * N3 && es < ne) ? enodes[es * N3 + (t - l * N3)] : -1;
    }
#pragma unroll
    for (int q = 0; q < NV; q++) {
      const int t = q * NT + tid;
      if (t < EPB * 32) {
        const int l = t >> 5, vt = (t >> 2) & 7, d = t & 3;
        const long es = bb * EPB + l;
        if (es < ne) {
          const long vi = everts[es * 8 + vt];
          pvx[q] = d < 3 ? vxyz[3 * vi + d] : (ph.has_k ? kv[vi] : 1.0);
        } else {  // unit cube for empty slots: finite geometry, nothing is scattered
          pvx[q] = d == 3 ? 1.0 : (double)((vt >> (2 - d)) & 1);
        }
      }
    }
  };
  auto load_val = [&]() {
    if constexpr (PREFETCH_VALUES) {
#pragma unroll
      for (int s = 0; s < NS; s++) {
        const int c = pidx[s];
        pval[s] = MODE == MODE_FUN ? (c >= 0 ? ustate[c] : T(0)) : ((c >= 0 && c < nowned) ? xdir[c] : T(0));
      }
    }
  };
  if (bt < hi) {
    load_idx(bt);
    load_val();
  }
  for (; bt < hi; bt += per) {
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const int t = s * NT + tid;
      if (t < EPB * N3) {
        IDX[t] = pidx[s];
        if constexpr (PREFETCH_VALUES) TA[t] = pval[s];
      }
    }
#pragma unroll
    for (int q = 0; q < NV; q++)
      if (q * NT + tid < EPB * 32) VX[q * NT + tid] = pvx[q];
    __syncthreads();
    T upre[MODE == MODE_JAC ? NS : 1];
    if constexpr (!PREFETCH_VALUES) {
#pragma unroll
      for (int s = 0; s < NS; s++) {
        const int t = s * NT + tid;
        if (t < EPB * N3) {
          const int c = IDX[t];
          TA[t] = (c >= 0 && c < nowned) ? xdir[c] : T(0);
          upre[s] = c >= 0 ? ustate[c] : T(0);
        }
      }
      __syncthreads();
    }
    const long nxt = bt + per;
    if (nxt < hi) load_idx(nxt);
    // a copy of the thread index the compiler cannot hoist: everything derived from it is
    // recomputed per batch instead of occupying registers across the loop
    int tid_body = tid;
    asm volatile("" : "+v"(tid_body));
    const int atomic = 1;
#define DOHP_BATCH_HOOK() \
  do {                    \
    if (nxt < hi) load_val(); \
  } while (0)
#define DOHP_SCATTER(c, val) \
  do {                       \
    if (atomic)              \
      atomicAdd(&y[c], val); \
    else                     \
      y[c] += val;           \
  } while (0)
#include "ellip_batch.inc"
#undef DOHP_SCATTER
#undef DOHP_BATCH_HOOK
    __syncthreads();  // the next batch overwrites the tiles
  }
}

template <typename T>
__global__ void zero_kernel(T *__restrict__ y, long n)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = T(0);
}

__global__ void closure_kernel(const double *__restrict__ u, const double *__restrict__ dcache, long nowned,
                               long nclosure, double *__restrict__ st, float *__restrict__ st32)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nclosure) return;
  const double v = i < nowned ? u[i] : dcache[i - nowned];
  st[i] = v;
  st32[i] = (float)v;
}

}  // namespace

struct _p_dohpEllip {
  dohpFS fs = nullptr;
  dohpEllipParam prm{};
  int scatter = 1;  // 0 = colours (deterministic), 1 = atomics
  bool has_state = false;
  int shape[3][2] = {{-1, -1}, {-1, -1}, {-1, -1}};  // tuned launch shape per [mode][FP32], -1 = not yet
  float shape_ms[3][2][3] = {};
  DevBuf<double> state, kvert, tab64, wphi;  // wphi: Gauss weights then (1 + z_q)/2
  DevBuf<float> state32, tab32;
  // host tables
  std::vector<double> B, Dh, w, phi;
};

namespace {

template <typename T>
const T *gtab(const _p_dohpEllip *op);
template <>
const double *gtab<double>(const _p_dohpEllip *op) { return op->tab64.ptr; }
template <>
const float *gtab<float>(const _p_dohpEllip *op) { return op->tab32.ptr; }

Phys make_phys(const _p_dohpEllip *op)
{
  Phys ph;
  ph.p = op->prm.p;
  ph.eps2 = op->prm.epsilon * op->prm.epsilon;
  ph.lambda = op->prm.lambda;
  ph.ea = op->prm.a;
  ph.eb = op->prm.b;
  ph.ec = op->prm.c;
  ph.exact = op->prm.exact;
  ph.has_k = op->kvert.ptr != nullptr;
  const double q4 = 2.0 * (op->prm.p - 2.0);
  ph.qpow = (q4 == std::floor(q4) && std::fabs(q4) <= 8.0) ? (int)q4 : 99;
  return ph;
}

template <int N, typename T, int MODE, int EPB>
void launch_shape(const _p_dohpEllip *op, const T *ust, const T *x, T *y, hipStream_t st, const Phys &ph)
{
  const dohpFS fs = op->fs;
  const dim3 block(wg_threads(N, EPB));
  if (op->scatter == 1) {
    const long ne = fs->nelem;
    hipLaunchKernelGGL((ellip_kernel<N, T, MODE, EPB>), dim3((ne + EPB - 1) / EPB), block, 0, st, nullptr, ne,
                       fs->d_elem_nodes.ptr, fs->d_elem_verts.ptr, fs->d_vcoords.ptr, op->kvert.ptr, ust, x, y,
                       fs->nowned, 1, gtab<T>(op), op->wphi.ptr, ph);
  } else {
    for (int col = 0; col < 8; col++) {
      const long ne = fs->color_off[col + 1] - fs->color_off[col];
      if (!ne) continue;
      hipLaunchKernelGGL((ellip_kernel<N, T, MODE, EPB>), dim3((ne + EPB - 1) / EPB), block, 0, st,
                         fs->d_color_elems.ptr + fs->color_off[col], ne, fs->d_elem_nodes.ptr, fs->d_elem_verts.ptr,
                         fs->d_vcoords.ptr, op->kvert.ptr, ust, x, y, fs->nowned, 0, gtab<T>(op), op->wphi.ptr, ph);
    }
  }
}

// Persistent launch: as many one-wavefront workgroups as fit on the device at once (a multiple of
// the 8 XCDs), each looping over its XCD's batches.
template <int N, typename T, int MODE, int EPB>
dErr launch_persistent(const _p_dohpEllip *op, const T *ust, const T *x, T *y, hipStream_t st, const Phys &ph)
{
  const dohpFS fs = op->fs;
  const int threads = wg_threads(N, EPB);
  int dev, per_cu = 0;
  hipDeviceProp_t prop;
  DOHP_HIP(hipGetDevice(&dev));
  DOHP_HIP(hipGetDeviceProperties(&prop, dev));
  DOHP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ellip_persistent_kernel<N, T, MODE, EPB>, threads, 0));
  const long nbatch = (fs->nelem + EPB - 1) / EPB;
  long grid = (long)prop.multiProcessorCount * (per_cu > 0 ? per_cu : 1);
  grid = std::min(grid, ((nbatch + 7) / 8) * 8);
  grid = std::max(8L, grid - grid % 8);
  hipLaunchKernelGGL((ellip_persistent_kernel<N, T, MODE, EPB>), dim3((unsigned)grid), dim3(threads), 0, st, nbatch,
                     (long)fs->nelem, fs->d_elem_nodes.ptr, fs->d_elem_verts.ptr, fs->d_vcoords.ptr, op->kvert.ptr, ust,
                     x, y, fs->nowned, gtab<T>(op), op->wphi.ptr, ph);
  return 0;
}

template <int N, typename T, int MODE>
dErr launch_variant(const _p_dohpEllip *op, int v, const T *ust, const T *x, T *y, hipStream_t st, const Phys &ph)
{
  if (v == 0) launch_shape<N, T, MODE, epb_wide(N)>(op, ust, x, y, st, ph);
  else if (v == 1) launch_shape<N, T, MODE, epb_narrow(N)>(op, ust, x, y, st, ph);
  else DOHP_CALL((launch_persistent<N, T, MODE, epb_narrow(N)>(op, ust, x, y, st, ph)));
  return 0;
}

// Autotuner: the first application of each (mode, precision) times the workgroup shapes (four
// wavefronts, one wavefront, one wavefront persistent) on a scratch output and keeps the fastest
// (the choice is part of the operator; results do not depend on it beyond the summation order
// of the atomic scatter).
template <int N, typename T, int MODE>
dErr launch_n(_p_dohpEllip *op, const T *ust, const T *x, T *y, hipStream_t st)
{
  const Phys ph = make_phys(op);
  int &shape = op->shape[MODE][sizeof(T) == 4];
  if (op->scatter != 1 && shape == 2) shape = -1;  // the persistent form needs the atomic scatter
  if (shape < 0) {
    const long n = op->fs->nowned;
    DevBuf<T> scratch;
    DOHP_CALL(scratch.alloc(n > 0 ? n : 1));
    hipEvent_t e0, e1;
    DOHP_HIP(hipEventCreate(&e0));
    DOHP_HIP(hipEventCreate(&e1));
    float best = 0;
    const int nvar = op->scatter == 1 ? 3 : 2;
    for (int v = 0; v < nvar; v++) {
      float ms = 0;
      for (int rep = 0; rep < 3; rep++) {  // the first run of each shape is a warm-up
        DOHP_HIP(hipEventRecord(e0, st));
        DOHP_CALL((launch_variant<N, T, MODE>(op, v, ust, x, scratch.ptr, st, ph)));
        DOHP_HIP(hipEventRecord(e1, st));
        DOHP_HIP(hipEventSynchronize(e1));
        float t;
        DOHP_HIP(hipEventElapsedTime(&t, e0, e1));
        if (rep > 0) ms += t;
      }
      op->shape_ms[MODE][sizeof(T) == 4][v] = ms / 2;
      if (v == 0 || ms < best) {
        best = ms;
        shape = v;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  DOHP_CALL((launch_variant<N, T, MODE>(op, shape, ust, x, y, st, ph)));
  DOHP_HIP(hipGetLastError());
  return 0;
}

template <typename T, int MODE>
dErr launch(_p_dohpEllip *op, const T *ust, const T *x, T *y, hipStream_t st)
{
  switch (op->fs->n) {
#define DOHP_CASE(NN) \
  case NN: return launch_n<NN, T, MODE>(op, ust, x, y, st);
    DOHP_CASE(2)
    DOHP_CASE(3)
    DOHP_CASE(4)
    DOHP_CASE(5)
    DOHP_CASE(6)
    DOHP_CASE(7)
    DOHP_CASE(8)
    DOHP_CASE(9)
    DOHP_CASE(10)
    DOHP_CASE(11)
#undef DOHP_CASE
  default: set_error("fused operator supports degree 1..10"); return dohpERR_ARG;
  }
}

template <typename T>
dErr zero(T *y, long n, hipStream_t st)
{
  if (n > 0) hipLaunchKernelGGL(zero_kernel<T>, dim3((n + 255) / 256), dim3(256), 0, st, y, n);
  DOHP_HIP(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" dErr dohpEllipCreate(dohpFS fs, const dohpEllipParam *prm, dohpEllip *out)
{
  DOHP_CHECK(fs && prm && out, dohpERR_ARG, "null argument");
  *out = nullptr;
  DOHP_CHECK(fs->n >= 2 && fs->n <= 11, dohpERR_ARG, "fused operator supports degree 1..10");
  const int N = fs->n;
  _p_dohpEllip *op = new _p_dohpEllip;
  op->fs = fs;
  op->prm = *prm;
  // FAST rule of order 2p: N Gauss points (fsruleset.c:32-37, tensorquad.c:84)
  double z[kMaxN], w[kMaxN], zl[kMaxN], wl[kMaxN];
  gauss_legendre(N, z, w);
  gauss_lobatto(N, zl, wl);
  op->B.resize(N * N);
  op->Dh.resize(N * N);
  lagrange_matrix(N, zl, N, z, op->B.data());
  collocation_derivative(N, z, op->Dh.data());
  op->w.assign(w, w + N);
  op->phi.resize(N);
  for (int q = 0; q < N; q++) op->phi[q] = 0.5 * (1.0 + z[q]);
  // geometry validation: negative Jacobian determinant is an error (jacobi.c:397)
  DevBuf<double> dz, mdet;
  dErr err = dz.upload(z, N);
  if (!err) err = mdet.alloc(fs->nelem);
  if (!err) {
    hipLaunchKernelGGL(min_det_kernel, dim3((fs->nelem + 127) / 128), dim3(128), 0, 0, N, fs->d_elem_verts.ptr,
                       fs->d_vcoords.ptr, (long)fs->nelem, dz.ptr, mdet.ptr);
    std::vector<double> h(fs->nelem);
    if (hipMemcpy(h.data(), mdet.ptr, sizeof(double) * fs->nelem, hipMemcpyDeviceToHost) != hipSuccess) {
      set_error("geometry validation failed");
      err = dohpERR_HIP;
    }
    for (int e = 0; !err && e < fs->nelem; e++)
      if (!(h[e] >= 0.0)) {
        set_error("Negative Jacobian at element " + std::to_string(e));
        err = dohpERR_NEGATIVE_JACOBIAN;
      }
  }
  {  // B, Dh, B^T, Dh^T in even-odd form (see eo_apply)
    const std::vector<double> t = eo_tables(N, op->B, op->Dh);
    std::vector<float> tf(t.begin(), t.end());
    if (!err) err = op->tab64.upload(t.data(), t.size());
    if (!err) err = op->tab32.upload(tf.data(), tf.size());
  }
  {
    std::vector<double> wp(op->w);
    wp.insert(wp.end(), op->phi.begin(), op->phi.end());
    if (!err) err = op->wphi.upload(wp.data(), wp.size());
  }
  if (!err) err = op->state.alloc(fs->nclosure);
  if (!err) err = op->state32.alloc(fs->nclosure);
  if (err) {
    delete op;
    return err;
  }
  *out = op;
  return 0;
}

extern "C" dErr dohpEllipDestroy(dohpEllip *op)
{
  if (op && *op) {
    delete *op;
    *op = nullptr;
  }
  return 0;
}

extern "C" dErr dohpEllipSetScatter(dohpEllip op, int mode)
{
  DOHP_CHECK(op && (mode == 0 || mode == 1), dohpERR_ARG, "scatter mode must be 0 (colours) or 1 (atomics)");
  op->scatter = mode;
  return 0;
}

extern "C" dErr dohpEllipGetTuning(dohpEllip op, int mode, int fp32, int *shape, float *ms)
{
  DOHP_CHECK(op && mode >= 0 && mode < 3 && (fp32 == 0 || fp32 == 1), dohpERR_ARG, "bad argument");
  if (shape) *shape = op->shape[mode][fp32];
  if (ms)
    for (int v = 0; v < 3; v++) ms[v] = op->shape_ms[mode][fp32][v];
  return 0;
}

extern "C" dErr dohpEllipSetCoefficient(dohpEllip op, const double *kvert)
{
  DOHP_CHECK(op, dohpERR_ARG, "null operator");
  if (!kvert) {
    op->kvert.release();
    return 0;
  }
  return op->kvert.upload(kvert, op->fs->nvert);
}

extern "C" dErr dohpEllipFunction(dohpEllip op, const double *u, double *f, void *stream)
{
  DOHP_CHECK(op && u && f, dohpERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  const dohpFS fs = op->fs;
  hipLaunchKernelGGL(closure_kernel, dim3((fs->nclosure + 255) / 256), dim3(256), 0, st, u, fs->d_dcache.ptr,
                     fs->nowned, fs->nclosure, op->state.ptr, op->state32.ptr);
  op->has_state = true;
  DOHP_CALL(zero(f, fs->nowned, st));
  return launch<double, MODE_FUN>(op, op->state.ptr, nullptr, f, st);
}

extern "C" dErr dohpEllipFunctionState_Private(dohpEllip op, const double *u, hipStream_t st)
{
  const dohpFS fs = op->fs;
  hipLaunchKernelGGL(closure_kernel, dim3((fs->nclosure + 255) / 256), dim3(256), 0, st, u, fs->d_dcache.ptr,
                     fs->nowned, fs->nclosure, op->state.ptr, op->state32.ptr);
  op->has_state = true;
  DOHP_HIP(hipGetLastError());
  return 0;
}

extern "C" dErr dohpEllipGetQ1Context_Private(dohpEllip op, dohpFS *fs, dohpEllipParam *prm, const double **kvert,
                                              const double **state)
{
  DOHP_CHECK(op, dohpERR_ARG, "null operator");
  *fs = op->fs;
  *prm = op->prm;
  *kvert = op->kvert.ptr;
  *state = op->state.ptr;
  return 0;
}

static bool is_linear(const _p_dohpEllip *op) { return op->prm.p == 2.0 && op->prm.lambda == 0.0; }

extern "C" dErr dohpEllipShellMatMult(dohpEllip op, const double *x, double *y, void *stream)
{
  DOHP_CHECK(op && x && y, dohpERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  DOHP_CALL(zero(y, op->fs->nowned, st));
  if (is_linear(op)) return launch<double, MODE_LIN>(op, nullptr, x, y, st);
  DOHP_CHECK(op->has_state, dohpERR_ARG, "no linearization state: evaluate the function first");
  return launch<double, MODE_JAC>(op, op->state.ptr, x, y, st);
}

extern "C" dErr dohpEllipShellMatMultAdd(dohpEllip op, const double *x, double *y, void *stream)
{
  DOHP_CHECK(op && x && y, dohpERR_ARG, "null argument");
  DOHP_CHECK(op->scatter == 1, dohpERR_ARG, "MatMultAdd accumulates with the atomic scatter only");
  hipStream_t st = (hipStream_t)stream;
  if (is_linear(op)) return launch<double, MODE_LIN>(op, nullptr, x, y, st);
  DOHP_CHECK(op->has_state, dohpERR_ARG, "no linearization state: evaluate the function first");
  return launch<double, MODE_JAC>(op, op->state.ptr, x, y, st);
}

extern "C" dErr dohpEllipShellMatMultF32(dohpEllip op, const float *x, float *y, void *stream)
{
  DOHP_CHECK(op && x && y, dohpERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  DOHP_CALL(zero(y, op->fs->nowned, st));
  if (is_linear(op)) return launch<float, MODE_LIN>(op, nullptr, x, y, st);
  DOHP_CHECK(op->has_state, dohpERR_ARG, "no linearization state: evaluate the function first");
  return launch<float, MODE_JAC>(op, op->state32.ptr, x, y, st);
}

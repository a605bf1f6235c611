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

// Node and vertex numbering of a block space, evaluated instead of read from the element maps:
// element e = (ex M1 + ey) M2 + ez has node (a, b, c) at lattice point (ex p + a, ey p + b, ez p + c)
// and vertex (ta, tb, tc) at ((ex + ta) V1 + ey + tb) V2 + ez + tc; the owned nodes are the box
// [I0, I0 + nx) x [J0, J0 + ny) x [K0, K0 + nz) of the lattice in lexicographic order (the closure
// numbering of dFSCreateBlock), and the state is kept in lattice order (Dirichlet values included).
struct BoxMap {
  int p, M1, M2, V1, V2, Ly, Lz, I0, J0, K0, nx, ny, nz;
};

__device__ __forceinline__ void elem_coords(const BoxMap &m, int e, int &ex, int &ey, int &ez)
{
  const int q = e / m.M2;
  ez = e - q * m.M2;
  ex = q / m.M1;
  ey = q - ex * m.M1;
}
__device__ __forceinline__ int owned_index(const BoxMap &m, int I, int J, int K)
{
  const int i = I - m.I0, j = J - m.J0, k = K - m.K0;
  return ((unsigned)i < (unsigned)m.nx && (unsigned)j < (unsigned)m.ny && (unsigned)k < (unsigned)m.nz)
             ? (i * m.ny + j) * m.nz + k
             : -1;
}
__device__ __forceinline__ long vertex_index(const BoxMap &m, int ex, int ey, int ez, int vt)
{
  return ((long)(ex + (vt >> 2)) * m.V1 + ey + ((vt >> 1) & 1)) * m.V2 + ez + (vt & 1);
}
// node r = (a n + b) n + c of an element whose first lattice point is (I, J, K)
template <int N>
__device__ __forceinline__ void node_lattice(int r, int &I, int &J, int &K)
{
  const int a = r / (N * N), r2 = r - a * N * N, b = r2 / N;
  I += a;
  J += b;
  K += r2 - b * N;
}

// Jacobian action: which field is contracted first.  The state first, reduced to a symmetric metric
// per point before the direction's contractions, needs fewer registers (used in FP32, and in FP64 from
// n = 9, where the direction-first order needs more than 256); the direction first needs fewer
// operations.
template <int N, typename T>
constexpr bool jac_state_first()
{
  return sizeof(T) == 4 || N >= 9;
}

// MF: the contractions as line passes on the matrix cores (n <= 8; 256 threads, one per line of the
// workgroup's tiles, lines past the elements in a pad tile).
template <int N, typename T, int MODE, int EPB, bool MF = false>
__global__ void __launch_bounds__(MF ? 256 : wg_threads(N, EPB))
__attribute__((amdgpu_waves_per_eu((MODE == MODE_FUN || MODE == MODE_QD) && N <= 7 ? 3 : 1))) ellip_kernel(const int *__restrict__ elist, long ne, BoxMap bm,
                                                    const int *__restrict__ enodes, const int *__restrict__ everts,
                                                    const double *__restrict__ vxyz, const double *__restrict__ kv,
                                                    const T *__restrict__ ustate, const T *__restrict__ xdir,
                                                    T *__restrict__ y, long nowned, int atomic,
                                                    const T *__restrict__ gtab, const double *__restrict__ gwphi, Phys ph,
                                                    int *__restrict__ nanflag)
{
  static_assert(!MF || (N <= kMfmaMaxN && EPB * N * N <= 256 && MODE != MODE_QD && MODE != MODE_FRC),
                "matrix-core passes: n <= 8, one line per thread, modes that scatter");
  constexpr int NN = N * N, N3 = NN * N, NT = MF ? 256 : wg_threads(N, EPB), NS = (EPB * N3 + NT - 1) / NT;
  constexpr int ET = EPB + (MF ? 1 : 0);  // element tiles (the pad tile last)
  __shared__ T TA[ET * N3], TGX[ET * N3], TGY[ET * N3];
  __shared__ int IDX[EPB * N3];
  __shared__ double VX[ET * 8 * 4];
  __shared__ int EB[ET * 4];  // per element slot: first lattice point, 1 if the slot holds an element

  const long lb = xcd_block(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x;
  for (int t = tid; t < ET * 8; t += NT) {
    const int l = t >> 3, vt = t & 7;
    const long es = lb * EPB + l;
    double *dst = &VX[(l * 8 + vt) * 4];
    if (es < ne) {
      const int e = elist ? elist[es] : (int)es;
      int ex = 0, ey = 0, ez = 0;
      if (!enodes) elem_coords(bm, e, ex, ey, ez);
      if (MODE != MODE_JQD) {  // (the stored quadrature data carries the geometry)
        const long vi = enodes ? (long)everts[(long)e * 8 + vt] : vertex_index(bm, ex, ey, ez, vt);
        dst[0] = vxyz[3 * vi];
        dst[1] = vxyz[3 * vi + 1];
        dst[2] = vxyz[3 * vi + 2];
        dst[3] = ph.has_k ? kv[vi] : 1.0;
      }
      if (vt == 0) {  // (a space from element maps keeps the element instead of its first lattice point)
        EB[4 * l] = enodes ? e : ex * bm.p;
        EB[4 * l + 1] = ey * bm.p;
        EB[4 * l + 2] = ez * bm.p;
        EB[4 * l + 3] = 1;
      }
    } else {  // unit cube for empty slots: finite geometry, nothing is scattered
      dst[0] = vt >> 2;
      dst[1] = (vt >> 1) & 1;
      dst[2] = vt & 1;
      dst[3] = 1.0;
      if (vt == 0) EB[4 * l + 3] = 0;
    }
  }
  __syncthreads();
  // gather of the first field, coalesced over the element's nodes (tile order = node order); the
  // Jacobian action keeps its second field in registers until the first one's contractions are done.
  // Every slot's indices first, then every load (slots without a value read entry 0 and drop it: no
  // branch separates the loads, so they are in flight together), then the tile.
  T pre2[MODE == MODE_JAC ? NS : 1];
  // stored-data Jacobian: the tensor of this thread's line, loaded with the gathered values (before
  // the barrier) so that both are in flight together
  T gq[MODE == MODE_JQD ? 6 : 1][MODE == MODE_JQD ? N : 1];
  {
    constexpr bool NEED_U = MODE == MODE_FUN || MODE == MODE_JAC || MODE == MODE_QD;
    constexpr bool NEED_X = MODE == MODE_LIN || MODE == MODE_JAC || MODE == MODE_JQD;
    constexpr bool STATE_FIRST =
        MODE == MODE_FUN || MODE == MODE_QD || (MODE == MODE_JAC && jac_state_first<N, T>());
    int own[NS], lat[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const int t = s * NT + tid;
      const bool in = t < EPB * N3;
      const int l = in ? t / N3 : 0;
      const bool live = in && EB[4 * l + 3];
      if (enodes) {  // element maps: closure index; the state is in closure order
        const int c = live ? enodes[(long)EB[4 * l] * N3 + (t - l * N3)] : -1;
        own[s] = c >= 0 && c < nowned ? c : -1;
        lat[s] = c;
      } else {
        int I = EB[4 * l], J = EB[4 * l + 1], K = EB[4 * l + 2];
        node_lattice<N>(t - l * N3, I, J, K);
        own[s] = live ? owned_index(bm, I, J, K) : -1;
        lat[s] = live ? (I * bm.Ly + J) * bm.Lz + K : -1;
      }
    }
    T xv[NS], uv[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) xv[s] = uv[s] = T(0);
    if (NEED_X && nowned > 0) {
#pragma unroll
      for (int s = 0; s < NS; s++) xv[s] = xdir[own[s] > 0 ? own[s] : 0];
    }
    // (the quadrature data of a linear operator is formed without a state)
    if (NEED_U && ustate) {
#pragma unroll
      for (int s = 0; s < NS; s++) uv[s] = ustate[lat[s] > 0 ? lat[s] : 0];
    }
    if constexpr (MODE == MODE_JQD) {  // line and element as in the body; an empty slot reads element 0
      const int ln = MF ? tid : tid % (EPB * NN), l = ln / NN;
      const long es = lb * EPB + l;
      const long el = es < ne ? (elist ? (long)elist[es] : es) : 0;
      const T *q = ustate + el * (ph.lambda != 0 ? 7 : 6) * N3 + (ln - l * NN);
#pragma unroll
      for (int k = 0; k < N; k++)
#pragma unroll
        for (int c = 0; c < 6; c++) gq[c][k] = q[c * N3 + k * NN];
    }
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const int t = s * NT + tid;
      if (t < EPB * N3) {
        const T v = own[s] >= 0 ? xv[s] : T(0), u = lat[s] >= 0 ? uv[s] : T(0);
        IDX[t] = own[s];
        TA[t] = STATE_FIRST ? u : v;
        if (MODE == MODE_JAC) pre2[s] = STATE_FIRST ? v : u;
      }
    }
  }
  __syncthreads();
  int tid_body = tid;  // derived values are formed here, not hoisted above the gather
  asm volatile("" : "+v"(tid_body));
  bool bad = false;  // a non-finite contribution was scattered (NaN or Inf in the inputs or the state)
#define DOHP_ELEM(l) \
  (lb * EPB + (l) < ne ? (elist ? (long)elist[lb * EPB + (l)] : lb * EPB + (l)) : -1L)
#define DOHP_BATCH_HOOK() ((void)0)
#define DOHP_SCATTER(c, val)            \
  do {                                  \
    bad |= !__builtin_isfinite(val);    \
    if (atomic)                         \
      atomicAdd(&y[c], val);            \
    else                                \
      y[c] += val;                      \
  } while (0)
#define DOHP_QD_PRELOADED
#include "ellip_batch.inc"
#undef DOHP_QD_PRELOADED
#undef DOHP_SCATTER
#undef DOHP_BATCH_HOOK
#undef DOHP_ELEM
  if (bad) atomicOr(nanflag, 1);
}

// Persistent form of the kernel (one wavefront per workgroup, atomic scatter): each workgroup
// loops over element batches of a contiguous range of its XCD and loads the next batch's node
// indices and vertices (and, for the linear and residual modes, its gathered values) into
// registers while computing the current one.
template <int N, typename T, int MODE, int EPB>
__global__ void __launch_bounds__(wg_threads(N, EPB))
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && MODE == MODE_LIN ? 3 : 1))) ellip_persistent_kernel(
    long nbatch, long ne, BoxMap bm, const double *__restrict__ vxyz, const double *__restrict__ kv,
    const T *__restrict__ ustate, const T *__restrict__ xdir, T *__restrict__ y, long nowned,
    const T *__restrict__ gtab, const double *__restrict__ gwphi, Phys ph, int *__restrict__ nanflag)
{
  constexpr int NN = N * N, N3 = NN * N, NT = wg_threads(N, EPB), NS = (EPB * N3 + NT - 1) / NT;
  // the Jacobian action prefetches its two fields (state, direction) only in FP32 (in FP64 the
  // registers cost occupancy)
  constexpr bool MF = false;  // (the matrix-core passes are a shape of the non-persistent kernel)
  constexpr bool PREFETCH_VALUES = MODE != MODE_JAC || sizeof(T) == 4;
  constexpr bool GATHER_X = MODE == MODE_LIN || MODE == MODE_JQD;  // else the state
  constexpr bool PREFETCH_STATE = MODE == MODE_JAC && PREFETCH_VALUES;
  __shared__ T TA[EPB * N3], TGX[EPB * N3], TGY[EPB * N3];
  __shared__ int IDX[EPB * N3];
  __shared__ double VX[EPB * 8 * 4];
  const int tid = threadIdx.x;
  // contiguous batch range of this workgroup's XCD (workgroups are dealt to the 8 XCDs in turn)
  const long per = gridDim.x / 8, xcd = blockIdx.x % 8, w = blockIdx.x / 8;
  const long hi = nbatch * (xcd + 1) / 8;
  long bt = nbatch * xcd / 8 + w;
  constexpr int NV = (EPB * 32 + NT - 1) / NT;  // vertex values (x, y, z, k of 8 vertices) per lane
  int pidx[PREFETCH_VALUES ? NS : 1];
  T pval[PREFETCH_VALUES ? NS : 1], pst[PREFETCH_STATE ? NS : 1];
  double pvx[NV];
  // node indices of slot t of batch bb: owned index (-1 if none) and lattice index (-1 past the end)
  auto slot_nodes = [&](long bb, int t, int &own, int &lat) {
    const int l = t / N3;
    const long es = bb * EPB + l;
    own = lat = -1;
    if (t < EPB * N3 && es < ne) {
      int ex, ey, ez;
      elem_coords(bm, (int)es, ex, ey, ez);
      int I = ex * bm.p, J = ey * bm.p, K = ez * bm.p;
      node_lattice<N>(t - l * N3, I, J, K);
      own = owned_index(bm, I, J, K);
      lat = (I * bm.Ly + J) * bm.Lz + K;
    }
  };
  auto load_idx = [&](long bb) {
#pragma unroll
    for (int q = 0; q < NV; q++) {
      const int t = q * NT + tid;
      if (t < EPB * 32) {
        const int l = t >> 5, vt = (t >> 2) & 7, d = t & 3;
        const long es = bb * EPB + l;
        if (MODE == MODE_JQD) {  // the stored quadrature data carries the geometry
          pvx[q] = 0.0;
        } else if (es < ne) {
          int ex, ey, ez;
          elem_coords(bm, (int)es, ex, ey, ez);
          const long vi = vertex_index(bm, ex, ey, ez, vt);
          pvx[q] = d < 3 ? vxyz[3 * vi + d] : (ph.has_k ? kv[vi] : 1.0);
        } else {  // unit cube for empty slots: finite geometry, nothing is scattered
          pvx[q] = d == 3 ? 1.0 : (double)((vt >> (2 - d)) & 1);
        }
      }
    }
  };
  auto load_val = [&](long bb) {
    if constexpr (PREFETCH_VALUES) {
      int tl = tid;  // slot positions are recomputed per batch, not kept live across the loop
      asm volatile("" : "+v"(tl));
#pragma unroll
      for (int s = 0; s < NS; s++) {
        int own, lat;
        slot_nodes(bb, s * NT + tl, own, lat);
        pidx[s] = own;
        pval[s] = GATHER_X ? (own >= 0 ? xdir[own] : T(0)) : (lat >= 0 && ustate ? ustate[lat] : T(0));
        if constexpr (PREFETCH_STATE) pst[s] = own >= 0 ? xdir[own] : T(0);
      }
    }
  };
  if (bt < hi) {
    load_idx(bt);
    load_val(bt);
  }
  bool bad = false;  // a non-finite contribution was scattered
  for (; bt < hi; bt += per) {
    if constexpr (PREFETCH_VALUES) {
#pragma unroll
      for (int s = 0; s < NS; s++) {
        const int t = s * NT + tid;
        if (t < EPB * N3) {
          IDX[t] = pidx[s];
          TA[t] = pval[s];
        }
      }
    }
#pragma unroll
    for (int q = 0; q < NV; q++)
      if (q * NT + tid < EPB * 32) VX[q * NT + tid] = pvx[q];
    T pre2[MODE == MODE_JAC ? NS : 1];
    if constexpr (PREFETCH_STATE) {
#pragma unroll
      for (int s = 0; s < NS; s++) pre2[s] = pst[s];
    }
    if constexpr (!PREFETCH_VALUES) {
      int tl = tid;
      asm volatile("" : "+v"(tl));
#pragma unroll
      for (int s = 0; s < NS; s++) {
        const int t = s * NT + tl;
        int own, lat;
        slot_nodes(bt, t, own, lat);
        if (t < EPB * N3) {
          IDX[t] = own;
          const T xv = own >= 0 ? xdir[own] : T(0), uv = lat >= 0 ? ustate[lat] : T(0);
          TA[t] = jac_state_first<N, T>() ? uv : xv;
          pre2[s] = jac_state_first<N, T>() ? xv : uv;
        }
      }
    }
    __syncthreads();
    const long nxt = bt + per;
    if (nxt < hi) load_idx(nxt);
    // a copy of the thread index the compiler cannot hoist: everything derived from it is
    // recomputed per batch instead of occupying registers across the loop
    int tid_body = tid;
    asm volatile("" : "+v"(tid_body));
    const int atomic = 1;
#define DOHP_ELEM(l) (bt * EPB + (l) < ne ? bt * EPB + (l) : -1L)
#define DOHP_BATCH_HOOK() \
  do {                    \
    if (nxt < hi) load_val(nxt); \
  } while (0)
#define DOHP_SCATTER(c, val)            \
  do {                                  \
    bad |= !__builtin_isfinite(val);    \
    if (atomic)                         \
      atomicAdd(&y[c], val);            \
    else                                \
      y[c] += val;                      \
  } while (0)
#include "ellip_batch.inc"
#undef DOHP_SCATTER
#undef DOHP_BATCH_HOOK
#undef DOHP_ELEM
    __syncthreads();  // the next batch overwrites the tiles
  }
  if (bad) atomicOr(nanflag, 1);
}

template <typename T>
__global__ void zero_kernel(T *__restrict__ y, long n)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = T(0);
}

// Linearization state in lattice order (closure order for a space from element maps, lat2clo = null): owned
// values from u, Dirichlet values from the cache.
__global__ void lattice_state_kernel(const double *__restrict__ u, const double *__restrict__ dcache,
                                     const int *__restrict__ lat2clo, long nowned, long nlat, double *__restrict__ st,
                                     float *__restrict__ st32)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nlat) return;
  const long c = lat2clo ? lat2clo[i] : i;
  const double v = c < nowned ? u[c] : dcache[c - nowned];
  st[i] = v;
  st32[i] = (float)v;
}

}  // namespace

struct _p_dohpEllip {
  dohp::Pool pool;  // work vectors of the solves on this operator (freed with it)
  dohpFS fs = nullptr;
  dohpEllipParam prm{};
  int scatter = 1;  // 0 = colours (deterministic), 1 = atomics
  int q1mode = 0;   // embedded-Q1 assembly: 0 = the reference's arithmetic, 1 = corrected sub-cell derivatives
  bool has_state = false;
  int shape[kModes][2] = {{-1, -1}, {-1, -1}, {-1, -1}, {-1, -1}, {-1, -1}, {-1, -1}};  // tuned per [mode][FP32]
  float shape_ms[kModes][2][kShapes] = {};
  // Jacobian (or linear operator) quadrature data, 6 values per point (7 with the Bratu term), formed
  // on the first application after each residual (the linear operator's once).  storage: 0 = never,
  // 1 = always, 2 = the faster of the two, timed on the first application in each precision (use_qd:
  // -1 = not timed yet, 0 = recompute, 1 = stored)
  int storage = 2;
  int use_qd[2] = {-1, -1};
  float qd_ms[2][2] = {};  // [FP32][recomputed, stored] milliseconds of that timing
  DevBuf<double> qd;
  DevBuf<float> qd32;
  bool qd_ok = false, qd32_ok = false;
  DevBuf<double> state, kvert, tab64, wphi;  // state in lattice order; wphi: Gauss weights then (1 + z_q)/2
  DevBuf<double> forcing;                     // assembled forcing term -(f, phi) on the owned nodes
  bool has_forcing = false;
  DevBuf<float> state32, tab32;
  // set by the element kernels when a scattered contribution is not finite (a NaN or Inf in the input
  // vector or the state); the checked entry points clear it, wait for their stream and read it back
  DevBuf<int> nanflag;
  int *nan_host = nullptr;  // pinned
  int checks = 1;
  const float *q1_cscale = nullptr;  // learned sub-cell derivative scales of later Q1 assemblies (caller's array)
  // host tables
  std::vector<double> B, Dh, w, phi;
  ~_p_dohpEllip()
  {
    if (nan_host) (void)hipHostFree(nan_host);
  }
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

BoxMap box_map(const _p_dohpFS *fs)
{
  BoxMap m;
  const int L[3] = {fs->p * fs->M[0] + 1, fs->p * fs->M[1] + 1, fs->p * fs->M[2] + 1};
  const int walls = fs->dirichlet_mask & 1, top = (fs->dirichlet_mask >> 1) & 1, bottom = (fs->dirichlet_mask >> 2) & 1;
  m.p = fs->p;
  m.M1 = fs->M[1];
  m.M2 = fs->M[2];
  m.V1 = fs->M[1] + 1;
  m.V2 = fs->M[2] + 1;
  m.Ly = L[1];
  m.Lz = L[2];
  m.I0 = walls;
  m.J0 = walls;
  m.K0 = bottom;
  m.nx = L[0] - 2 * walls;
  m.ny = L[1] - 2 * walls;
  m.nz = L[2] - bottom - top;
  return m;
}

template <int N, typename T, int MODE, int EPB, bool MF = false>
void launch_shape(const _p_dohpEllip *op, const T *ust, const T *x, T *y, hipStream_t st, const Phys &ph)
{
  const dohpFS fs = op->fs;
  const dim3 block(MF ? 256 : wg_threads(N, EPB));
  // a space from element maps is read through them (block spaces evaluate their numbering)
  const int *enodes = fs->is_block() ? nullptr : fs->d_elem_nodes.ptr;
  const int *everts = fs->is_block() ? nullptr : fs->d_elem_verts.ptr;
  if (op->scatter == 1) {
    const long ne = fs->nelem;
    hipLaunchKernelGGL((ellip_kernel<N, T, MODE, EPB, MF>), dim3((ne + EPB - 1) / EPB), block, 0, st, nullptr, ne,
                       box_map(fs), enodes, everts, fs->d_vcoords.ptr, op->kvert.ptr, ust, x, y, fs->nowned, 1,
                       gtab<T>(op), op->wphi.ptr, ph, op->nanflag.ptr);
  } else {
    for (int col = 0; col < fs->ncolor; col++) {
      const long ne = fs->color_off[col + 1] - fs->color_off[col];
      if (!ne) continue;
      hipLaunchKernelGGL((ellip_kernel<N, T, MODE, EPB, MF>), dim3((ne + EPB - 1) / EPB), block, 0, st,
                         fs->d_color_elems.ptr + fs->color_off[col], ne, box_map(fs), enodes, everts,
                         fs->d_vcoords.ptr, op->kvert.ptr, ust, x, y, fs->nowned, 0, gtab<T>(op), op->wphi.ptr, ph,
                         op->nanflag.ptr);
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
  int dev, cus;
  DOHP_CALL(current_device_cus(dev, cus));
  static int occupancy[64] = {};  // workgroups per CU of this instance, per device (asked once)
  if (!occupancy[dev]) {
    int per_cu = 0;
    DOHP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ellip_persistent_kernel<N, T, MODE, EPB>, threads,
                                                          0));
    occupancy[dev] = per_cu > 0 ? per_cu : 1;
  }
  const long nbatch = (fs->nelem + EPB - 1) / EPB;
  long grid = (long)cus * occupancy[dev];
  grid = std::min(grid, ((nbatch + 7) / 8) * 8);
  grid = std::max(8L, grid - grid % 8);
  hipLaunchKernelGGL((ellip_persistent_kernel<N, T, MODE, EPB>), dim3((unsigned)grid), dim3(threads), 0, st, nbatch,
                     (long)fs->nelem, box_map(fs), fs->d_vcoords.ptr, op->kvert.ptr, ust, x, y, fs->nowned,
                     gtab<T>(op), op->wphi.ptr, ph, op->nanflag.ptr);
  return 0;
}

// Degrees whose fused kernels have the matrix-core shape (shape 4).
constexpr bool has_mfma_shape(int N) { return N >= 4 && N <= kMfmaMaxN; }

template <int N, typename T, int MODE>
dErr launch_variant(const _p_dohpEllip *op, int v, const T *ust, const T *x, T *y, hipStream_t st, const Phys &ph)
{
  if (v == 0) launch_shape<N, T, MODE, epb_wide(N)>(op, ust, x, y, st, ph);
  else if (v == 1) launch_shape<N, T, MODE, epb_narrow(N)>(op, ust, x, y, st, ph);
  else if (v == 2) DOHP_CALL((launch_persistent<N, T, MODE, epb_narrow(N)>(op, ust, x, y, st, ph)));
  else if (v == 3) DOHP_CALL((launch_persistent<N, T, MODE, epb_wide(N)>(op, ust, x, y, st, ph)));
  else if constexpr (has_mfma_shape(N)) launch_shape<N, T, MODE, epb_wide(N), true>(op, ust, x, y, st, ph);
  else DOHP_CHECK(false, dohpERR_ARG, "no matrix-core shape at this degree");
  return 0;
}

// Autotuner: the first application of each (mode, precision) times the workgroup shapes (four
// wavefronts, one wavefront, the persistent forms of both, and in FP32 from n = 4 to 8 the
// four-wavefront shape with the contractions on the matrix cores) on a scratch output and keeps the fastest
// (the choice is part of the operator; results do not depend on it beyond the summation order
// of the atomic scatter).
template <int N, typename T, int MODE>
dErr launch_n(_p_dohpEllip *op, const T *ust, const T *x, T *y, hipStream_t st)
{
  const Phys ph = make_phys(op);
  int &shape = op->shape[MODE][sizeof(T) == 4];
  // the persistent forms need the atomic scatter and a block space
  if ((op->scatter != 1 || !op->fs->is_block()) && (shape == 2 || shape == 3)) shape = -1;
  if (shape < 0) {
    const long n = op->fs->nowned;
    DevBuf<T> scratch;
    DOHP_CALL(scratch.alloc(n > 0 ? n : 1));
    hipEvent_t e0, e1;
    DOHP_HIP(hipEventCreate(&e0));
    DOHP_HIP(hipEventCreate(&e1));
    float best = 0;
    for (int v = 0; v < kShapes; v++) {
      if ((v == 2 || v == 3) && (op->scatter != 1 || !op->fs->is_block())) continue;
      // (the matrix-core shape is timed in FP32 only: FP64 MFMA issues at the FP64 vector rate, so its padded
      // 16x16 tiles cannot win there; it stays available through dohpEllipSetShape)
      if (v == 4 && (!has_mfma_shape(N) || sizeof(T) != 4)) continue;
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

bool is_linear(const _p_dohpEllip *op) { return op->prm.p == 2.0 && op->prm.lambda == 0.0; }

// One launch of the four-wavefront shape, outside the tuner (modes whose output is not a vector of
// owned nodes: the forcing term's is, but it runs once).
template <int MODE>
dErr launch_fixed(const _p_dohpEllip *op, const double *ust, double *y, hipStream_t st)
{
  const Phys ph = make_phys(op);
  switch (op->fs->n) {
#define DOHP_CASE(NN)                                                                  \
  case NN:                                                                             \
    launch_shape<NN, double, MODE, epb_wide(NN)>(op, ust, nullptr, y, st, ph);       \
    break;
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
  DOHP_HIP(hipGetLastError());
  return 0;
}

// Quadrature data of the current linearization (of the mesh and k alone for the linear operator),
// formed once and reused by every application until the state or the coefficient changes: the
// stash that the reference's residual leaves for its Jacobian action (ellip.c:311, 333-342, 460-467),
// here reduced with the geometry to one symmetric tensor per point.
dErr ensure_qd(_p_dohpEllip *op, hipStream_t st)
{
  if (op->qd_ok) return 0;
  const dohpFS fs = op->fs;
  // one element's worth of slack at the end: the kernels may load the Bratu component of the last
  // element's data ahead of the (uniform) test that it exists
  const size_t n = ((size_t)fs->nelem * (op->prm.lambda != 0 ? 7 : 6) + 1) * fs->n * fs->n * fs->n;
  if (!op->qd.holds(n)) DOHP_CALL(op->qd.alloc(n));
  DOHP_CALL(launch_fixed<MODE_QD>(op, is_linear(op) ? nullptr : op->state.ptr, op->qd.ptr, st));
  op->qd_ok = true;
  op->qd32_ok = false;
  return 0;
}

__global__ void round_f32_kernel(const double *__restrict__ a, float *__restrict__ b, long n)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

dErr ensure_qd32(_p_dohpEllip *op, hipStream_t st)
{
  if (op->qd32_ok && op->qd_ok) return 0;
  DOHP_CALL(ensure_qd(op, st));
  const long n = (long)op->qd.n;
  if (!op->qd32.holds(op->qd.n)) DOHP_CALL(op->qd32.alloc(op->qd.n));
  hipLaunchKernelGGL(round_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, op->qd.ptr, op->qd32.ptr,
                     n);
  DOHP_HIP(hipGetLastError());
  op->qd32_ok = true;
  return 0;
}

void drop_qd(_p_dohpEllip *op) { op->qd_ok = op->qd32_ok = false; }

// The forcing term of the residual (exact solutions 0..4), assembled on first use.
dErr assemble_forcing(_p_dohpEllip *op, hipStream_t st)
{
  const dohpFS fs = op->fs;
  DOHP_CALL(op->forcing.alloc(fs->nowned));
  hipLaunchKernelGGL(zero_kernel<double>, dim3((fs->nowned + 255) / 256), dim3(256), 0, st, op->forcing.ptr,
                     fs->nowned);
  DOHP_CALL(launch_fixed<MODE_FRC>(op, nullptr, op->forcing.ptr, st));
  op->has_forcing = true;
  return 0;
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
  {  // B, Dh, B^T, Dh^T in even-odd form (see eo_apply), then the matrix-core line passes' fragments
    std::vector<double> t = eo_tables(N, op->B, op->Dh);
    if (N <= kMfmaMaxN) {
      const std::vector<double> f = mfma_frags(N, op->B, op->Dh);
      t.insert(t.end(), f.begin(), f.end());
    }
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
  if (!err) err = op->nanflag.alloc(1);
  if (!err && hipMemset(op->nanflag.ptr, 0, sizeof(int)) != hipSuccess) err = dohpERR_HIP;
  if (!err && hipHostMalloc((void **)&op->nan_host, sizeof(int), hipHostMallocDefault) != hipSuccess) {
    op->nan_host = nullptr;
    set_error("pinned allocation failed");
    err = dohpERR_OOM;
  }
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

// tuning modes of the C API: 0 linear, 1 residual, 2 Jacobian recomputed, 3 apply from quadrature data
static const int kApiMode[4] = {MODE_LIN, MODE_FUN, MODE_JAC, MODE_JQD};

extern "C" dErr dohpEllipGetTuning(dohpEllip op, int mode, int fp32, int *shape, float *ms)
{
  DOHP_CHECK(op && mode >= 0 && mode < 4 && (fp32 == 0 || fp32 == 1), dohpERR_ARG, "bad argument");
  const int m = kApiMode[mode];
  if (shape) *shape = op->shape[m][fp32];
  if (ms)
    for (int v = 0; v < kShapes; v++) ms[v] = op->shape_ms[m][fp32][v];
  return 0;
}

extern "C" dErr dohpEllipSetShape(dohpEllip op, int mode, int fp32, int shape)
{
  DOHP_CHECK(op && mode >= 0 && mode < 4 && (fp32 == 0 || fp32 == 1) && shape >= -1 && shape < kShapes, dohpERR_ARG,
             "bad argument");
  DOHP_CHECK((shape != 2 && shape != 3) || op->scatter == 1, dohpERR_ARG,
             "the persistent shapes need the atomic scatter");
  DOHP_CHECK((shape != 2 && shape != 3) || op->fs->is_block(), dohpERR_ARG,
             "the persistent shapes need a block space (dFSCreateBlock)");
  DOHP_CHECK(shape != 4 || (op->fs->n >= 4 && op->fs->n <= kMfmaMaxN), dohpERR_ARG,
             "the matrix-core shape needs degree 3..7");
  op->shape[kApiMode[mode]][fp32] = shape;
  return 0;
}

extern "C" dErr dohpEllipSetCoefficient(dohpEllip op, const double *kvert)
{
  DOHP_CHECK(op, dohpERR_ARG, "null operator");
  drop_qd(op);
  if (!kvert) {
    op->kvert.release();
    return 0;
  }
  return op->kvert.upload(kvert, op->fs->nvert);
}

namespace {

// Checked form of an entry point (checks on): clears the non-finite flag, runs f, waits for the stream
// and reports a non-finite value met by the element kernels.  Checks off: f alone, asynchronous.
template <typename F>
dErr checked(_p_dohpEllip *op, hipStream_t st, const char *what, F f)
{
  if (!op->checks) return f();
  DOHP_CALL(zero(op->nanflag.ptr, 1L, st));
  DOHP_CALL(f());
  DOHP_HIP(hipMemcpyAsync(op->nan_host, op->nanflag.ptr, sizeof(int), hipMemcpyDeviceToHost, st));
  DOHP_HIP(hipStreamSynchronize(st));
  DOHP_CHECK(*op->nan_host == 0, dohpERR_NAN,
             std::string(what) + ": non-finite value met on the device (NaN or Inf in the input vector or the state)");
  return 0;
}

dErr ellip_function(_p_dohpEllip *op, const double *u, double *f, hipStream_t st)
{
  const dohpFS fs = op->fs;
  hipLaunchKernelGGL(lattice_state_kernel, dim3((fs->nclosure + 255) / 256), dim3(256), 0, st, u, fs->d_dcache.ptr,
                     fs->d_lat2clo.ptr, fs->nowned, fs->nclosure, op->state.ptr, op->state32.ptr);
  op->has_state = true;
  drop_qd(op);
  if (op->prm.exact >= 0) {  // y = forcing term, then the state's terms are added
    if (!op->has_forcing) DOHP_CALL(assemble_forcing(op, st));
    DOHP_HIP(hipMemcpyAsync(f, op->forcing.ptr, sizeof(double) * fs->nowned, hipMemcpyDeviceToDevice, st));
  } else {
    DOHP_CALL(zero(f, fs->nowned, st));
  }
  return launch<double, MODE_FUN>(op, op->state.ptr, nullptr, f, st);
}

}  // namespace

extern "C" dErr dohpEllipFunction(dohpEllip op, const double *u, double *f, void *stream)
{
  DOHP_CHECK(op && u && f, dohpERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  return checked(op, st, "dohpEllipFunction", [&]() { return ellip_function(op, u, f, st); });
}

extern "C" dErr dohpEllipSetChecks(dohpEllip op, int mode)
{
  DOHP_CHECK(op && (mode == 0 || mode == 1), dohpERR_ARG, "checks must be 0 (off) or 1 (on)");
  op->checks = mode;
  return 0;
}

extern "C" dErr dohpEllipFunctionState_Private(dohpEllip op, const double *u, hipStream_t st)
{
  const dohpFS fs = op->fs;
  hipLaunchKernelGGL(lattice_state_kernel, dim3((fs->nclosure + 255) / 256), dim3(256), 0, st, u, fs->d_dcache.ptr,
                     fs->d_lat2clo.ptr, fs->nowned, fs->nclosure, op->state.ptr, op->state32.ptr);
  op->has_state = true;
  drop_qd(op);
  DOHP_HIP(hipGetLastError());
  return 0;
}

dohp::Pool *dohpEllipPool_Private(dohpEllip op) { return &op->pool; }

extern "C" dErr dohpEllipGetQ1Context_Private(dohpEllip op, dohpFS *fs, dohpEllipParam *prm, const double **kvert,
                                              const double **state, int *q1mode)
{
  DOHP_CHECK(op, dohpERR_ARG, "null operator");
  *fs = op->fs;
  *prm = op->prm;
  *kvert = op->kvert.ptr;
  *state = op->state.ptr;
  if (q1mode) *q1mode = op->q1mode;
  return 0;
}

const float *dohpEllipGetQ1CellScales_Private(dohpEllip op) { return op->q1_cscale; }

extern "C" dErr dohpEllipSetQ1CellScales(dohpEllip op, const float *scales_dev)
{
  DOHP_CHECK(op, dohpERR_ARG, "null operator");
  op->q1_cscale = scales_dev;
  return 0;
}

extern "C" dErr dohpEllipSetQ1Mode(dohpEllip op, int mode)
{
  DOHP_CHECK(op && (mode == 0 || mode == 1), dohpERR_ARG, "Q1 mode must be 0 (reference) or 1 (corrected)");
  op->q1mode = mode;
  return 0;
}

namespace {

// y += J x (A x for the linear operator): from the stored quadrature data, or recomputed from the state
// and the geometry at every application (storage 0)
template <typename T>
struct Prec;
template <>
struct Prec<double> {
  static dErr ensure(_p_dohpEllip *op, hipStream_t st) { return ensure_qd(op, st); }
  static const double *qd(const _p_dohpEllip *op) { return op->qd.ptr; }
  static const double *state(const _p_dohpEllip *op) { return op->state.ptr; }
};
template <>
struct Prec<float> {
  static dErr ensure(_p_dohpEllip *op, hipStream_t st) { return ensure_qd32(op, st); }
  static const float *qd(const _p_dohpEllip *op) { return op->qd32.ptr; }
  static const float *state(const _p_dohpEllip *op) { return op->state32.ptr; }
};

template <typename T>
dErr apply_recomputed(_p_dohpEllip *op, const T *x, T *y, hipStream_t st)
{
  if (is_linear(op)) return launch<T, MODE_LIN>(op, nullptr, x, y, st);
  return launch<T, MODE_JAC>(op, Prec<T>::state(op), x, y, st);
}

template <typename T>
dErr apply_stored(_p_dohpEllip *op, const T *x, T *y, hipStream_t st)
{
  DOHP_CALL(Prec<T>::ensure(op, st));
  return launch<T, MODE_JQD>(op, Prec<T>::qd(op), x, y, st);
}

// milliseconds of one call of f on stream st (after one untimed call, which also tunes its launch shape)
template <typename F>
dErr time_call(F f, hipStream_t st, float &ms)
{
  hipEvent_t e0, e1;
  DOHP_CALL(f());
  DOHP_HIP(hipEventCreate(&e0));
  DOHP_HIP(hipEventCreate(&e1));
  DOHP_HIP(hipEventRecord(e0, st));
  for (int r = 0; r < 2; r++) DOHP_CALL(f());
  DOHP_HIP(hipEventRecord(e1, st));
  DOHP_HIP(hipEventSynchronize(e1));
  DOHP_HIP(hipEventElapsedTime(&ms, e0, e1));
  ms /= 2;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 0;
}

// y += J x (A x for the linear operator): from the stored quadrature data or recomputed from the state and
// the geometry.  With storage 2 the first application in each precision times both forms on a scratch
// output and keeps the faster (per application; forming the data is paid once per linearization).
template <typename T>
dErr apply_add(_p_dohpEllip *op, const T *x, T *y, hipStream_t st)
{
  if (!is_linear(op)) DOHP_CHECK(op->has_state, dohpERR_ARG, "no linearization state: evaluate the function first");
  const int p32 = sizeof(T) == 4;
  int use = op->storage == 2 ? op->use_qd[p32] : op->storage;
  if (use < 0) {
    DevBuf<T> scratch;
    DOHP_CALL(scratch.alloc(op->fs->nowned > 0 ? op->fs->nowned : 1));
    float t_qd = 0, t_re = 0;
    // the data may not fit in device memory: then the recomputing form is kept (any other failure is
    // returned)
    const dErr fit_err = Prec<T>::ensure(op, st);
    if (fit_err && fit_err != dohpERR_OOM) return fit_err;
    const bool fits = fit_err == 0;
    if (fits) {
      DOHP_CALL(time_call([&]() { return apply_stored<T>(op, x, scratch.ptr, st); }, st, t_qd));
    } else {
      drop_qd(op);
      op->qd.release();
      op->qd32.release();
    }
    DOHP_CALL(time_call([&]() { return apply_recomputed<T>(op, x, scratch.ptr, st); }, st, t_re));
    use = op->use_qd[p32] = fits && t_qd < t_re ? 1 : 0;
    op->qd_ms[p32][0] = t_re;
    op->qd_ms[p32][1] = t_qd;
    if (!use && op->use_qd[1 - p32] != 1) {  // neither precision reads it: the memory is returned
      drop_qd(op);
      op->qd.release();
      op->qd32.release();
    }
  }
  return use ? apply_stored<T>(op, x, y, st) : apply_recomputed<T>(op, x, y, st);
}

}  // namespace

// Unchecked applications for the library's own solvers (their Krylov loops test the norms instead).
dErr dohpEllipMatMult_Private(dohpEllip op, const double *x, double *y, hipStream_t st)
{
  DOHP_CALL(zero(y, op->fs->nowned, st));
  return apply_add(op, x, y, st);
}
dErr dohpEllipMatMult_Private(dohpEllip op, const float *x, float *y, hipStream_t st)
{
  DOHP_CALL(zero(y, op->fs->nowned, st));
  return apply_add(op, x, y, st);
}

extern "C" dErr dohpEllipShellMatMult(dohpEllip op, const double *x, double *y, void *stream)
{
  DOHP_CHECK(op && x && y, dohpERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  return checked(op, st, "dohpEllipShellMatMult", [&]() { return dohpEllipMatMult_Private(op, x, y, st); });
}

extern "C" dErr dohpEllipShellMatMultAdd(dohpEllip op, const double *x, double *y, void *stream)
{
  DOHP_CHECK(op && x && y, dohpERR_ARG, "null argument");
  DOHP_CHECK(op->scatter == 1, dohpERR_ARG, "MatMultAdd accumulates with the atomic scatter only");
  hipStream_t st = (hipStream_t)stream;
  return checked(op, st, "dohpEllipShellMatMultAdd", [&]() { return apply_add(op, x, y, st); });
}

extern "C" dErr dohpEllipShellMatMultF32(dohpEllip op, const float *x, float *y, void *stream)
{
  DOHP_CHECK(op && x && y, dohpERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  return checked(op, st, "dohpEllipShellMatMultF32", [&]() { return dohpEllipMatMult_Private(op, x, y, st); });
}

extern "C" dErr dohpEllipGetStorage(dohpEllip op, int fp32, int *stored, float *ms)
{
  DOHP_CHECK(op && (fp32 == 0 || fp32 == 1), dohpERR_ARG, "bad argument");
  if (stored) *stored = op->storage == 2 ? op->use_qd[fp32] : op->storage;
  if (ms) {
    ms[0] = op->qd_ms[fp32][0];
    ms[1] = op->qd_ms[fp32][1];
  }
  return 0;
}

extern "C" dErr dohpEllipSetStorage(dohpEllip op, int mode)
{
  DOHP_CHECK(op && mode >= 0 && mode <= 2, dohpERR_ARG,
             "storage must be 0 (recompute), 1 (quadrature data) or 2 (the faster, timed)");
  op->storage = mode;
  op->use_qd[0] = op->use_qd[1] = -1;
  if (mode == 0) {  // the memory is returned
    drop_qd(op);
    op->qd.release();
    op->qd32.release();
  }
  return 0;
}

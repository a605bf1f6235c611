// Batched matrix-free Poisson operator over elements of mixed order, applied in one persistent
// launch.
//
// Segments of the batch share one packed vector:
//   - a block space (dFSCreateBlock): y = A x on its owned nodes, A the homogeneous-Dirichlet
//     stiffness of -div(k grad u), the Jacobian of ellip at p = 2 (EllipShellMatMult,
//     src/fs/tests/ellip.c:442-472);
//   - independent hexahedra of per-element degree 1..10, each with its own n^3 GLL nodes (the
//     element-local expanded space, element slices at offsets sum n_e^3 as src/fs/impls/cont/cont.c:372):
//     y_e = K_e x_e with K_e the element stiffness, i.e. dEFSApply GRAD, jw k grad, GRAD_T
//     (src/jacobi/impls/tensor/efstopo.c:368-423, src/jacobi/interface/jacobi.c:383-414) at the
//     element's own order, as the mixed-order dEFS arrays of src/jacobi/tests/ex1.c:282-309.  The
//     reference rejects continuity across order jumps (src/jacobi/impls/tensor/tensor.c:417), so
//     different orders meet only in this element-local form.
// Block-space segments come first in the packed vector (they are cleared before the atomic
// scatter), element segments after them (plain stores).  Elements are bucketed by order; a work
// list of items (order and tile, first element, count, store kind) is consumed by one persistent
// kernel through an atomic counter, highest order first.  The tile of each order (elements packed
// over the four wavefronts of a workgroup with workgroup barriers, or packed per wavefront with
// wavefront-local synchronization) is chosen by timing both on the first application in each
// precision.  From degree 4 up, FP32 applies may also offer the tuner a third tile whose contractions are
// wavefront-level GEMMs on the matrix cores (BF16 operands, FP32 accumulation; dohpBatchSetMFMA).
#include <algorithm>
#include <map>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "element.h"

using namespace dohp;

namespace {

constexpr int kBT = 256;                  // threads of a batch workgroup
constexpr int kNMin = 2, kNMax = 11;      // nodes per direction (degree 1..10)
enum { PACK_BLOCK = 0, PACK_WAVE = 1, PACK_MFMA = 2 };
constexpr int kPacks = 3;

// Offsets of an order's tables in the concatenated tables of n = 2..11.
constexpr int tab_off(int N)
{
  int s = 0;
  for (int m = kNMin; m < N; m++) s += kEoOps * eo_size(m);
  return s;
}
constexpr int wphi_off(int N) { return N * (N - 1) - 2; }  // sum of 2m over m = 2..N-1
constexpr int ftab_off(int N)                               // full B, Dh, B^T, Dh^T per order
{
  int s = 0;
  for (int m = kNMin; m < N; m++) s += 4 * m * m;
  return s;
}
constexpr int epw(int N) { return 64 / (N * N); }           // elements per wavefront (0 if n^2 > 64)
constexpr int item_elems(int N, int pack) { return pack == PACK_WAVE ? 4 * epw(N) : kBT / (N * N); }
constexpr int cfg_id(int N, int pack) { return (N - kNMin) * kPacks + pack; }
constexpr int kMfmaMinN = 5;  // MFMA contractions from degree 4 up
inline bool pack_valid(int N, int pack, int tsize, int mfma)
{
  if (pack == PACK_WAVE) return epw(N) > 0;
  if (pack == PACK_MFMA) return mfma > 0 && tsize == 4 && N >= kMfmaMinN;
  return true;
}
inline size_t cfg_lds(int N, int pack, int tsize)
{
  const size_t E = item_elems(N, pack), N3 = (size_t)N * N * N;
  return E * 32 * sizeof(double) + 3 * E * N3 * tsize + E * N3 * sizeof(int);
}

struct BArgs {
  const int4 *items;
  int nitems;
  int *counter;
  const int *gidx;     // per bucket: [elements][n^3] packed index of each node, -1 = Dirichlet
  const int *everts;   // per bucket: [elements][8] packed vertex index, tensor order
  const double *vxyz;  // [vertices][3]
  const double *kv;    // [vertices] coefficient k
  const double *wphi;  // per order: Gauss weights then (1 + z_q)/2
  const float *ftab;   // per order: B, Dh, B^T, Dh^T as full n x n matrices [q][m]
  int passes;          // MFMA tiles: 1 = BF16 operands, 3 = operands split into two BF16 parts
  long goff[kNMax + 1], eoff[kNMax + 1];
};

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short bf16_bits(float v) { return __builtin_bit_cast(short, (__bf16)v); }
__device__ __forceinline__ float bf16_value(short h) { return __builtin_bit_cast(float, (unsigned)(unsigned short)h << 16); }

// A 1-D operator Op[q][m] (n x n) as the B operand of v_mfma_f32_16x16x16_bf16: lane l holds
// Op[q = l % 16][m = 4 (l / 16) + i], i = 0..3, zero outside n x n; hi is its BF16 rounding, lo the BF16
// rounding of the rest.
template <int N>
__device__ __forceinline__ void mfma_operator(const float *__restrict__ op, int lane, bf16x4 &hi, bf16x4 &lo)
{
  const int q = lane & 15, kg = lane >> 4;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int m = 4 * kg + i;
    const float v = (q < N && m < N) ? op[q * N + m] : 0.f;
    hi[i] = bf16_bits(v);
    lo[i] = bf16_bits(v - bf16_value(hi[i]));
  }
}

// One contraction along DIR of the E element tiles of a workgroup as wavefront-level GEMMs: the lines of a tile
// (n^2 of them) are the GEMM rows, taken 16 at a time; row block x operator is one 16x16x16 BF16 MFMA with FP32
// accumulation (three with split operands: hi.hi + lo.hi + hi.lo).  Wavefront w takes the (element, row block)
// pairs w, w + 4, ...  out = Op in (ACC: out += Op in); in == out is allowed, since a block's rows are read
// before they are written and blocks share no lines.
template <int N, int DIR, bool ACC, int E>
__device__ __forceinline__ void mfma_contract(const float *in, float *out, bf16x4 oh, bf16x4 ol, int passes,
                                              int wave, int lane)
{
  constexpr int NN = N * N, N3 = NN * N, NB = (NN + 15) / 16;
  const int r16 = lane & 15, kg = lane >> 4;
  for (int p = wave; p < E * NB; p += 4) {
    const int e = p / NB, rb = p - e * NB;
    const float *ti = in + e * N3;
    float *to = out + e * N3;
    const int r = rb * 16 + r16, ra = r / N, rr = r - ra * N;
    bf16x4 xh, xl;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int m = 4 * kg + i;
      const float v = (r < NN && m < N) ? ti[line_at<N, DIR>(ra, rr, m)] : 0.f;
      xh[i] = bf16_bits(v);
      xl[i] = bf16_bits(v - bf16_value(xh[i]));
    }
    floatx4 c = {0.f, 0.f, 0.f, 0.f};
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xh, oh, c, 0, 0, 0);
    if (passes > 1) {
      c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xl, oh, c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xh, ol, c, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int r2 = rb * 16 + 4 * kg + i, a2 = r2 / N, b2 = r2 - a2 * N;
      if (r2 < NN && r16 < N) {
        float *o = &to[line_at<N, DIR>(a2, b2, r16)];
        *o = ACC ? *o + c[i] : c[i];
      }
    }
  }
}

// One item: up to item_elems(N, PACK) consecutive elements of the order-N bucket.  Loader groups
// are the whole workgroup (PACK_BLOCK) or each wavefront for its own elements (PACK_WAVE).
template <int N, typename T, int PACK>
__device__ __forceinline__ void batch_item(const BArgs &g, const int4 it, const T *__restrict__ tab,
                                           const T *__restrict__ x, T *__restrict__ y, char *smem)
{
  constexpr int NN = N * N, N3 = NN * N, E = item_elems(N, PACK);
  constexpr bool WS = PACK == PACK_WAVE;
  constexpr int EW = WS ? E / 4 : E, GT = WS ? 64 : kBT, NS = (EW * N3 + GT - 1) / GT;
  double *VX = (double *)smem;
  T *TA = (T *)(VX + E * 32), *TGX = TA + E * N3, *TGY = TGX + E * N3;
  int *IDX = (int *)(TGY + E * N3);
  // the thread index is laundered so that nothing derived from it is hoisted out of the item loop
  // (each order would keep its own copies live across every other order's code)
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int grp = WS ? tid / 64 : 0, gt = WS ? tid % 64 : tid, e0 = grp * EW;
  const int first = it.y, count = it.z, excl = it.w;
  const int *__restrict__ gidx = g.gidx + g.goff[N] + (long)first * N3;
  const int *__restrict__ ev = g.everts + g.eoff[N] + (long)first * 8;

  // gather: node indices and values (coalesced over the element node lists), then vertices
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int t = s * GT + gt;
    if (t < EW * N3) {
      const int ts = e0 * N3 + t, l = ts / N3;
      const int c = l < count ? gidx[ts] : -1;
      IDX[ts] = c;
      TA[ts] = c >= 0 ? x[c] : T(0);
    }
  }
  for (int t = gt; t < EW * 8; t += GT) {
    const int l = e0 + (t >> 3), vt = t & 7;
    double *dst = &VX[(l * 8 + vt) * 4];
    if (l < count) {
      const long vi = ev[l * 8 + vt];
      dst[0] = g.vxyz[3 * vi];
      dst[1] = g.vxyz[3 * vi + 1];
      dst[2] = g.vxyz[3 * vi + 2];
      dst[3] = g.kv[vi];
    } else {  // unit cube for empty slots: finite geometry, nothing is scattered
      dst[0] = vt >> 2;
      dst[1] = (vt >> 1) & 1;
      dst[2] = vt & 1;
      dst[3] = 1.0;
    }
  }
  tile_sync<WS>();

  // line of this thread: lanes past the group's lines repeat lines of their own wavefront
  int line;
  if constexpr (WS) {
    constexpr int LW = EW * NN;
    line = grp * LW + gt % LW;
  } else {
    constexpr int L = E * NN, L0 = ((L - 1) / 64) * 64;
    line = tid < L ? tid : L0 + (tid - L0) % (L - L0);
  }
  int lb = line;  // derived values are formed here, not hoisted above the gather
  asm volatile("" : "+v"(lb));
  const int le = lb / NN, ab = lb - le * NN, a = ab / N, b = ab - a * N;
  T *A = &TA[le * N3], *GX = &TGX[le * N3], *GY = &TGY[le * N3];
  typedef const __attribute__((address_space(4))) double ConstD;
  const double *wl = g.wphi + wphi_off(N);
  const ConstD *wq = (const ConstD *)sgpr_fence(wl), *phq = wq + N;
  const T wij = (T)(wl[a] * wl[b]);

  T xq[N], hx[N], hy[N], hz[N];
  const int wave = tid >> 6, lane = tid & 63;
  const float *fop = g.ftab + ftab_off(N);
  if constexpr (PACK == PACK_MFMA) {
    // interpolation along z, y, x, then the collocated derivatives along x and y into GX and GY, and along z in
    // place (the linear operator needs no values at the points)
    bf16x4 oh, ol;
    mfma_operator<N>(fop, lane, oh, ol);
    mfma_contract<N, DIR_Z, false, E>(TA, TA, oh, ol, g.passes, wave, lane);
    __syncthreads();
    mfma_contract<N, DIR_Y, false, E>(TA, TA, oh, ol, g.passes, wave, lane);
    __syncthreads();
    mfma_contract<N, DIR_X, false, E>(TA, TA, oh, ol, g.passes, wave, lane);
    __syncthreads();
    mfma_operator<N>(fop + NN, lane, oh, ol);
    mfma_contract<N, DIR_X, false, E>(TA, TGX, oh, ol, g.passes, wave, lane);
    mfma_contract<N, DIR_Y, false, E>(TA, TGY, oh, ol, g.passes, wave, lane);
    __syncthreads();
    mfma_contract<N, DIR_Z, false, E>(TA, TA, oh, ol, g.passes, wave, lane);
    __syncthreads();
    ld_line<N, DIR_Z>(GX, a, b, hx);
    ld_line<N, DIR_Z>(GY, a, b, hy);
    ld_line<N, DIR_Z>(A, a, b, hz);
  } else {
    forward<N, T, WS>(tab + tab_off(N), A, GX, GY, a, b, xq, hx, hy, hz);
  }
  double Pc[2][3], Kc[2];
  const ColGeo<T> cg = column_geometry<N, T>(&VX[le * 32], wl + N, a, b, Pc, Kc);
  T vq[N], d0[N], d1[N], d2[N];
#pragma unroll
  for (int k = 0; k < N; k++) {
    T r0[3], r1[3], r2[3], det;
    cg.eval((T)phq[k], r0, r1, r2, det);
    const T kq = (T)(Kc[0] + (Kc[1] - Kc[0]) * phq[k]);
    // Dv = jw k grad(x) = w k (sum_m ghat_m r_m) / det;  df_m = jinv[m][:] . Dv = r_m . Dv / det
    const T sc = wij * (T)wq[k] * kq * recip(det);
    const T a0 = hx[k], a1 = hy[k], a2 = hz[k];
    T s3[3];
#pragma unroll
    for (int d = 0; d < 3; d++) s3[d] = a0 * r0[d] + a1 * r1[d] + a2 * r2[d];
    vq[k] = T(0);
    d0[k] = sc * (r0[0] * s3[0] + r0[1] * s3[1] + r0[2] * s3[2]);
    d1[k] = sc * (r1[0] * s3[0] + r1[1] * s3[1] + r1[2] * s3[2]);
    d2[k] = sc * (r2[0] * s3[0] + r2[1] * s3[1] + r2[2] * s3[2]);
  }
  if constexpr (PACK == PACK_MFMA) {
    // fluxes back into the tiles (each thread its own z line), then Dh^T along z in place, x and y added, and
    // B^T along x, y, z
    st_line<N, DIR_Z>(GX, a, b, d0);
    st_line<N, DIR_Z>(GY, a, b, d1);
    st_line<N, DIR_Z>(A, a, b, d2);
    __syncthreads();
    bf16x4 oh, ol;
    mfma_operator<N>(fop + 3 * NN, lane, oh, ol);
    mfma_contract<N, DIR_Z, false, E>(TA, TA, oh, ol, g.passes, wave, lane);
    __syncthreads();
    mfma_contract<N, DIR_X, true, E>(TGX, TA, oh, ol, g.passes, wave, lane);
    __syncthreads();
    mfma_contract<N, DIR_Y, true, E>(TGY, TA, oh, ol, g.passes, wave, lane);
    __syncthreads();
    mfma_operator<N>(fop + 2 * NN, lane, oh, ol);
    mfma_contract<N, DIR_X, false, E>(TA, TA, oh, ol, g.passes, wave, lane);
    __syncthreads();
    mfma_contract<N, DIR_Y, false, E>(TA, TA, oh, ol, g.passes, wave, lane);
    __syncthreads();
    mfma_contract<N, DIR_Z, false, E>(TA, TA, oh, ol, g.passes, wave, lane);
    __syncthreads();
    (void)vq;
  } else {
    backward<N, T, WS>(tab + tab_off(N), A, GX, GY, a, b, vq, d0, d1, d2);
  }

  // scatter: atomics into block spaces, plain stores into element-local slices.  Nodes shared by
  // consecutive elements of the group (z neighbours in a block space) are folded into the later
  // element's slot first, and the slots walk (a, b, element, c), so the atomics cover runs of
  // consecutive, distinct nodes.
  int tl = gt;
  asm volatile("" : "+v"(tl));
  if constexpr (EW > 1) {
    if (!excl) {
      for (int t = tl; t < (EW - 1) * NN; t += GT) {
        const int le2 = e0 + t / NN, ab2 = t % NN, top = le2 * N3 + ab2 * N + N - 1, bot = top + N3 - N + 1;
        const int ct = IDX[top];
        if (ct >= 0 && ct == IDX[bot]) {
          TA[bot] += TA[top];
          IDX[top] = -2;
        }
      }
    }
    tile_sync<WS>();
  }
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int t = s * GT + tl;
    if (t < EW * N3) {
      const int ts = e0 * N3 + brick_slot<N, EW>(t), c = IDX[ts];
      if (c >= 0) {
        const T v = TA[ts];
        if (excl) y[c] = v;
        else atomicAdd(&y[c], v);
      }
    }
  }
}

// Persistent batch kernel: every workgroup takes items from the shared counter until the list is
// exhausted (every workgroup reaches the end).
template <typename T>
__global__ void __launch_bounds__(kBT) batch_kernel(BArgs g, const T *__restrict__ tab, const T *__restrict__ x,
                                                    T *__restrict__ y)
{
  extern __shared__ __attribute__((aligned(16))) char smem[];
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(g.counter, 1);
    __syncthreads();
    const int i = s_item;
    if (i >= g.nitems) break;
    const int4 it = g.items[i];
    switch (it.x) {
#define DOHP_BCASE(NN_)                                                                          \
  case cfg_id(NN_, PACK_BLOCK): batch_item<NN_, T, PACK_BLOCK>(g, it, tab, x, y, smem); break;
#define DOHP_BCASE2(NN_)                                                                         \
  DOHP_BCASE(NN_)                                                                                \
  case cfg_id(NN_, PACK_WAVE): batch_item<NN_, T, PACK_WAVE>(g, it, tab, x, y, smem); break;
#define DOHP_BCASEM(NN_)                                                                         \
  case cfg_id(NN_, PACK_MFMA):                                                                   \
    if constexpr (sizeof(T) == 4) batch_item<NN_, T, PACK_MFMA>(g, it, tab, x, y, smem);        \
    break;
      DOHP_BCASE2(2)
      DOHP_BCASE2(3)
      DOHP_BCASE2(4)
      DOHP_BCASE2(5)
      DOHP_BCASE2(6)
      DOHP_BCASE2(7)
      DOHP_BCASE2(8)
      DOHP_BCASE(9)
      DOHP_BCASE(10)
      DOHP_BCASE(11)
      DOHP_BCASEM(5)
      DOHP_BCASEM(6)
      DOHP_BCASEM(7)
      DOHP_BCASEM(8)
      DOHP_BCASEM(9)
      DOHP_BCASEM(10)
      DOHP_BCASEM(11)
#undef DOHP_BCASEM
#undef DOHP_BCASE2
#undef DOHP_BCASE
      default: break;
    }
    __syncthreads();  // tiles and the item slot are reused
  }
}

// Clears the block-space prefix of y and resets the work counter.
template <typename T>
__global__ void batch_prep_kernel(T *__restrict__ y, long n, int *__restrict__ counter)
{
  const long i0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i0 == 0) *counter = 0;
  for (long i = i0; i < n; i += (long)gridDim.x * blockDim.x) y[i] = T(0);
}

struct SpaceSeg {
  int n;
  long nowned, nvert;
  int nelem;
  std::vector<int> elem_nodes, elem_verts;
  std::vector<double> vcoords, kv;
};
struct ElemSeg {
  int nelem;
  std::vector<int> n;               // nodes per direction of each element
  std::vector<double> vcoords, kv;  // [nelem][8][3], [nelem][8]
  long nexp;
};

}  // namespace

struct _p_dohpBatch {
  std::vector<int> seg_kind;  // 0 = block space, 1 = elements (index into the lists below)
  std::vector<int> seg_idx;
  std::vector<SpaceSeg> spaces;
  std::vector<ElemSeg> elems;
  bool ready = false;
  std::vector<long> seg_off, seg_len;
  long n = 0, nzero = 0;
  long bucket_n[kNMax + 1] = {}, bucket_excl0[kNMax + 1] = {};  // elements, first element-segment element
  BArgs args{};
  DevBuf<int> gidx, everts, counter;
  DevBuf<double> vxyz, kv, wphi, tab64;
  DevBuf<float> tab32;
  // per precision: tile per order (-1 = not tuned), measured ms per tile, item list
  int pack[2][kNMax + 1];
  float pack_ms[2][kNMax + 1][kPacks];
  int mfma = 0;  // 0 = VALU contractions only; 1 or 3 = MFMA tiles (BF16 passes) offered to the FP32 tuner
  int forced[2][kNMax + 1];  // tile fixed by the caller, -1 = tuned
  DevBuf<float> ftab;
  DevBuf<int4> items[2];
  int nitems[2] = {0, 0};
  size_t lds[2] = {0, 0};
  std::map<std::pair<int, size_t>, long> grid[2];  // persistent grid per (device, dynamic LDS bytes)
  _p_dohpBatch()
  {
    for (int p = 0; p < 2; p++)
      for (int N = 0; N <= kNMax; N++) {
        pack[p][N] = forced[p][N] = -1;
        for (int k = 0; k < kPacks; k++) pack_ms[p][N][k] = -1.f;
      }
  }
};

namespace {

dErr batch_setup(_p_dohpBatch *b)
{
  if (b->ready) return 0;
  const int nseg = (int)b->seg_kind.size();
  b->seg_off.assign(nseg, 0);
  b->seg_len.assign(nseg, 0);
  long off = 0;
  for (int pass = 0; pass < 2; pass++)
    for (int s = 0; s < nseg; s++) {
      if (b->seg_kind[s] != pass) continue;
      const long len = pass == 0 ? b->spaces[b->seg_idx[s]].nowned : b->elems[b->seg_idx[s]].nexp;
      b->seg_off[s] = off;
      b->seg_len[s] = len;
      off += len;
      if (pass == 0) b->nzero = off;
    }
  b->n = off;
  DOHP_CHECK(b->n < (1L << 31), dohpERR_ARG, "batch exceeds 2^31 packed unknowns");
  // vertices: spaces, then element segments
  std::vector<double> vx, kv;
  std::vector<long> vbase(nseg, 0);
  for (int s = 0; s < nseg; s++) {
    vbase[s] = (long)kv.size();
    if (b->seg_kind[s] == 0) {
      const SpaceSeg &S = b->spaces[b->seg_idx[s]];
      vx.insert(vx.end(), S.vcoords.begin(), S.vcoords.end());
      kv.insert(kv.end(), S.kv.begin(), S.kv.end());
    } else {
      const ElemSeg &S = b->elems[b->seg_idx[s]];
      vx.insert(vx.end(), S.vcoords.begin(), S.vcoords.end());
      kv.insert(kv.end(), S.kv.begin(), S.kv.end());
    }
  }
  DOHP_CHECK(kv.size() < (1UL << 31), dohpERR_ARG, "batch exceeds 2^31 vertices");
  // buckets by order: block-space elements, then element-segment elements
  std::vector<int> gidx, ev;
  for (int N = kNMin; N <= kNMax; N++) {
    const int N3 = N * N * N;
    b->args.goff[N] = (long)gidx.size();
    b->args.eoff[N] = (long)ev.size();
    long cnt = 0;
    for (int pass = 0; pass < 2; pass++) {
      if (pass == 1) b->bucket_excl0[N] = cnt;
      for (int s = 0; s < nseg; s++) {
        if (b->seg_kind[s] != pass) continue;
        const long so = b->seg_off[s];
        if (pass == 0) {
          const SpaceSeg &S = b->spaces[b->seg_idx[s]];
          if (S.n != N) continue;
          for (int e = 0; e < S.nelem; e++) {
            for (int i = 0; i < N3; i++) {
              const int c = S.elem_nodes[(long)e * N3 + i];
              gidx.push_back(c < S.nowned ? (int)(so + c) : -1);
            }
            for (int t = 0; t < 8; t++) ev.push_back((int)(vbase[s] + S.elem_verts[e * 8 + t]));
            cnt++;
          }
        } else {
          const ElemSeg &S = b->elems[b->seg_idx[s]];
          long eo = 0;
          for (int e = 0; e < S.nelem; e++) {
            const int n = S.n[e], n3 = n * n * n;
            if (n == N) {
              for (int i = 0; i < n3; i++) gidx.push_back((int)(so + eo + i));
              for (int t = 0; t < 8; t++) ev.push_back((int)(vbase[s] + e * 8 + t));
              cnt++;
            }
            eo += n3;
          }
        }
      }
    }
    b->bucket_n[N] = cnt;
  }
  // tables of every order
  std::vector<double> t64(tab_off(kNMax + 1), 0.0), wp(wphi_off(kNMax + 1), 0.0);
  for (int N = kNMin; N <= kNMax; N++) {
    double z[kMaxN], w[kMaxN], zl[kMaxN], wl[kMaxN];
    gauss_legendre(N, z, w);
    gauss_lobatto(N, zl, wl);
    std::vector<double> B(N * N), Dh(N * N);
    lagrange_matrix(N, zl, N, z, B.data());
    collocation_derivative(N, z, Dh.data());
    const std::vector<double> t = eo_tables(N, B, Dh);
    std::copy(t.begin(), t.end(), t64.begin() + tab_off(N));
    for (int q = 0; q < N; q++) {
      wp[wphi_off(N) + q] = w[q];
      wp[wphi_off(N) + N + q] = 0.5 * (1.0 + z[q]);
    }
  }
  std::vector<float> t32(t64.begin(), t64.end()), tf(ftab_off(kNMax + 1), 0.f);
  for (int N = kNMin; N <= kNMax; N++) {
    double z[kMaxN], w[kMaxN], zl[kMaxN], wl[kMaxN];
    gauss_legendre(N, z, w);
    gauss_lobatto(N, zl, wl);
    std::vector<double> B(N * N), Dh(N * N);
    lagrange_matrix(N, zl, N, z, B.data());
    collocation_derivative(N, z, Dh.data());
    float *f = &tf[ftab_off(N)];
    for (int q = 0; q < N; q++)
      for (int m = 0; m < N; m++) {
        f[q * N + m] = (float)B[q * N + m];
        f[N * N + q * N + m] = (float)Dh[q * N + m];
        f[2 * N * N + q * N + m] = (float)B[m * N + q];
        f[3 * N * N + q * N + m] = (float)Dh[m * N + q];
      }
  }
  DOHP_CALL(b->ftab.upload(tf.data(), tf.size()));
  b->args.ftab = b->ftab.ptr;
  DOHP_CALL(b->gidx.upload(gidx.data(), gidx.size()));
  DOHP_CALL(b->everts.upload(ev.data(), ev.size()));
  DOHP_CALL(b->vxyz.upload(vx.data(), vx.size()));
  DOHP_CALL(b->kv.upload(kv.data(), kv.size()));
  DOHP_CALL(b->wphi.upload(wp.data(), wp.size()));
  DOHP_CALL(b->tab64.upload(t64.data(), t64.size()));
  DOHP_CALL(b->tab32.upload(t32.data(), t32.size()));
  DOHP_CALL(b->counter.alloc(1));
  b->args.gidx = b->gidx.ptr;
  b->args.everts = b->everts.ptr;
  b->args.vxyz = b->vxyz.ptr;
  b->args.kv = b->kv.ptr;
  b->args.wphi = b->wphi.ptr;
  b->args.counter = b->counter.ptr;
  b->ready = true;
  return 0;
}

// Items of the orders with pack[N] >= 0, highest order first; each item holds elements of one
// store kind.
std::vector<int4> build_items(const _p_dohpBatch *b, const int *pack)
{
  std::vector<int4> it;
  for (int N = kNMax; N >= kNMin; N--) {
    if (pack[N] < 0 || b->bucket_n[N] == 0) continue;
    const int E = item_elems(N, pack[N]);
    const long ranges[2][2] = {{0, b->bucket_excl0[N]}, {b->bucket_excl0[N], b->bucket_n[N]}};
    for (int k = 0; k < 2; k++)
      for (long f = ranges[k][0]; f < ranges[k][1]; f += E)
        it.push_back(make_int4(cfg_id(N, pack[N]), (int)f, (int)std::min<long>(E, ranges[k][1] - f), k));
  }
  return it;
}

template <typename T>
const T *tabs(const _p_dohpBatch *b);
template <>
const double *tabs<double>(const _p_dohpBatch *b) { return b->tab64.ptr; }
template <>
const float *tabs<float>(const _p_dohpBatch *b) { return b->tab32.ptr; }

template <typename T>
dErr batch_launch(_p_dohpBatch *b, const int4 *items, int nitems, size_t lds, const T *x, T *y, hipStream_t st)
{
  hipLaunchKernelGGL(batch_prep_kernel<T>, dim3(b->nzero > 0 ? std::min<long>((b->nzero + 255) / 256, 2048) : 1),
                     dim3(256), 0, st, y, b->nzero, b->counter.ptr);
  if (nitems == 0) return 0;
  int dev = 0, cus = 0;
  DOHP_CALL(current_device_cus(dev, cus));
  DOHP_CHECK(lds <= 160 * 1024, dohpERR_ARG, "batch tile exceeds the workgroup LDS limit");
  // (the kernel's dynamic-LDS limit is shared by every batch object: set for this launch's size)
  DOHP_HIP(hipFuncSetAttribute((const void *)batch_kernel<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  // grid per (device, dynamic LDS), asked once: the launch shape does not change between applications
  long &grid_cached = b->grid[sizeof(T) == 4][{dev, lds}];
  if (!grid_cached) {
    int per_cu = 0;
    DOHP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel<T>, kBT, lds));
    grid_cached = (long)cus * std::max(per_cu, 1);
  }
  const long grid = std::min<long>(nitems, grid_cached);
  BArgs a = b->args;
  a.items = items;
  a.nitems = nitems;
  a.passes = b->mfma > 1 ? 3 : 1;
  hipLaunchKernelGGL(batch_kernel<T>, dim3(grid), dim3(kBT), lds, st, a, tabs<T>(b), x, y);
  DOHP_HIP(hipGetLastError());
  return 0;
}

// First application in a precision: time both tiles of every order present (three launches each,
// the first a warm-up) on the caller's output, keep the faster, build the work list.
template <typename T>
dErr batch_tune(_p_dohpBatch *b, const T *x, T *y, hipStream_t st)
{
  const int pr = sizeof(T) == 4;
  if (b->nitems[pr] > 0 || b->items[pr].ptr) return 0;
  hipEvent_t e0, e1;
  DOHP_HIP(hipEventCreate(&e0));
  DOHP_HIP(hipEventCreate(&e1));
  dErr err = 0;
  for (int N = kNMin; N <= kNMax && !err; N++) {
    if (b->bucket_n[N] == 0) continue;
    float best = 1e30f;
    if (b->forced[pr][N] >= 0) {
      DOHP_CHECK(pack_valid(N, b->forced[pr][N], sizeof(T), b->mfma), dohpERR_ARG,
                 "tile not available for degree " + std::to_string(N - 1));
      b->pack[pr][N] = b->forced[pr][N];
      continue;
    }
    for (int p = 0; p < kPacks && !err; p++) {
      if (!pack_valid(N, p, sizeof(T), b->mfma)) continue;
      int one[kNMax + 1];
      for (int m = 0; m <= kNMax; m++) one[m] = m == N ? p : -1;
      const std::vector<int4> it = build_items(b, one);
      DevBuf<int4> dit;
      err = dit.upload(it.data(), it.size());
      float ms = 0.f;
      for (int r = 0; r < 3 && !err; r++) {
        if (hipEventRecord(e0, st) != hipSuccess) err = dohpERR_HIP;
        if (!err) err = batch_launch<T>(b, dit.ptr, (int)it.size(), cfg_lds(N, p, sizeof(T)), x, y, st);
        if (!err && hipEventRecord(e1, st) != hipSuccess) err = dohpERR_HIP;
        if (!err && hipEventSynchronize(e1) != hipSuccess) err = dohpERR_HIP;
        float t = 0.f;
        if (!err && hipEventElapsedTime(&t, e0, e1) != hipSuccess) err = dohpERR_HIP;
        if (r > 0) ms += t / 2;
      }
      if (err) break;
      b->pack_ms[pr][N][p] = ms;
      if (ms < best) {
        best = ms;
        b->pack[pr][N] = p;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (err == dohpERR_HIP) set_error("batch tuning: HIP error");
  if (err) return err;
  const std::vector<int4> it = build_items(b, b->pack[pr]);
  size_t lds = 0;
  for (int N = kNMin; N <= kNMax; N++)
    if (b->pack[pr][N] >= 0 && b->bucket_n[N] > 0) lds = std::max(lds, cfg_lds(N, b->pack[pr][N], sizeof(T)));
  DOHP_CALL(b->items[pr].upload(it.data(), it.size()));
  b->nitems[pr] = (int)it.size();
  b->lds[pr] = lds;
  return 0;
}

template <typename T>
dErr batch_mult(_p_dohpBatch *b, const T *x, T *y, void *stream)
{
  DOHP_CHECK(b && x && y, dohpERR_ARG, "null argument");
  DOHP_CALL(batch_setup(b));
  hipStream_t st = (hipStream_t)stream;
  DOHP_CALL(batch_tune<T>(b, x, y, st));
  const int pr = sizeof(T) == 4;
  return batch_launch<T>(b, b->items[pr].ptr, b->nitems[pr], b->lds[pr], x, y, st);
}

}  // namespace

extern "C" dErr dohpBatchCreate(dohpBatch *out)
{
  DOHP_CHECK(out, dohpERR_ARG, "null argument");
  *out = new _p_dohpBatch;
  return 0;
}

extern "C" dErr dohpBatchDestroy(dohpBatch *b)
{
  if (b && *b) {
    delete *b;
    *b = nullptr;
  }
  return 0;
}

extern "C" dErr dohpBatchAddSpace(dohpBatch b, dohpFS fs, const double *kvert, int *segment)
{
  DOHP_CHECK(b && fs, dohpERR_ARG, "null argument");
  DOHP_CHECK(!b->ready, dohpERR_ARG, "batch already set up");
  DOHP_CHECK(fs->n >= kNMin && fs->n <= kNMax, dohpERR_ARG, "batch supports degree 1..10");
  SpaceSeg S;
  S.n = fs->n;
  S.nowned = fs->nowned;
  S.nvert = fs->nvert;
  S.nelem = fs->nelem;
  S.elem_nodes = fs->elem_nodes;
  S.elem_verts = fs->elem_verts;
  S.vcoords = fs->vcoords;
  if (kvert) S.kv.assign(kvert, kvert + fs->nvert);
  else S.kv.assign(fs->nvert, 1.0);
  b->seg_kind.push_back(0);
  b->seg_idx.push_back((int)b->spaces.size());
  b->spaces.push_back(std::move(S));
  if (segment) *segment = (int)b->seg_kind.size() - 1;
  return 0;
}

extern "C" dErr dohpBatchAddElements(dohpBatch b, int nelem, const int *degree, const double *elem_vx,
                                     const double *kvert, int *segment)
{
  DOHP_CHECK(b && degree && elem_vx && nelem >= 0, dohpERR_ARG, "bad argument");
  DOHP_CHECK(!b->ready, dohpERR_ARG, "batch already set up");
  ElemSeg S;
  S.nelem = nelem;
  S.n.resize(nelem);
  S.nexp = 0;
  for (int e = 0; e < nelem; e++) {
    DOHP_CHECK(degree[e] >= 1 && degree[e] <= 10, dohpERR_ARG, "batch supports degree 1..10");
    S.n[e] = degree[e] + 1;
    S.nexp += (long)S.n[e] * S.n[e] * S.n[e];
  }
  S.vcoords.assign(elem_vx, elem_vx + 24L * nelem);
  if (kvert) S.kv.assign(kvert, kvert + 8L * nelem);
  else S.kv.assign(8L * nelem, 1.0);
  // geometry validation at the Gauss points of each element's rule (jacobi.c:397)
  std::vector<double> md(nelem);
  {
    DevBuf<int> dv;
    DevBuf<double> dx, dz, dm;
    std::vector<int> idv(8L * nelem);
    for (long i = 0; i < 8L * nelem; i++) idv[i] = (int)i;
    DOHP_CALL(dv.upload(idv.data(), idv.size()));
    DOHP_CALL(dx.upload(S.vcoords.data(), S.vcoords.size()));
    DOHP_CALL(dm.alloc(nelem > 0 ? nelem : 1));
    for (int N = kNMin; N <= kNMax; N++) {
      std::vector<int> sel;
      for (int e = 0; e < nelem; e++)
        if (S.n[e] == N) sel.push_back(e);
      if (sel.empty()) continue;
      double z[kMaxN], w[kMaxN];
      gauss_legendre(N, z, w);
      DOHP_CALL(dz.upload(z, N));
      hipLaunchKernelGGL(min_det_kernel, dim3((nelem + 127) / 128), dim3(128), 0, 0, N, dv.ptr, dx.ptr, (long)nelem,
                         dz.ptr, dm.ptr);
      std::vector<double> h(nelem);
      DOHP_HIP(hipMemcpy(h.data(), dm.ptr, sizeof(double) * nelem, hipMemcpyDeviceToHost));
      for (int e : sel) md[e] = h[e];
    }
  }
  for (int e = 0; e < nelem; e++)
    DOHP_CHECK(md[e] >= 0.0, dohpERR_NEGATIVE_JACOBIAN, "Negative Jacobian at element " + std::to_string(e));
  b->seg_kind.push_back(1);
  b->seg_idx.push_back((int)b->elems.size());
  b->elems.push_back(std::move(S));
  if (segment) *segment = (int)b->seg_kind.size() - 1;
  return 0;
}

extern "C" dErr dohpBatchSetUp(dohpBatch b) { DOHP_CHECK(b, dohpERR_ARG, "null batch"); return batch_setup(b); }

extern "C" dErr dohpBatchGetSizes(dohpBatch b, long *n, int *nsegments)
{
  DOHP_CHECK(b, dohpERR_ARG, "null batch");
  DOHP_CALL(batch_setup(b));
  if (n) *n = b->n;
  if (nsegments) *nsegments = (int)b->seg_kind.size();
  return 0;
}

extern "C" dErr dohpBatchGetSegment(dohpBatch b, int segment, long *offset, long *length)
{
  DOHP_CHECK(b, dohpERR_ARG, "null batch");
  DOHP_CALL(batch_setup(b));
  DOHP_CHECK(segment >= 0 && segment < (int)b->seg_kind.size(), dohpERR_ARG, "no such segment");
  if (offset) *offset = b->seg_off[segment];
  if (length) *length = b->seg_len[segment];
  return 0;
}

extern "C" dErr dohpBatchMult(dohpBatch b, const double *x, double *y, void *stream)
{
  return batch_mult<double>(b, x, y, stream);
}

extern "C" dErr dohpBatchMultF32(dohpBatch b, const float *x, float *y, void *stream)
{
  return batch_mult<float>(b, x, y, stream);
}

extern "C" dErr dohpBatchSetMFMA(dohpBatch b, int passes)
{
  DOHP_CHECK(b && (passes == 0 || passes == 1 || passes == 3), dohpERR_ARG, "MFMA passes must be 0, 1 or 3");
  if (b->mfma != passes) {  // the FP32 tiles are tuned again on the next application
    b->mfma = passes;
    b->items[1].release();
    b->nitems[1] = 0;
    for (int N = 0; N <= kNMax; N++) {
      b->pack[1][N] = -1;
      for (int k = 0; k < kPacks; k++) b->pack_ms[1][N][k] = -1.f;
    }
  }
  return 0;
}

extern "C" dErr dohpBatchSetTile(dohpBatch b, int degree, int fp32, int tile)
{
  DOHP_CHECK(b && degree >= 1 && degree <= 10 && (fp32 == 0 || fp32 == 1) && tile >= -1 && tile < kPacks,
             dohpERR_ARG, "bad argument");
  b->forced[fp32][degree + 1] = tile;
  b->items[fp32].release();
  b->nitems[fp32] = 0;
  b->pack[fp32][degree + 1] = -1;
  return 0;
}

extern "C" dErr dohpBatchGetTuning(dohpBatch b, int degree, int fp32, int *tile, float *ms)
{
  DOHP_CHECK(b && degree >= 1 && degree <= 10 && (fp32 == 0 || fp32 == 1), dohpERR_ARG, "bad argument");
  const int N = degree + 1;
  if (tile) *tile = b->pack[fp32][N];
  if (ms) {
    for (int k = 0; k < kPacks; k++) ms[k] = b->pack_ms[fp32][N][k];
  }
  return 0;
}

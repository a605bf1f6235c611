// Device building blocks of the fused element kernels, shared by the single-operator kernels
// (ellip.hip) and the batched mixed-order kernel (batch.hip): the even-odd 1-D operators, the
// line-per-thread contractions through LDS, the column geometry of the trilinear map and the
// pointwise helpers.  Internal header: everything has internal linkage.
#pragma once

#include <cmath>
#include <vector>

#include "common.h"

namespace {

// Element kernel modes: linear operator, residual without its forcing term, Jacobian action, the
// forcing term alone (assembled once per operator: it does not depend on the state), the per-point
// quadrature data of the Jacobian (MODE_QD, written once per linearization) and the Jacobian action
// (or linear operator) that reads it (MODE_JQD).
enum { MODE_LIN = 0, MODE_FUN = 1, MODE_JAC = 2, MODE_FRC = 3, MODE_QD = 4, MODE_JQD = 5 };
constexpr int kModes = 6;
// launch shapes of the fused kernels the tuner chooses from (four-wavefront, one-wavefront, their
// persistent forms, and the four-wavefront shape with the contractions on the matrix cores)
constexpr int kShapes = 5;


struct Phys {
  double p, eps2, lambda, ea, eb, ec;
  int exact;
  int has_k;
  int qpow;  // 4 (p-2)/2 when it is an integer in [-8, 8], else 99 (general power)
};

// Workgroup shapes the tuner chooses from: elements per workgroup filling four wavefronts, or one
// wavefront (one element per wavefront from n = 6 up; two wavefronts for n >= 9).
constexpr int epb_wide(int N) { return (256 / (N * N)) > 0 ? 256 / (N * N) : 1; }
constexpr int epb_narrow(int N) { return (64 / (N * N)) > 0 ? 64 / (N * N) : 1; }

// Exact-solution forcing f = -div(eta grad u) (ellip.c:42-117), from u's gradient and Hessian.
__device__ double forcing(const Phys &ph, double x, double y, double z)
{
  const double a = ph.ea, b = ph.eb, c = ph.ec;
  double du[3] = {0, 0, 0}, H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  switch (ph.exact) {
  case 0: {
    const double cx = cos(a * x), sx = sin(a * x), ey = exp(b * y), cz = cos(c * z), sz = sin(c * z);
    du[0] = -a * sx * ey * sz; du[1] = b * cx * ey * sz; du[2] = c * cx * ey * cz;
    H[0] = -a * a * cx * ey * sz; H[4] = b * b * cx * ey * sz; H[8] = -c * c * cx * ey * sz;
    H[1] = H[3] = -a * b * sx * ey * sz; H[2] = H[6] = -a * c * sx * ey * cz; H[5] = H[7] = b * c * cx * ey * cz;
  } break;
  case 1:
    du[0] = 2 * a * x; du[1] = 2 * b * y; du[2] = -2 * c * z;
    H[0] = 2 * a; H[4] = 2 * b; H[8] = -2 * c;
    break;
  case 2: du[0] = a; du[1] = b; du[2] = c; break;
  case 3:
    du[0] = 3 * a * x * x + c * (1 - z * z); du[1] = 2 * b * y * z; du[2] = b * y * y - 2 * c * z * x;
    H[0] = 6 * a * x; H[4] = 2 * b * z; H[8] = -2 * c * x;
    H[2] = H[6] = -2 * c * z; H[5] = H[7] = 2 * b * y;
    break;
  case 4: {
    const double pi = M_PI;
    return 3 * pi * pi * sin(pi * x) * sin(pi * y) * sin(pi * z);
  }
  default: return 0.0;
  }
  const double e2g = ph.eps2 + 0.5 * (du[0] * du[0] + du[1] * du[1] + du[2] * du[2]);
  const double eta = pow(e2g, (ph.p - 2) / 2);
  double quad = 0;
  for (int r = 0; r < 3; r++)
    for (int s = 0; s < 3; s++) quad += du[r] * H[r * 3 + s] * du[s];
  return -eta * (H[0] + H[4] + H[8]) - (ph.p - 2) / 2 * pow(e2g, (ph.p - 4) / 2) * quad;
}

// Polynomial form of the column geometry: with c0, c1 linear in g = (1+z_k)/2 and c2 constant,
// r0 = c1 x c2 and r1 = c2 x c0 are linear in g, r2 = c0 x c1 and det = c0 . r0 quadratic.  The
// coefficients are formed in FP64 from the vertex coordinates, once per column; evaluating them
// costs 14 multiply-adds per Gauss point.
template <typename G>
struct ColGeo {
  G p0[3], q0[3], p1[3], q1[3], p2[3], q2[3], s2[3], e0, e1, e2;
  __device__ inline void eval(G g, G (&r0)[3], G (&r1)[3], G (&r2)[3], G &det) const
  {
#pragma unroll
    for (int d = 0; d < 3; d++) {
      r0[d] = p0[d] + g * q0[d];
      r1[d] = p1[d] + g * q1[d];
      r2[d] = p2[d] + g * (q2[d] + g * s2[d]);
    }
    det = e0 + g * (e1 + g * e2);
  }
};

__device__ inline void cross3(const double (&a)[3], const double (&b)[3], double (&c)[3])
{
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

template <typename G>
__device__ inline ColGeo<G> make_colgeo(const double (&Rc)[2][3], const double (&Sc)[2][3], const double (&c2)[3])
{
  double a0[3], b0[3], a1[3], b1[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    a0[d] = Rc[0][d];
    b0[d] = Rc[1][d] - Rc[0][d];
    a1[d] = Sc[0][d];
    b1[d] = Sc[1][d] - Sc[0][d];
  }
  double p0[3], q0[3], p1[3], q1[3], p2[3], t0[3], t1[3], s2[3];
  cross3(a1, c2, p0);
  cross3(b1, c2, q0);
  cross3(c2, a0, p1);
  cross3(c2, b0, q1);
  cross3(a0, a1, p2);
  cross3(a0, b1, t0);
  cross3(b0, a1, t1);
  cross3(b0, b1, s2);
  ColGeo<G> cg;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    cg.p0[d] = (G)p0[d]; cg.q0[d] = (G)q0[d];
    cg.p1[d] = (G)p1[d]; cg.q1[d] = (G)q1[d];
    cg.p2[d] = (G)p2[d]; cg.q2[d] = (G)(t0[d] + t1[d]); cg.s2[d] = (G)s2[d];
  }
  cg.e0 = (G)(a0[0] * p0[0] + a0[1] * p0[1] + a0[2] * p0[2]);
  cg.e1 = (G)(a0[0] * q0[0] + a0[1] * q0[1] + a0[2] * q0[2] + b0[0] * p0[0] + b0[1] * p0[1] + b0[2] * p0[2]);
  cg.e2 = (G)(b0[0] * q0[0] + b0[1] * q0[1] + b0[2] * q0[2]);
  return cg;
}

// Reciprocal of a positive determinant: hardware estimate refined by Newton steps to full
// precision (FP64) or used as is (FP32, 1 ulp).
__device__ inline double recip(double d)
{
  double x = __builtin_amdgcn_rcp(d);
  double e = fma(-d, x, 1.0);
  x = fma(x, e, x);
  e = fma(-d, x, 1.0);
  return fma(x, e, x);
}
__device__ inline float recip(float d) { return __builtin_amdgcn_rcpf(d); }

// ---------------------------------------------------------------------------------------------
// 1-D operators in even-odd form.  The Gauss points and the GLL nodes are symmetric about 0, so
// B[n-1-q][n-1-m] = B[q][m] and the collocated derivative has Dh[n-1-q][n-1-m] = -Dh[q][m] (the
// same holds for the transposes).  With e[m] = in[m] + in[n-1-m] and o[m] = in[m] - in[n-1-m]
// (m < h = n/2), rows q and n-1-q of out = A in share
//     S = sum_m P[q][m] e[m] (+ Pm[q] in[h] for odd n),   D = sum_m M[q][m] o[m],
//     out[q] = S + D,   out[n-1-q] = S - D (symmetric A) or D - S (antisymmetric A),
// with P = (A[q][m] + A[q][n-1-m]) / 2 and M = (A[q][m] - A[q][n-1-m]) / 2; the middle row of odd
// n is R . e (+ R[h] in[h]) for symmetric A and R . o for antisymmetric A.  Packed per operator:
// [P: h*h][M: h*h][Pm: h][R: h+1] (the last two only for odd n).  The coefficients are uniform
// over the workgroup and are read with scalar loads as instruction operands.
constexpr int eo_size(int N) { return 2 * (N / 2) * (N / 2) + (N & 1) * (2 * (N / 2) + 1); }
// B, the collocated derivative Dh, their transposes, and the product Dh B (derivative of the
// interpolant, one pass) with its transpose.
enum { OP_B = 0, OP_D = 1, OP_BT = 2, OP_DT = 3, OP_DB = 4, OP_DBT = 5 };
constexpr int kEoOps = 6;

// Table loads issued after this point stay after it, so each contraction holds only its own
// coefficients in scalar registers.
template <typename T>
__device__ __forceinline__ const T *sgpr_fence(const T *p)
{
  asm volatile("" : "+s"(p));
  return p;
}

template <int N, int SYM, typename T>
__device__ __forceinline__ void eo_apply(const T *__restrict__ tab, int op, const T (&in)[N], T (&out)[N])
{
  constexpr int H = N / 2, ODD = N & 1, HH = H * H;
  typedef const __attribute__((address_space(4))) T ConstT;  // constant address space: scalar loads
  const ConstT *A = (const ConstT *)(sgpr_fence(tab) + op * eo_size(N));
  T e[H > 0 ? H : 1], o[H > 0 ? H : 1];
#pragma unroll
  for (int m = 0; m < H; m++) {
    e[m] = in[m] + in[N - 1 - m];
    o[m] = in[m] - in[N - 1 - m];
  }
#pragma unroll
  for (int q = 0; q < H; q++) {
    T sv = 0, dv = 0;
#pragma unroll
    for (int m = 0; m < H; m++) {
      sv = fma(A[q * H + m], e[m], sv);
      dv = fma(A[HH + q * H + m], o[m], dv);
    }
    if (ODD) sv = fma(A[2 * HH + q], in[H], sv);
    out[q] = sv + dv;
    out[N - 1 - q] = SYM > 0 ? sv - dv : dv - sv;
  }
  if (ODD) {
    T r = 0;
#pragma unroll
    for (int m = 0; m < H; m++) r = fma(A[2 * HH + H + m], SYM > 0 ? e[m] : o[m], r);
    if (SYM > 0) r = fma(A[2 * HH + 2 * H], in[H], r);
    out[H] = r;
  }
}

// Element tiles [i][j][k] (k fastest, the reference's node order).  A thread (a, b) of an
// element owns one line per phase: a z line (i, j) = (a, b), a y line (i, k) = (a, b) or an
// x line (j, k) = (a, b).  No register value crosses from one phase to the next, so any one-to-one
// assignment of lines to threads within a phase is valid.  At n = 7 the y lines are dealt out as
// (i, k) = (b, (7 - a) mod 7): with (a, b) the lanes of a wavefront start at a * 49 + b, which
// puts several lanes of each 16 on the same LDS bank pair (8-byte slots); the permuted starts
// spread them (2.9 -> 2.0 accesses per bank per pass on 32 banks, counted on the host; the FP64
// residual and recomputing kernels at degree 6 run 1.3-1.9 % faster).  At the other orders no
// permutation of this kind measured faster, and (a, b) is kept.
enum { DIR_X = 0, DIR_Y = 1, DIR_Z = 2 };
template <int N, int DIR>
__device__ __forceinline__ int line_at(int a, int b, int s)
{
  if constexpr (DIR == DIR_Y && N == 7) return (b * N + s) * N + (a == 0 ? 0 : N - a);
  return DIR == DIR_Z ? (a * N + b) * N + s : DIR == DIR_Y ? (a * N + s) * N + b : (s * N + a) * N + b;
}
template <int N, int DIR, typename T>
__device__ __forceinline__ void ld_line(const T *tile, int a, int b, T (&v)[N])
{
#pragma unroll
  for (int s = 0; s < N; s++) v[s] = tile[line_at<N, DIR>(a, b, s)];
}
template <int N, int DIR, typename T>
__device__ __forceinline__ void st_line(T *tile, int a, int b, const T (&v)[N])
{
#pragma unroll
  for (int s = 0; s < N; s++) tile[line_at<N, DIR>(a, b, s)] = v[s];
}

// Barrier between line phases: the whole workgroup (WS = false), or only the wavefront when every
// tile it touches belongs to it (WS = true: the fences wait for its LDS accesses and keep the
// compiler from moving them across; the wavefront runs in lockstep).
template <bool WS>
__device__ __forceinline__ void tile_sync()
{
  if constexpr (WS) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  } else {
    __syncthreads();
  }
}

// Interpolation and reference gradient of the nodal tile A at the Gauss points.  On return the
// thread holds, along its z line (q1, q2) = (a, b): values q, gradients gx, gy, gz.  A is
// overwritten; GX and GY are scratch.  Every thread of the workgroup must call it.  Three line
// phases: z (B); y (B, and Dh B into GY); x (B and Dh B into GX, B of GY); the z derivative stays
// in registers.
template <int N, typename T, bool WS = false>
__device__ __forceinline__ void forward(const T *tab, T *A, T *GX, T *GY, int a, int b, T (&q)[N], T (&gx)[N],
                                        T (&gy)[N], T (&gz)[N])
{
  T in[N], out[N];
  ld_line<N, DIR_Z>(A, a, b, in);
  eo_apply<N, 1>(tab, OP_B, in, out);
  st_line<N, DIR_Z>(A, a, b, out);
  tile_sync<WS>();
  ld_line<N, DIR_Y>(A, a, b, in);
  eo_apply<N, 1>(tab, OP_B, in, out);
  st_line<N, DIR_Y>(A, a, b, out);
  eo_apply<N, -1>(tab, OP_DB, in, out);
  st_line<N, DIR_Y>(GY, a, b, out);
  tile_sync<WS>();
  ld_line<N, DIR_X>(A, a, b, in);
  eo_apply<N, 1>(tab, OP_B, in, out);
  st_line<N, DIR_X>(A, a, b, out);
  eo_apply<N, -1>(tab, OP_DB, in, out);
  st_line<N, DIR_X>(GX, a, b, out);
  ld_line<N, DIR_X>(GY, a, b, in);
  eo_apply<N, 1>(tab, OP_B, in, out);
  st_line<N, DIR_X>(GY, a, b, out);
  tile_sync<WS>();
  ld_line<N, DIR_Z>(A, a, b, q);
  eo_apply<N, -1>(tab, OP_D, q, gz);
  ld_line<N, DIR_Z>(GX, a, b, gx);
  ld_line<N, DIR_Z>(GY, a, b, gy);
}

// Transposed applies: from the pointwise values v and the mapped fluxes d0, d1, d2 along the
// thread's z line, the element's nodal contributions B^T (v + Dh_x^T d0 + Dh_y^T d1 + Dh_z^T d2)
// land in tile A.  Three line phases after the z one: x (B^T of the sum, plus (Dh B)^T of d0; B^T
// of d1), y (B^T, plus (Dh B)^T of d1), z (B^T).  Ends with a barrier.
template <int N, typename T, bool WS = false>
__device__ __forceinline__ void backward(const T *tab, T *A, T *GX, T *GY, int a, int b, const T (&v)[N],
                                         const T (&d0)[N], const T (&d1)[N], const T (&d2)[N])
{
  T w[N], in[N], out[N];
  eo_apply<N, -1>(tab, OP_DT, d2, w);
#pragma unroll
  for (int s = 0; s < N; s++) w[s] += v[s];
  st_line<N, DIR_Z>(A, a, b, w);
  st_line<N, DIR_Z>(GY, a, b, d1);
  st_line<N, DIR_Z>(GX, a, b, d0);
  tile_sync<WS>();
  ld_line<N, DIR_X>(A, a, b, in);
  eo_apply<N, 1>(tab, OP_BT, in, w);
  ld_line<N, DIR_X>(GX, a, b, in);
  eo_apply<N, -1>(tab, OP_DBT, in, out);
#pragma unroll
  for (int s = 0; s < N; s++) w[s] += out[s];
  st_line<N, DIR_X>(A, a, b, w);
  ld_line<N, DIR_X>(GY, a, b, in);
  eo_apply<N, 1>(tab, OP_BT, in, out);
  st_line<N, DIR_X>(GY, a, b, out);
  tile_sync<WS>();
  ld_line<N, DIR_Y>(A, a, b, in);
  eo_apply<N, 1>(tab, OP_BT, in, w);
  ld_line<N, DIR_Y>(GY, a, b, in);
  eo_apply<N, -1>(tab, OP_DBT, in, out);
#pragma unroll
  for (int s = 0; s < N; s++) w[s] += out[s];
  st_line<N, DIR_Y>(A, a, b, w);
  tile_sync<WS>();
  ld_line<N, DIR_Z>(A, a, b, in);
  eo_apply<N, 1>(tab, OP_BT, in, out);
  st_line<N, DIR_Z>(A, a, b, out);
  tile_sync<WS>();
}

// ---- Line passes on the matrix cores (n = 5..8) ----
//
// A one-dimensional pass applies an n x n operator (or two stacked) to every line of the
// workgroup's tiles along one direction: C[o][L] = sum_k Op[o][k] In[k][L].  Each wavefront takes
// the 64 lines of its own threads as four blocks of 16 columns of one 16x16x4 MFMA product per
// K step (v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64, the precision of the kernel, so nothing is
// rounded below it).  Operand layout of the 16x16x4 products: lane l holds A[l % 16][l / 16] of the
// operator, B[l / 16][l % 16] of the lines and, after the product, C[row(l / 16, r)][l % 16],
// r = 0..3, with row = 4 (l / 16) + r in FP32 and l / 16 + 4 r in FP64 (measured on gfx950).  Line inputs are one or two fields (K = 8 per field: positions s = 0..n-1 along the
// line, the rest zero), outputs one or two fields (rows 0..7 into the first, 8..15 into the
// second).  A pass reads and writes only its own lines, so it can update a field in place; passes
// in different directions are separated by workgroup barriers.  Every lane of the wavefront must
// be active (the workgroup has 256 threads, one per line; lines past the elements are a pad tile).
constexpr int kMfmaMaxN = 8;
enum { MF_B = 0, MF_BDB = 128, MF_BT = 256, MF_BTDBT = 384, kMfmaFrag = 640 };  // operator fragments [t][64]

template <typename T>
struct Mfma16;
template <>
struct Mfma16<float> {
  typedef float acc __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc mma(float a, float b, acc c)
  {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static constexpr __device__ int row(int kq, int r) { return 4 * kq + r; }
};
template <>
struct Mfma16<double> {
  typedef double acc __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc mma(double a, double b, acc c)
  {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static constexpr __device__ int row(int kq, int r) { return kq + 4 * r; }
};

// One pass of this wavefront's lines in direction DIR; own = the first point of the thread's own line
// in the tiles (the line of its lane's column in block j is lane 16 j + l % 16's own line), frag = the
// operator's fragments.  (Issuing the four blocks' reads together, then four independent product
// chains, measured slower: more registers, the same matrix-pipe time.)
template <int N, int DIR, typename T, int NIN, int NOUT>
__device__ __forceinline__ void mfma_lines(const T *__restrict__ frag, const T *in0, const T *in1, T *out0, T *out1,
                                           int own, int lane)
{
  static_assert(N <= kMfmaMaxN, "the line pass holds at most 8 points per field");
  constexpr int STR = DIR == DIR_Z ? 1 : DIR == DIR_Y ? N : N * N, KS = 2 * NIN;
  typedef typename Mfma16<T>::acc acc;
  const int kq = lane >> 4;
  T op[KS];
#pragma unroll
  for (int t = 0; t < KS; t++) op[t] = frag[t * 64 + lane];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int base = __builtin_amdgcn_ds_bpermute((16 * j + (lane & 15)) << 2, own);
    acc c = {T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int t = 0; t < KS; t++) {
      const int s = 4 * (t & 1) + kq;
      const T *src = t < 2 ? in0 : in1;
      const T v = src[base + (s < N ? s : N - 1) * STR];  // (the read stays on the line; past it the value is 0)
      c = Mfma16<T>::mma(op[t], s < N ? v : T(0), c);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int o = Mfma16<T>::row(kq, r), s = o & 7;
      T *dst = NOUT == 2 && o >= 8 ? out1 : out0;
      if ((NOUT == 2 || o < 8) && s < N) dst[base + s * STR] = c[r];
    }
  }
}

// forward() on the matrix cores: z (B), y (B and Dh B), x (B and Dh B; B of GY); the z derivative
// of the thread's own line stays on the vector units.  TA, TGX, TGY are the workgroup's tiles, le the
// thread's tile.  Called by all 256 threads.
template <int N, typename T>
__device__ __forceinline__ void forward_mfma(const T *tab, const T *frag, T *TA, T *TGX, T *TGY, int le, int a,
                                             int b, T (&q)[N], T (&gx)[N], T (&gy)[N], T (&gz)[N])
{
  const int lane = threadIdx.x & 63, e0 = le * N * N * N;
  const int bz = e0 + line_at<N, DIR_Z>(a, b, 0), by = e0 + line_at<N, DIR_Y>(a, b, 0),
            bx = e0 + line_at<N, DIR_X>(a, b, 0);
  mfma_lines<N, DIR_Z, T, 1, 1>(frag + MF_B, TA, TA, TA, TA, bz, lane);
  __syncthreads();
  mfma_lines<N, DIR_Y, T, 1, 2>(frag + MF_BDB, TA, TA, TA, TGY, by, lane);
  __syncthreads();
  mfma_lines<N, DIR_X, T, 1, 2>(frag + MF_BDB, TA, TA, TA, TGX, bx, lane);
  mfma_lines<N, DIR_X, T, 1, 1>(frag + MF_B, TGY, TGY, TGY, TGY, bx, lane);
  __syncthreads();
  ld_line<N, DIR_Z>(&TA[e0], a, b, q);
  eo_apply<N, -1>(tab, OP_D, q, gz);
  ld_line<N, DIR_Z>(&TGX[e0], a, b, gx);
  ld_line<N, DIR_Z>(&TGY[e0], a, b, gy);
}

// backward() on the matrix cores: the z transposed derivative of the thread's own line on the
// vector units, then x (B^T of the sum plus (Dh B)^T of d0; B^T of d1), y (B^T plus (Dh B)^T of
// d1), z (B^T).  Ends with a barrier; the result is in TA.
template <int N, typename T>
__device__ __forceinline__ void backward_mfma(const T *tab, const T *frag, T *TA, T *TGX, T *TGY, int le, int a,
                                              int b, const T (&v)[N], const T (&d0)[N], const T (&d1)[N],
                                              const T (&d2)[N])
{
  const int lane = threadIdx.x & 63, e0 = le * N * N * N;
  const int bz = e0 + line_at<N, DIR_Z>(a, b, 0), by = e0 + line_at<N, DIR_Y>(a, b, 0),
            bx = e0 + line_at<N, DIR_X>(a, b, 0);
  T w[N];
  eo_apply<N, -1>(tab, OP_DT, d2, w);
#pragma unroll
  for (int s = 0; s < N; s++) w[s] += v[s];
  st_line<N, DIR_Z>(&TA[e0], a, b, w);
  st_line<N, DIR_Z>(&TGY[e0], a, b, d1);
  st_line<N, DIR_Z>(&TGX[e0], a, b, d0);
  __syncthreads();
  mfma_lines<N, DIR_X, T, 2, 1>(frag + MF_BTDBT, TA, TGX, TA, TA, bx, lane);
  mfma_lines<N, DIR_X, T, 1, 1>(frag + MF_BT, TGY, TGY, TGY, TGY, bx, lane);
  __syncthreads();
  mfma_lines<N, DIR_Y, T, 2, 1>(frag + MF_BTDBT, TA, TGY, TA, TA, by, lane);
  __syncthreads();
  mfma_lines<N, DIR_Z, T, 1, 1>(frag + MF_BT, TA, TA, TA, TA, bz, lane);
  __syncthreads();
}

// x^e for x > 0.  Quarter-integer exponents (ph.qpow = 4e in [-8, 8]) use the hardware
// reciprocal square root refined by Newton steps; other exponents call pow.
__device__ inline double rsqrt_nr(double x)
{
  double y = __builtin_amdgcn_rsq(x);
  double r = fma(-x * y, y, 1.0);
  y = fma(0.5 * y, r, y);
  r = fma(-x * y, y, 1.0);
  return fma(0.5 * y, r, y);
}
__device__ inline float rsqrt_nr(float x) { return __builtin_amdgcn_rsqf(x); }

// The general power, outside the kernels' inlined code: the quarter-integer exponents of the common p take the
// fast path of powe, and this call's registers do not constrain the allocation of the rest.
__device__ __noinline__ double pow_general(double x, double e) { return pow(x, e); }
__device__ __noinline__ float pow_general(float x, float e) { return powf(x, e); }

template <typename T>
__device__ inline T powe(T x, T e, int qpow)
{
  if (qpow == 0) return T(1);
  if (qpow > 8 || qpow < -8) return pow_general(x, e);
  const T r2 = rsqrt_nr(x);         // x^(-1/2)
  const T r4 = r2 * rsqrt_nr(r2);   // x^(-1/4)
  const int n = qpow < 0 ? -qpow : qpow;
  const T base = qpow < 0 ? r4 : T(1) / r4;
  T acc = T(1);
#pragma unroll 1
  for (int t = 0; t < n; t++) acc *= base;
  return acc;
}

// sqrt for the rank-one factor: negative arguments (p > 2) give NaN as sqrt does.
template <typename T>
__device__ inline T sqrt_pos(T x) { return x > T(0) ? x * rsqrt_nr(x) : sqrt(x); }

// Column geometry of the trilinear map along the z line (xi0 = z_a, xi1 = z_b) of an element
// whose 8 vertices (x, y, z, k) sit at X[32]; also the line's coordinate and coefficient ends.
template <int N, typename T>
__device__ __forceinline__ ColGeo<T> column_geometry(const double *X, const double *phi, int a, int b,
                                                     double (&Pc)[2][3], double (&Kc)[2])
{
  const double fa1 = phi[a], fa0 = 1.0 - fa1, fb1 = phi[b], fb0 = 1.0 - fb1;
  double Rc[2][3], Sc[2][3];
#pragma unroll
  for (int c = 0; c < 2; c++) {
#pragma unroll
    for (int d = 0; d < 4; d++) {
      const double x00 = X[(0 + c) * 4 + d], x01 = X[(2 + c) * 4 + d], x10 = X[(4 + c) * 4 + d],
                   x11 = X[(6 + c) * 4 + d];
      const double pc = fa0 * (fb0 * x00 + fb1 * x01) + fa1 * (fb0 * x10 + fb1 * x11);
      if (d < 3) {
        Pc[c][d] = pc;
        Rc[c][d] = 0.5 * (fb0 * (x10 - x00) + fb1 * (x11 - x01));
        Sc[c][d] = 0.5 * (fa0 * (x01 - x00) + fa1 * (x11 - x10));
      } else {
        Kc[c] = pc;
      }
    }
  }
  const double c2[3] = {0.5 * (Pc[1][0] - Pc[0][0]), 0.5 * (Pc[1][1] - Pc[0][1]), 0.5 * (Pc[1][2] - Pc[0][2])};
  return make_colgeo<T>(Rc, Sc, c2);
}

// Order of the gather and scatter slots of a workgroup's EPB elements: consecutive slots walk the
// z lines of consecutive elements, (a, b, element, c), so that when the elements are neighbours along
// z a wavefront's loads and atomic adds cover long runs of consecutive nodes (atomic throughput
// depends on it).  Returns the tile index (element * n^3 + (a n + b) n + c) of slot t.
template <int N, int EPB>
__device__ __forceinline__ int brick_slot(int t)
{
  constexpr int RUN = EPB * N;
  const int ab = t / RUN, r = t - ab * RUN, le = r / N;
  return le * N * N * N + ab * N + (r - le * N);
}

// Threads of a workgroup holding EPB elements: one per line, EPB * n^2.  A workgroup smaller than a
// wavefront is rounded up to one: its extra lanes repeat lines of the same wavefront, which runs in
// lockstep, so they load, compute and store the same values at the same time as the lines' owners.
// (Repeating lines of another wavefront would race with in-place line updates.)
constexpr int wg_threads(int N, int EPB) { return EPB * N * N < 64 ? 64 : EPB * N * N; }

// Minimum of det J over the Gauss points of every element (geometry validation).
__global__ void min_det_kernel(int N, const int *__restrict__ everts, const double *__restrict__ vxyz, long nelem,
                               const double *__restrict__ z, double *__restrict__ mindet)
{
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  double X[8][3];
  for (int t = 0; t < 8; t++)
    for (int d = 0; d < 3; d++) X[t][d] = vxyz[3L * everts[e * 8 + t] + d];
  double m = 1e300;
  for (int a = 0; a < N; a++)
    for (int b = 0; b < N; b++)
      for (int c = 0; c < N; c++) {
        const double f[3][2] = {{(1 - z[a]) / 2, (1 + z[a]) / 2}, {(1 - z[b]) / 2, (1 + z[b]) / 2},
                                {(1 - z[c]) / 2, (1 + z[c]) / 2}};
        double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int t = 0; t < 8; t++) {
          const int ta = t >> 2, tb = (t >> 1) & 1, tc = t & 1;
          const double s0 = (ta ? 0.5 : -0.5) * f[1][tb] * f[2][tc];
          const double s1 = f[0][ta] * (tb ? 0.5 : -0.5) * f[2][tc];
          const double s2 = f[0][ta] * f[1][tb] * (tc ? 0.5 : -0.5);
          for (int d = 0; d < 3; d++) {
            J[d][0] += X[t][d] * s0;
            J[d][1] += X[t][d] * s1;
            J[d][2] += X[t][d] * s2;
          }
        }
        const double det = J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) -
                           J[0][1] * (J[1][0] * J[2][2] - J[2][0] * J[1][2]) +
                           J[0][2] * (J[1][0] * J[2][1] - J[2][0] * J[1][1]);
        m = det < m || det != det ? det : m;
      }
  mindet[e] = m;
}

// B, Dh, B^T, Dh^T, Dh B, (Dh B)^T of an n-point rule in the even-odd form of eo_apply,
// kEoOps * eo_size(n) entries; B and Dh are [q][m] row-major.
inline std::vector<double> db_matrix(int N, const std::vector<double> &B, const std::vector<double> &Dh)
{
  std::vector<double> DB(N * N, 0.0);
  for (int q = 0; q < N; q++)
    for (int m = 0; m < N; m++)
      for (int j = 0; j < N; j++) DB[q * N + m] += Dh[q * N + j] * B[j * N + m];
  return DB;
}

// Operator fragments of the matrix-core line passes (kMfmaFrag entries, [set + t * 64 + lane]): lane l
// holds Op[o = l % 16][k = 4 t + l / 16] of
//   MF_B      B (rows o < n),                          K = one field
//   MF_BDB    B in rows 0..7, Dh B in rows 8..15,      K = one field
//   MF_BT     B^T,                                     K = one field
//   MF_BTDBT  B^T over the first field, (Dh B)^T over the second (t = 2, 3)
// with positions past n zero.
inline std::vector<double> mfma_frags(int N, const std::vector<double> &B, const std::vector<double> &Dh)
{
  const std::vector<double> DB = db_matrix(N, B, Dh);
  std::vector<double> f(kMfmaFrag, 0.0);
  for (int t = 0; t < 4; t++)
    for (int l = 0; l < 64; l++) {
      const int o = l % 16, s = 4 * (t & 1) + l / 16;
      if (s >= N) continue;
      if (t < 2) {
        if (o < N) f[MF_B + t * 64 + l] = B[o * N + s];
        if ((o & 7) < N) f[MF_BDB + t * 64 + l] = (o < 8 ? B : DB)[(o & 7) * N + s];
        if (o < N) f[MF_BT + t * 64 + l] = B[s * N + o];
      }
      if (o < N) f[MF_BTDBT + t * 64 + l] = (t < 2 ? B : DB)[s * N + o];
    }
  return f;
}

inline std::vector<double> eo_tables(int N, const std::vector<double> &B, const std::vector<double> &Dh)
{
  const int H = N / 2, S = eo_size(N);
  const std::vector<double> DB = db_matrix(N, B, Dh);
  std::vector<double> t(kEoOps * S, 0.0);
  for (int opi = 0; opi < kEoOps; opi++) {
    auto A = [&](int q, int m) {
      const std::vector<double> &M = (opi == OP_B || opi == OP_BT) ? B : (opi == OP_D || opi == OP_DT) ? Dh : DB;
      return (opi == OP_B || opi == OP_D || opi == OP_DB) ? M[q * N + m] : M[m * N + q];
    };
    const bool sym = opi == OP_B || opi == OP_BT;
    double *T0 = &t[opi * S];
    for (int q = 0; q < H; q++)
      for (int m = 0; m < H; m++) {
        T0[q * H + m] = 0.5 * (A(q, m) + A(q, N - 1 - m));
        T0[H * H + q * H + m] = 0.5 * (A(q, m) - A(q, N - 1 - m));
      }
    if (N & 1) {
      for (int q = 0; q < H; q++) T0[2 * H * H + q] = A(q, H);
      for (int m = 0; m < H; m++) T0[2 * H * H + H + m] = A(H, m);
      T0[2 * H * H + 2 * H] = sym ? A(H, H) : 0.0;
    }
  }
  return t;
}

}  // namespace

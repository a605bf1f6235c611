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

using namespace dohp;

namespace {

enum { MODE_LIN = 0, MODE_FUN = 1, MODE_JAC = 2 };

// Row stride of the LDS tables and element tiles: even, so rows start 16-byte aligned.
constexpr int padded(int N) { return (N + 1) & ~1; }

template <int N, typename T>
struct Tab {
  double w[N];   // Gauss weights
  double phi[N]; // (1 + z_q)/2
};

// Basis tables staged in LDS once per workgroup, rows padded to NP entries:
//   B[q][j] (GLL nodal basis at Gauss point q), Dh[q][r] (collocated derivative at the Gauss
//   points), and their transposes, so that every row or column a thread needs is contiguous.
template <int N, typename T>
struct LdsTab {
  static constexpr int NP = padded(N);
  T B[N * NP], Dh[N * NP], BT[N * NP], DhT[N * NP];
};

struct Phys {
  double p, eps2, lambda, ea, eb, ec;
  int exact;
  int has_k;
};

constexpr int elems_per_block(int N) { return (256 / (N * N)) > 0 ? 256 / (N * N) : 1; }

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

// Logical block index: consecutive workgroups are dealt round-robin to the eight XCDs, so
// give each XCD a contiguous range of elements (neighbours share nodes through its L2).
__device__ inline long xcd_block(long b, long nb)
{
  const long q = nb / 8, r = nb % 8, x = b % 8, pos = b / 8;
  return x < r ? x * (q + 1) + pos : r * (q + 1) + (x - r) * q + pos;
}

// Forward: column values col[] at the nodes -> uq (Gauss points) and reference gradient.
// s1/s2 are this element's LDS tiles [i][j][k] with column stride NP.
template <int N, typename T>
__device__ inline void interp_grad(const LdsTab<N, T> &tb, const T (&col)[N], T (&uq)[N], T (&gx)[N], T (&gy)[N],
                                   T (&gz)[N], T *s1, T *s2, int i, int j)
{
  constexpr int NP = padded(N);
  T t[N], row[N];
#pragma unroll
  for (int q = 0; q < N; q++) {
    T s = 0;
#pragma unroll
    for (int k = 0; k < N; k++) s += tb.B[q * NP + k] * col[k];
    t[q] = s;
  }
#pragma unroll
  for (int k = 0; k < N; k++) s1[(i * N + j) * NP + k] = t[k];
#pragma unroll
  for (int m = 0; m < N; m++) row[m] = tb.B[j * NP + m];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; k++) t[k] = 0;
#pragma unroll
  for (int m = 0; m < N; m++)
#pragma unroll
    for (int k = 0; k < N; k++) t[k] += row[m] * s1[(i * N + m) * NP + k];
#pragma unroll
  for (int k = 0; k < N; k++) s2[(i * N + j) * NP + k] = t[k];
#pragma unroll
  for (int m = 0; m < N; m++) row[m] = tb.B[i * NP + m];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; k++) uq[k] = 0;
#pragma unroll
  for (int m = 0; m < N; m++)
#pragma unroll
    for (int k = 0; k < N; k++) uq[k] += row[m] * s2[(m * N + j) * NP + k];
#pragma unroll
  for (int q = 0; q < N; q++) {
    T s = 0;
#pragma unroll
    for (int k = 0; k < N; k++) s += tb.Dh[q * NP + k] * uq[k];
    gz[q] = s;
  }
#pragma unroll
  for (int k = 0; k < N; k++) s1[(i * N + j) * NP + k] = uq[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; k++) gy[k] = gx[k] = 0;
#pragma unroll
  for (int m = 0; m < N; m++) {
    const T dj = tb.Dh[j * NP + m], di = tb.Dh[i * NP + m];
#pragma unroll
    for (int k = 0; k < N; k++) {
      gy[k] += dj * s1[(i * N + m) * NP + k];
      gx[k] += di * s1[(m * N + j) * NP + k];
    }
  }
}

template <int N, typename T>
__device__ inline T tb_l(const T *tab, int r, int c) { return tab[r * padded(N) + c]; }

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

template <int N, typename T, int MODE>
__global__ void __launch_bounds__(256) ellip_kernel(const int *__restrict__ elist, long ne,
                                                    const int *__restrict__ enodes, const int *__restrict__ everts,
                                                    const double *__restrict__ vxyz, const double *__restrict__ kv,
                                                    const T *__restrict__ ustate, const T *__restrict__ xdir,
                                                    T *__restrict__ y, long nowned, int atomic,
                                                    const T *__restrict__ gtab, Tab<N, T> tb, Phys ph)
{
  constexpr int NN = N * N, N3 = NN * N, EPB = elems_per_block(N), NP = padded(N), CS = NN * NP;
  constexpr int N3X = MODE == MODE_LIN ? 1 : EPB * CS;
  __shared__ LdsTab<N, T> TB;
  __shared__ T S1[EPB * CS], S2[EPB * CS], S3[N3X];
  __shared__ int IDX[EPB * N3];
  __shared__ double VX[EPB * 8 * 4];

  const long lb = xcd_block(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x;
  const int le = tid / NN, ij = tid - le * NN, i = ij / N, j = ij - i * N;
  for (int t = tid; t < 4 * N * NP; t += blockDim.x) (&TB.B[0])[t] = gtab[t];
  for (int t = tid; t < EPB * N3; t += blockDim.x) {
    const int l = t / N3;
    const long es = lb * EPB + l;
    int v = -1;
    if (es < ne) {
      const long e = elist ? elist[es] : es;
      v = enodes[e * N3 + (t - l * N3)];
    }
    IDX[t] = v;
  }
  for (int t = tid; t < EPB * 8; t += blockDim.x) {
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
  // gather (coalesced over the element's node list)
  for (int t = tid; t < EPB * N3; t += blockDim.x) {
    const int c = IDX[t];
    T a = 0, b = 0;
    if (c >= 0) {
      if (MODE == MODE_FUN) a = ustate[c];
      if (MODE == MODE_JAC) { a = ustate[c]; b = c < nowned ? xdir[c] : T(0); }
      if (MODE == MODE_LIN) a = c < nowned ? xdir[c] : T(0);
    }
    const int l = t / N3, r = t - l * N3, col = r / N, k = r - col * N;
    S1[l * CS + col * NP + k] = a;
    if (MODE == MODE_JAC) S3[l * CS + col * NP + k] = b;
  }
  __syncthreads();
  T *s1 = &S1[le * CS], *s2 = &S2[le * CS];
  T ucol[N], xcol[N];
#pragma unroll
  for (int k = 0; k < N; k++) {
    ucol[k] = s1[(i * N + j) * NP + k];
    if (MODE == MODE_JAC) xcol[k] = S3[le * CS + (i * N + j) * NP + k];
  }
  __syncthreads();

  // geometry of the trilinear map along this thread's column (xi0 = z_i, xi1 = z_j)
  const double *X = &VX[le * 32];
  const double fa1 = tb.phi[i], fa0 = 1.0 - fa1, fb1 = tb.phi[j], fb0 = 1.0 - fb1;
  double Pc[2][3], Rc[2][3], Sc[2][3], Kc[2];
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
  const ColGeo<T> cg = make_colgeo<T>(Rc, Sc, c2);
  const T wij = (T)(tb.w[i] * tb.w[j]);

  T uq[N], gx[N], gy[N], gz[N];
  interp_grad<N, T>(TB, ucol, uq, gx, gy, gz, s1, s2, i, j);

  T vq[N], d0[N], d1[N], d2[N];
  if (MODE == MODE_LIN) {
#pragma unroll
    for (int k = 0; k < N; k++) {
      T r0[3], r1[3], r2[3], det;
      cg.eval((T)tb.phi[k], r0, r1, r2, det);
      const T kq = ph.has_k ? (T)(Kc[0] + (Kc[1] - Kc[0]) * tb.phi[k]) : T(1);
      // Dv = jw k grad(x) = w k (sum_m ghat_m r_m) / det;  df_m = jinv[m][:] . Dv = r_m . Dv / det
      const T sc = wij * (T)tb.w[k] * kq * recip(det);
      const T a0 = gx[k], a1 = gy[k], a2 = gz[k];
      T s3[3];
#pragma unroll
      for (int d = 0; d < 3; d++) s3[d] = a0 * r0[d] + a1 * r1[d] + a2 * r2[d];
      vq[k] = T(0);
      d0[k] = sc * (r0[0] * s3[0] + r0[1] * s3[1] + r0[2] * s3[2]);
      d1[k] = sc * (r1[0] * s3[0] + r1[1] * s3[1] + r1[2] * s3[2]);
      d2[k] = sc * (r2[0] * s3[0] + r2[1] * s3[1] + r2[2] * s3[2]);
    }
  } else {
    // Pointwise physics of the state in three phases so that one copy of the transcendental
    // code is emitted: (a) |grad u|^2 and jw per point into this thread's own LDS slots,
    // (b) eta, the rank-one Jacobian factor, forcing and Bratu term in a rolled loop, (c) the
    // per-point store.  Own-slot LDS traffic needs no barrier.
    __syncthreads();  // S1/S2/S3 are free once every column is in registers
    T *E = &S1[le * CS + (i * N + j) * NP], *F1 = &S2[le * CS + (i * N + j) * NP], *U = &S3[le * CS + (i * N + j) * NP];
    T gu0[N], gu1[N], gu2[N];
#pragma unroll
    for (int k = 0; k < N; k++) {
      T r0[3], r1[3], r2[3], det;
      cg.eval((T)tb.phi[k], r0, r1, r2, det);
      const T idet = recip(det);
      const T a0 = gx[k], a1 = gy[k], a2 = gz[k];
      gu0[k] = (a0 * r0[0] + a1 * r1[0] + a2 * r2[0]) * idet;
      gu1[k] = (a0 * r0[1] + a1 * r1[1] + a2 * r2[1]) * idet;
      gu2[k] = (a0 * r0[2] + a1 * r1[2] + a2 * r2[2]) * idet;
      E[k] = (T)ph.eps2 + T(0.5) * (gu0[k] * gu0[k] + gu1[k] * gu1[k] + gu2[k] * gu2[k]);
      F1[k] = wij * (T)tb.w[k] * det;  // jw
      U[k] = uq[k];
    }
#pragma unroll 1
    for (int k = 0; k < N; k++) {
      const T espg = E[k], jw = F1[k];
      const T kq = ph.has_k ? (T)(Kc[0] + (Kc[1] - Kc[0]) * tb.phi[k]) : T(1);
      const T eta = pow(espg, T((ph.p - 2) / 2));
      const T lexp = ph.lambda != 0 ? T(ph.lambda) * exp(U[k]) : T(0);
      if (MODE == MODE_FUN) {
        double f = 0.0;
        if (ph.exact >= 0) {
          const double g1 = tb.phi[k], g0 = 1.0 - g1;
          f = forcing(ph, Pc[0][0] * g0 + Pc[1][0] * g1, Pc[0][1] * g0 + Pc[1][1] * g1, Pc[0][2] * g0 + Pc[1][2] * g1);
        }
        E[k] = jw * eta * kq;               // Dv = jw eta k grad u
        U[k] = -jw * (T)f + -jw * lexp;     // v = -jw f - jw lambda e^u
      } else {
        // Dv = jw k (eta grad x - (s . grad x) s), s = sqrt(-(p-2)/2 eta / espg) grad u; folded:
        // Dv = E grad x - (sv . grad x) sv with sv = sqrt(jw k) s
        E[k] = jw * eta * kq;
        F1[k] = sqrt(-T((ph.p - 2) / 2) * eta / espg * jw * kq);
        U[k] = -jw * lexp;
      }
    }
    if (MODE == MODE_FUN) {
#pragma unroll
      for (int k = 0; k < N; k++) {
        T r0[3], r1[3], r2[3], det;
        cg.eval((T)tb.phi[k], r0, r1, r2, det);
        const T sc = E[k];
        const T Dv[3] = {sc * gu0[k], sc * gu1[k], sc * gu2[k]};
        const T idet = recip(det);
        vq[k] = U[k];
        d0[k] = (r0[0] * Dv[0] + r0[1] * Dv[1] + r0[2] * Dv[2]) * idet;
        d1[k] = (r1[0] * Dv[0] + r1[1] * Dv[1] + r1[2] * Dv[2]) * idet;
        d2[k] = (r2[0] * Dv[0] + r2[1] * Dv[1] + r2[2] * Dv[2]) * idet;
      }
    } else {
      // store: etaw, sv (3), lamjw per point; the state's fields are dead after this
      T etaw[N], sv0[N], sv1[N], sv2[N], lam[N];
#pragma unroll
      for (int k = 0; k < N; k++) {
        const T c1 = F1[k];
        etaw[k] = E[k];
        sv0[k] = c1 * gu0[k];
        sv1[k] = c1 * gu1[k];
        sv2[k] = c1 * gu2[k];
        lam[k] = U[k];
      }
      __syncthreads();
      T xq[N], hx[N], hy[N], hz[N];
      interp_grad<N, T>(TB, xcol, xq, hx, hy, hz, s1, s2, i, j);
#pragma unroll
      for (int k = 0; k < N; k++) {
        T r0[3], r1[3], r2[3], det;
        cg.eval((T)tb.phi[k], r0, r1, r2, det);
        const T idet = recip(det);
        const T b0 = hx[k], b1 = hy[k], b2 = hz[k];
        T gxd[3];
#pragma unroll
        for (int d = 0; d < 3; d++) gxd[d] = (b0 * r0[d] + b1 * r1[d] + b2 * r2[d]) * idet;
        const T dot = sv0[k] * gxd[0] + sv1[k] * gxd[1] + sv2[k] * gxd[2];
        const T Dv[3] = {etaw[k] * gxd[0] - dot * sv0[k], etaw[k] * gxd[1] - dot * sv1[k],
                         etaw[k] * gxd[2] - dot * sv2[k]};
        vq[k] = lam[k] * xq[k];
        d0[k] = (r0[0] * Dv[0] + r0[1] * Dv[1] + r0[2] * Dv[2]) * idet;
        d1[k] = (r1[0] * Dv[0] + r1[1] * Dv[1] + r1[2] * Dv[2]) * idet;
        d2[k] = (r2[0] * Dv[0] + r2[1] * Dv[1] + r2[2] * Dv[2]) * idet;
      }
    }
  }

  // transposed collocated derivative, combined at the Gauss points
  T t[N], row[N], rw2[N];
#pragma unroll
  for (int k = 0; k < N; k++) {
    T s = vq[k];
#pragma unroll
    for (int m = 0; m < N; m++) s += tb_l<N, T>(TB.DhT, k, m) * d2[m];
    t[k] = s;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; k++) {
    s1[(i * N + j) * NP + k] = d1[k];
    s2[(i * N + j) * NP + k] = d0[k];
  }
#pragma unroll
  for (int m = 0; m < N; m++) {
    row[m] = TB.DhT[j * NP + m];
    rw2[m] = TB.DhT[i * NP + m];
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < N; m++)
#pragma unroll
    for (int k = 0; k < N; k++) t[k] += row[m] * s1[(i * N + m) * NP + k] + rw2[m] * s2[(m * N + j) * NP + k];
  // B^T along z, y, x
  T t2[N];
#pragma unroll
  for (int k = 0; k < N; k++) {
    T s = 0;
#pragma unroll
    for (int q = 0; q < N; q++) s += tb_l<N, T>(TB.BT, k, q) * t[q];
    t2[k] = s;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; k++) s1[(i * N + j) * NP + k] = t2[k];
#pragma unroll
  for (int m = 0; m < N; m++) row[m] = TB.BT[j * NP + m];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; k++) t2[k] = 0;
#pragma unroll
  for (int m = 0; m < N; m++)
#pragma unroll
    for (int k = 0; k < N; k++) t2[k] += row[m] * s1[(i * N + m) * NP + k];
#pragma unroll
  for (int k = 0; k < N; k++) s2[(i * N + j) * NP + k] = t2[k];
#pragma unroll
  for (int m = 0; m < N; m++) row[m] = TB.BT[i * NP + m];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; k++) t[k] = 0;
#pragma unroll
  for (int m = 0; m < N; m++)
#pragma unroll
    for (int k = 0; k < N; k++) t[k] += row[m] * s2[(m * N + j) * NP + k];
#pragma unroll
  for (int k = 0; k < N; k++) s1[(i * N + j) * NP + k] = t[k];
  __syncthreads();
  // scatter (coalesced over the element's node list); Dirichlet rows are discarded
  for (int tt = tid; tt < EPB * N3; tt += blockDim.x) {
    const int c = IDX[tt];
    if (c >= 0 && c < nowned) {
      const int l = tt / N3, r = tt - l * N3, col = r / N, k = r - col * N;
      const T val = S1[l * CS + col * NP + k];
      if (atomic) atomicAdd(&y[c], val);
      else y[c] += val;
    }
  }
}

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
  DevBuf<double> state, kvert, tab64;
  DevBuf<float> state32, tab32;
  // host tables
  std::vector<double> B, Dh, w, phi;
};

namespace {

template <int N, typename T>
Tab<N, T> make_tab(const _p_dohpEllip *op)
{
  Tab<N, T> tb;
  for (int q = 0; q < N; q++) {
    tb.w[q] = op->w[q];
    tb.phi[q] = op->phi[q];
  }
  return tb;
}

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
  return ph;
}

template <int N, typename T, int MODE>
dErr launch_n(const _p_dohpEllip *op, const T *ust, const T *x, T *y, hipStream_t st)
{
  const dohpFS fs = op->fs;
  constexpr int EPB = elems_per_block(N);
  const Tab<N, T> tb = make_tab<N, T>(op);
  const Phys ph = make_phys(op);
  const dim3 block(EPB * N * N);
  if (op->scatter == 1) {
    const long ne = fs->nelem;
    hipLaunchKernelGGL((ellip_kernel<N, T, MODE>), dim3((ne + EPB - 1) / EPB), block, 0, st, nullptr, ne,
                       fs->d_elem_nodes.ptr, fs->d_elem_verts.ptr, fs->d_vcoords.ptr, op->kvert.ptr, ust, x, y,
                       fs->nowned, 1, gtab<T>(op), tb, ph);
  } else {
    for (int col = 0; col < 8; col++) {
      const long ne = fs->color_off[col + 1] - fs->color_off[col];
      if (!ne) continue;
      hipLaunchKernelGGL((ellip_kernel<N, T, MODE>), dim3((ne + EPB - 1) / EPB), block, 0, st,
                         fs->d_color_elems.ptr + fs->color_off[col], ne, fs->d_elem_nodes.ptr, fs->d_elem_verts.ptr,
                         fs->d_vcoords.ptr, op->kvert.ptr, ust, x, y, fs->nowned, 0, gtab<T>(op), tb, ph);
    }
  }
  DOHP_HIP(hipGetLastError());
  return 0;
}

template <typename T, int MODE>
dErr launch(const _p_dohpEllip *op, const T *ust, const T *x, T *y, hipStream_t st)
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
  {  // B, Dh, B^T, Dh^T with rows padded to an even length
    const int NP = (N + 1) & ~1;
    std::vector<double> t(4 * N * NP, 0.0);
    for (int q = 0; q < N; q++)
      for (int r = 0; r < N; r++) {
        t[0 * N * NP + q * NP + r] = op->B[q * N + r];
        t[1 * N * NP + q * NP + r] = op->Dh[q * N + r];
        t[2 * N * NP + r * NP + q] = op->B[q * N + r];
        t[3 * N * NP + r * NP + q] = op->Dh[q * N + r];
      }
    std::vector<float> tf(t.begin(), t.end());
    if (!err) err = op->tab64.upload(t.data(), t.size());
    if (!err) err = op->tab32.upload(tf.data(), tf.size());
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

extern "C" dErr dohpEllipShellMatMultF32(dohpEllip op, const float *x, float *y, void *stream)
{
  DOHP_CHECK(op && x && y, dohpERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  DOHP_CALL(zero(y, op->fs->nowned, st));
  if (is_linear(op)) return launch<float, MODE_LIN>(op, nullptr, x, y, st);
  DOHP_CHECK(op->has_state, dohpERR_ARG, "no linearization state: evaluate the function first");
  return launch<float, MODE_JAC>(op, op->state32.ptr, x, y, st);
}

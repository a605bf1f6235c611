// Embedded-Q1 preconditioning matrix (EllipJacobian, src/fs/tests/ellip.c:474-536) on the GPU,
// and the device CSR matrix it fills.
//
// The element's GLL sub-cells carry Q1 bases; each is integrated on its 2x2x2 Gauss points of the
// SPARSE rule (src/jacobi/quadrature/impls/tensor/tensorquad.c:32-49,73,120-169).  As in the
// reference (src/jacobi/impls/tensor/efstopo.c:660-717):
//   * sub-cell derivatives are +-1 per unit element-reference coordinate (no 1/h factor);
//   * the inverse Jacobian used for sub-cell point (patch, q) is the element-natural one at index
//     patch*8+q, because the iterator passes its element-ordered array
//     (src/fs/interface/fsrulesetit.c:684) to a routine that indexes it in patch order;
//   * u, grad u and the weights are taken at the correct patch points.
// dohpEllipSetQ1Mode(op, 1) assembles the corrected matrix instead (the Q1 finite-element stiffness of
// the GLL sub-mesh): sub-cell derivatives +-1/h and the inverse Jacobian at each patch point.
// The 8x8 sub-cell matrices are added into a CSR matrix over owned nodes (Dirichlet rows and
// columns are dropped, src/fs/interface/fsbuild.c:31-34).  One thread per sub-cell, in eight launches by the parity
// of the sub-cell's lattice position: sub-cells of one colour share no node, so the blocks are added without atomics
// in a fixed order and the matrix is bitwise reproducible.
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "common.h"

using namespace dohp;


namespace {

constexpr int kMaxP = 10;  // degree limit of the sub-cell kernel (n <= 11)

struct Q1Tab {
  int p, n, nq;                     // degree, nodes per direction, sparse points per direction (2p)
  int corrected;                    // 1: sub-cell derivatives +-1/h and the inverse Jacobian at the point
  double hsub[kMaxP];               // GLL sub-interval widths (element-reference coordinates)
  double zq[2 * kMaxP], wq[2 * kMaxP];
  double Bs[2 * kMaxP][kMaxP + 1];  // GLL basis at the sparse points
  double Ds[2 * kMaxP][kMaxP + 1];  // its derivative
};

struct Q1Phys {
  double p, eps2, lambda;
  int has_k;
  int M1, M2, Ly, Lz;  // block extents: elements, and node lattice (the state is in lattice order)
};

// One colour of the sub-cell lattice: sub-cells whose global lattice position (e_d p + a_d) has parity c_d in every
// direction d share no node, so a launch over one colour adds its blocks without atomics; the eight colours in
// turn give a fixed summation order (the assembled matrix is bitwise reproducible).
struct CellColor {
  int c[3];    // parity per direction
  int cnt[3];  // sub-cells of this colour per direction
  __host__ __device__ long n() const { return (long)cnt[0] * cnt[1] * cnt[2]; }
};

__device__ inline void jac_at(const double (&X)[8][3], double x0, double x1, double x2, double (&J)[3][3])
{
  const double f[3][2] = {{0.5 * (1 - x0), 0.5 * (1 + x0)}, {0.5 * (1 - x1), 0.5 * (1 + x1)},
                          {0.5 * (1 - x2), 0.5 * (1 + x2)}};
  for (int d = 0; d < 3; d++) J[d][0] = J[d][1] = J[d][2] = 0.0;
  for (int t = 0; t < 8; t++) {
    const int a = t >> 2, b = (t >> 1) & 1, c = t & 1;
    const double s0 = (a ? 0.5 : -0.5) * f[1][b] * f[2][c];
    const double s1 = f[0][a] * (b ? 0.5 : -0.5) * f[2][c];
    const double s2 = f[0][a] * f[1][b] * (c ? 0.5 : -0.5);
    for (int d = 0; d < 3; d++) {
      J[d][0] += X[t][d] * s0;
      J[d][1] += X[t][d] * s1;
      J[d][2] += X[t][d] * s2;
    }
  }
}

// Cofactor inverse (dGeomInvert3, include/dohpgeom.h:57-75).
__device__ inline double invert3(const double (&a)[3][3], double (&b)[3][3])
{
  const double b0 = a[1][1] * a[2][2] - a[2][1] * a[1][2];
  const double b3 = -(a[1][0] * a[2][2] - a[2][0] * a[1][2]);
  const double b6 = a[1][0] * a[2][1] - a[2][0] * a[1][1];
  const double det = a[0][0] * b0 + a[0][1] * b3 + a[0][2] * b6;
  const double id = 1.0 / det;
  b[0][0] = id * b0;
  b[0][1] = -id * (a[0][1] * a[2][2] - a[2][1] * a[0][2]);
  b[0][2] = id * (a[0][1] * a[1][2] - a[1][1] * a[0][2]);
  b[1][0] = id * b3;
  b[1][1] = id * (a[0][0] * a[2][2] - a[2][0] * a[0][2]);
  b[1][2] = -id * (a[0][0] * a[1][2] - a[1][0] * a[0][2]);
  b[2][0] = id * b6;
  b[2][1] = -id * (a[0][0] * a[2][1] - a[2][0] * a[0][1]);
  b[2][2] = id * (a[0][0] * a[1][1] - a[1][0] * a[0][1]);
  return det;
}

__device__ inline int find_col(const int *__restrict__ col, int lo, int hi, int c)
{
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (col[mid] < c) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// position of entry (r, c) of a symmetric 8x8 block in its row-major upper triangle
__host__ __device__ constexpr int sym_index(int r, int c)
{
  return r <= c ? r * 8 - r * (r - 1) / 2 + (c - r) : c * 8 - c * (c - 1) / 2 + (r - c);
}

// pow and exp out of line: their registers do not constrain the allocation of the sub-cell block
__device__ __noinline__ double q1_pow(double x, double e) { return pow(x, e); }
__device__ __noinline__ double q1_exp(double x) { return exp(x); }

// Sub-cell kernel of the embedded-Q1 assembly for n = N nodes per direction, one thread per GLL sub-cell: the
// 1-D tables sit in device memory (zq, wq, hsub, Bs, Ds at tab, read through the caches), the state's values at the
// sub-cell's 8 points accumulate over the element's x planes (no per-thread copies of tables or partial sums),
// the points' data waits in LDS and the 8x8 block in registers.
template <int N>
__global__ void __launch_bounds__(128) q1_cell_kernel(int nelem, const int *__restrict__ enodes,
                                                     const int *__restrict__ everts, const double *__restrict__ vxyz,
                                                     const double *__restrict__ kv, const double *__restrict__ ustate,
                                                     long nowned, const int *__restrict__ rowptr,
                                                     const int *__restrict__ colidx, double *__restrict__ val,
                                                     double *__restrict__ st, int corrected,
                                                     const double *__restrict__ tab, Q1Phys ph,
                                                     const float *__restrict__ cscale, CellColor col)
{
  constexpr int P = N - 1, NP3 = P * P * P, NQ = 2 * P;
  // per thread: u and the reference gradient at the 8 points
  __shared__ double PD[32][128];
  const double *zq = tab, *wq = tab + NQ, *hsub = tab + 2 * NQ, *Bs = tab + 2 * NQ + P, *Ds = Bs + NQ * N;
  const int lt = threadIdx.x;
  const long t = (long)blockIdx.x * blockDim.x + lt;
  if (t >= col.n()) return;
  // sub-cell of this colour: global lattice position G_d = c_d + 2 t_d, element e_d = G_d / p, sub-cell G_d % p
  const int t2 = (int)(t % col.cnt[2]), t1 = (int)((t / col.cnt[2]) % col.cnt[1]), t0 = (int)(t / ((long)col.cnt[1] * col.cnt[2]));
  const int G0 = col.c[0] + 2 * t0, G1 = col.c[1] + 2 * t1, G2 = col.c[2] + 2 * t2;
  const long e0 = G0 / P, e1 = G1 / P, e2 = G2 / P;
  const int pa = G0 - (int)e0 * P, pb = G1 - (int)e1 * P, pc = G2 - (int)e2 * P;
  const long e = (e0 * ph.M1 + e1) * ph.M2 + e2;
  const int pidx = (pa * P + pb) * P + pc;
  const long tid = e * NP3 + pidx;  // the sub-cell's index (learned scales)
  (void)nelem;
  const int *ev = everts + e * 8;
  const long lat0 = ((e0 * P) * ph.Ly + e1 * P) * ph.Lz + e2 * P;
  {
    const double *bz0 = Bs + (2 * pc) * N, *bz1 = bz0 + N, *dz0 = Ds + (2 * pc) * N, *dz1 = dz0 + N;
    const double *by0 = Bs + (2 * pb) * N, *by1 = by0 + N, *dy0 = Ds + (2 * pb) * N, *dy1 = dy0 + N;
    const double *bx0 = Bs + (2 * pa) * N, *bx1 = bx0 + N, *dx0 = Ds + (2 * pa) * N, *dx1 = dx0 + N;
    double PU[8], PG0[8], PG1[8], PG2[8];
#pragma unroll
    for (int q = 0; q < 8; q++) PU[q] = PG0[q] = PG1[q] = PG2[q] = 0.0;
    // the state enters the block only through eta(grad u) and s (p != 2) and lambda e^u: for the linear operator
    // (p = 2, lambda = 0) eta = 1 and s = 0 whatever the state, and its interpolation is skipped
    const int na = ustate && (ph.p != 2.0 || ph.lambda != 0.0) ? N : 0;
#pragma unroll 1
    for (int a = 0; a < na; a++) {
      double Y[2][2][3];  // [jj][kk][BB | DB | BD]
#pragma unroll
      for (int jj = 0; jj < 2; jj++)
#pragma unroll
        for (int kk = 0; kk < 2; kk++) Y[jj][kk][0] = Y[jj][kk][1] = Y[jj][kk][2] = 0.0;
#pragma unroll 1
      for (int b = 0; b < N; b++) {
        const double *ucol = ustate + lat0 + ((long)a * ph.Ly + b) * ph.Lz;
        double s00 = 0, s01 = 0, s10 = 0, s11 = 0;  // [kk][B | D]
#pragma unroll
        for (int c = 0; c < N; c++) {
          const double uv = ucol[c];
          s00 += bz0[c] * uv;
          s01 += dz0[c] * uv;
          s10 += bz1[c] * uv;
          s11 += dz1[c] * uv;
        }
        const double B[2] = {by0[b], by1[b]}, D[2] = {dy0[b], dy1[b]};
#pragma unroll
        for (int jj = 0; jj < 2; jj++) {
          Y[jj][0][0] += B[jj] * s00;
          Y[jj][0][1] += D[jj] * s00;
          Y[jj][0][2] += B[jj] * s01;
          Y[jj][1][0] += B[jj] * s10;
          Y[jj][1][1] += D[jj] * s10;
          Y[jj][1][2] += B[jj] * s11;
        }
      }
      const double Bx[2] = {bx0[a], bx1[a]}, Dx[2] = {dx0[a], dx1[a]};
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const int ii = q >> 2, jj = (q >> 1) & 1, kk = q & 1;
        PU[q] += Bx[ii] * Y[jj][kk][0];
        PG0[q] += Dx[ii] * Y[jj][kk][0];
        PG1[q] += Bx[ii] * Y[jj][kk][1];
        PG2[q] += Bx[ii] * Y[jj][kk][2];
      }
    }
#pragma unroll
    for (int q = 0; q < 8; q++) {
      PD[q][lt] = PU[q];
      PD[8 + q][lt] = PG0[q];
      PD[16 + q][lt] = PG1[q];
      PD[24 + q][lt] = PG2[q];
    }
  }
  const double i1[2][2] = {{0.78867513459481287, 0.21132486540518713}, {0.21132486540518713, 0.78867513459481287}};
  const double d1[2][2] = {{-1., 1.}, {-1., 1.}};
  // sub-cell derivative scale per reference direction: 1 (the reference's), 1/h (corrected), times the learned
  // per-sub-cell factors when given
  double sx = corrected ? 1.0 / hsub[pa] : 1.0, sy = corrected ? 1.0 / hsub[pb] : 1.0,
         sz = corrected ? 1.0 / hsub[pc] : 1.0;
  if (cscale) {
    sx *= (double)cscale[3 * tid];
    sy *= (double)cscale[3 * tid + 1];
    sz *= (double)cscale[3 * tid + 2];
  }
  // the block is symmetric (EllipPointwiseJacobian below gives K[r][c] = -w lambda e^u I_r I_c + eta w D_r.D_c
  // - w (s.D_r)(s.D_c)): its upper triangle, row-major, is accumulated
  double K[36];
#pragma unroll
  for (int t = 0; t < 36; t++) K[t] = 0.0;
#pragma unroll 1
  for (int q = 0; q < 8; q++) {
    const int ii = q >> 2, jj = (q >> 1) & 1, kk = q & 1;
    const int qi = 2 * pa + ii, qj = 2 * pb + jj, qk = 2 * pc + kk;
    double X[8][3];  // (read here, through the caches, so that they are not live in the block's accumulation)
    for (int t = 0; t < 8; t++)
      for (int d = 0; d < 3; d++) X[t][d] = vxyz[3 * (long)ev[t] + d];
    double J[3][3], Ji[3][3], cj[3][3];
    jac_at(X, zq[qi], zq[qj], zq[qk], J);
    const double det = invert3(J, Ji);
    const double w = wq[qi] * wq[qj] * wq[qk] * det;
    const double g0 = PD[8 + q][lt], g1 = PD[16 + q][lt], g2 = PD[24 + q][lt];
    double Du[3];
    for (int d = 0; d < 3; d++) Du[d] = g0 * Ji[0][d] + g1 * Ji[1][d] + g2 * Ji[2][d];
    double kq = 0;
    {
      const double f0[2] = {0.5 * (1 - zq[qi]), 0.5 * (1 + zq[qi])};
      const double f1[2] = {0.5 * (1 - zq[qj]), 0.5 * (1 + zq[qj])};
      const double f2[2] = {0.5 * (1 - zq[qk]), 0.5 * (1 + zq[qk])};
      for (int t = 0; t < 8; t++) kq += (ph.has_k ? kv[ev[t]] : 1.0) * f0[t >> 2] * f1[(t >> 1) & 1] * f2[t & 1];
    }
    if (corrected) {  // the inverse Jacobian at the point itself
      for (int r = 0; r < 3; r++)
        for (int d = 0; d < 3; d++) cj[r][d] = Ji[r][d];
    } else {  // element-natural point pidx*8+q, decoded on the (2p)^3 sparse lattice
      const int qe = pidx * 8 + q;
      const int ei = qe / (NQ * NQ), ej = (qe / NQ) % NQ, ek = qe % NQ;
      double J2[3][3];
      jac_at(X, zq[ei], zq[ej], zq[ek], J2);
      invert3(J2, cj);
    }
    // EllipPointwiseComputeStore (ellip.c:333-343) with the coefficient factor
    const double gamma = 0.5 * (Du[0] * Du[0] + Du[1] * Du[1] + Du[2] * Du[2]);
    const double espg = ph.eps2 + gamma;
    double eta = q1_pow(espg, (ph.p - 2) / 2);
    const double smdc = sqrt(-((ph.p - 2) / 2) * eta / espg) * sqrt(kq);
    eta *= kq;
    const double smd[3] = {smdc * Du[0], smdc * Du[1], smdc * Du[2]};
    const double lexp = ph.lambda != 0 ? ph.lambda * q1_exp(PD[q][lt]) : 0.0;
    double I8[8], D8[8][3];
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const int a0 = b >> 2, a1 = (b >> 1) & 1, a2 = b & 1;
      const double dX = sx * d1[ii][a0] * i1[jj][a1] * i1[kk][a2];
      const double dY = sy * i1[ii][a0] * d1[jj][a1] * i1[kk][a2];
      const double dZ = sz * i1[ii][a0] * i1[jj][a1] * d1[kk][a2];
      I8[b] = i1[ii][a0] * i1[jj][a1] * i1[kk][a2];
#pragma unroll
      for (int d = 0; d < 3; d++) D8[b][d] = dX * cj[0][d] + dY * cj[1][d] + dZ * cj[2][d];
    }
    // EllipPointwiseJacobian (ellip.c:369-377) on basis function c, tested against basis function r:
    // v_c = -w lambda e^u I_c, Dv_c = eta w D_c - w (s.D_c) s
    const double etaw = eta * w, lw = -w * lexp;
    double sd[8];
#pragma unroll
    for (int b = 0; b < 8; b++) sd[b] = smd[0] * D8[b][0] + smd[1] * D8[b][1] + smd[2] * D8[b][2];
#pragma unroll
    for (int r = 0, t = 0; r < 8; r++)
#pragma unroll
      for (int c = r; c < 8; c++, t++)
        K[t] += I8[r] * (lw * I8[c]) + etaw * (D8[r][0] * D8[c][0] + D8[r][1] * D8[c][1] + D8[r][2] * D8[c][2]) -
                w * sd[r] * sd[c];
  }
  const int *en = &enodes[e * N * N * N];
  int gidx[8];
#pragma unroll
  for (int b = 0; b < 8; b++) {
    const int a0 = b >> 2, a1 = (b >> 1) & 1, a2 = b & 1;
    gidx[b] = en[((pa + a0) * N + pb + a1) * N + pc + a2];
  }
#pragma unroll
  for (int r = 0; r < 8; r++) {
    const int gr = gidx[r];
    if (gr >= nowned) continue;
    if (st) {  // stencil form [27][rows]: the entry's slot is the lattice offset from the row node
#pragma unroll
      for (int c = 0; c < 8; c++) {
        const int gc = gidx[c];
        if (gc >= nowned) continue;
        const int sl = ((c >> 2) - (r >> 2) + 1) * 9 + (((c >> 1) & 1) - ((r >> 1) & 1) + 1) * 3 + ((c & 1) - (r & 1) + 1);
        st[sl * nowned + gr] += K[sym_index(r, c)];  // (no other sub-cell of this colour holds the entry)
      }
    } else {
      const int lo = rowptr[gr], hi = rowptr[gr + 1];
#pragma unroll
      for (int c = 0; c < 8; c++) {
        const int gc = gidx[c];
        if (gc >= nowned) continue;
        val[find_col(colidx, lo, hi, gc)] += K[sym_index(r, c)];
      }
    }
  }
}

template <typename T>
__global__ void csr_spmv_kernel(long nrows, const int *__restrict__ rowptr, const int *__restrict__ col,
                                const T *__restrict__ val, const T *__restrict__ x, T *__restrict__ y)
{
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  T s = 0;
  for (int k = rowptr[r]; k < rowptr[r + 1]; k++) s += val[k] * x[col[k]];
  y[r] = s;
}

__global__ void csr_diag_kernel(long nrows, const int *__restrict__ rowptr, const int *__restrict__ col,
                                const double *__restrict__ val, double *__restrict__ d)
{
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  double s = 0;
  for (int k = rowptr[r]; k < rowptr[r + 1]; k++)
    if (col[k] == r) s = val[k];
  d[r] = s;
}

__global__ void to_f32_kernel(const double *__restrict__ a, float *__restrict__ b, long n)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

__global__ void copy_f64_kernel(const double *__restrict__ a, double *__restrict__ b, long n)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void zero_f64_kernel(double *__restrict__ a, long n)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

// GLL node a of the element from the sub-interval widths.
__device__ inline double gll_node(const double *hsub, int a)
{
  double z = -1.0;
  for (int i = 0; i < a; i++) z += hsub[i];
  return z;
}

// Features of each GLL sub-cell for the learned derivative scales (kQ1CellInputs values): the log widths of its
// three reference sub-intervals, its centre in element-reference coordinates, the log physical lengths of the three
// reference directions at the centre minus their mean, log det J minus their sum (skew, <= 0), and log k at the
// centre minus the mean log k of the element's vertices (0 without a coefficient).
template <int N>
__global__ void q1_cell_features_kernel(int nelem, const int *__restrict__ everts, const double *__restrict__ vxyz,
                                        const double *__restrict__ kv, const double *__restrict__ tab,
                                        float *__restrict__ X)
{
  constexpr int P = N - 1, NP3 = P * P * P, NQ = 2 * P;
  const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= (long)nelem * NP3) return;
  const long e = tid / NP3;
  const int pidx = (int)(tid - e * NP3), pa = pidx / (P * P), pb = (pidx / P) % P, pc = pidx % P;
  const double *hsub = tab + 2 * NQ;
  const int *ev = everts + e * 8;
  double Xv[8][3];
  for (int t = 0; t < 8; t++)
    for (int d = 0; d < 3; d++) Xv[t][d] = vxyz[3 * (long)ev[t] + d];
  const double c0 = gll_node(hsub, pa) + 0.5 * hsub[pa], c1 = gll_node(hsub, pb) + 0.5 * hsub[pb],
               c2 = gll_node(hsub, pc) + 0.5 * hsub[pc];
  double J[3][3], Ji[3][3];
  jac_at(Xv, c0, c1, c2, J);
  const double det = invert3(J, Ji);
  double len[3], mean = 0;
  for (int a = 0; a < 3; a++) {
    len[a] = 0.5 * log(J[0][a] * J[0][a] + J[1][a] * J[1][a] + J[2][a] * J[2][a]);
    mean += len[a] / 3;
  }
  double lk = 0;
  if (kv) {
    const double f[3][2] = {{0.5 * (1 - c0), 0.5 * (1 + c0)}, {0.5 * (1 - c1), 0.5 * (1 + c1)},
                            {0.5 * (1 - c2), 0.5 * (1 + c2)}};
    double kc = 0, lmean = 0;
    for (int t = 0; t < 8; t++) {
      kc += kv[ev[t]] * f[0][t >> 2] * f[1][(t >> 1) & 1] * f[2][t & 1];
      lmean += log(kv[ev[t]]) / 8;
    }
    lk = log(kc) - lmean;
  }
  float *x = X + tid * kQ1CellInputs;
  x[0] = (float)log(hsub[pa]);
  x[1] = (float)log(hsub[pb]);
  x[2] = (float)log(hsub[pc]);
  x[3] = (float)c0;
  x[4] = (float)c1;
  x[5] = (float)c2;
  for (int a = 0; a < 3; a++) x[6 + a] = (float)(len[a] - mean);
  x[9] = (float)(log(det) - 3 * mean);
  x[10] = (float)lk;
}

// Gradient of the bilinear form w^T K(sigma) z of the linear operator's embedded-Q1 matrix (ellip p = 2, lambda = 0)
// in the sub-cell derivative scales: the sub-cell block is K = sum_q W_q D_q D_q^T with D_q = sum_a sigma_a U_qa
// (U_qa: the +-1 sub-cell derivative along reference direction a times the inverse Jacobian row a, as assembled), so
// d(w^T K z)/d sigma_a = sum_q W_q [(U_qa^T w) . (D_q^T z) + (D_q^T w) . (U_qa^T z)].  g[cell][a] += coef times it.
// w and z are vectors over the owned nodes (zero on Dirichlet nodes, whose rows and columns are dropped).
template <int N>
__global__ void __launch_bounds__(128) q1_scale_grad_kernel(int nelem, const int *__restrict__ enodes,
                                                           const int *__restrict__ everts,
                                                           const double *__restrict__ vxyz,
                                                           const double *__restrict__ kv, long nowned, int corrected,
                                                           const double *__restrict__ tab,
                                                           const float *__restrict__ cscale,
                                                           const double *__restrict__ wv, const double *__restrict__ zv,
                                                           double coef, float *__restrict__ g)
{
  constexpr int P = N - 1, NP3 = P * P * P, NQ = 2 * P;
  const double *zq = tab, *wq = tab + NQ, *hsub = tab + 2 * NQ;
  const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= (long)nelem * NP3) return;
  const long e = tid / NP3;
  const int pidx = (int)(tid - e * NP3), pa = pidx / (P * P), pb = (pidx / P) % P, pc = pidx % P;
  const int *ev = everts + e * 8;
  const int *en = &enodes[e * N * N * N];
  double w8[8], z8[8];
  for (int b = 0; b < 8; b++) {
    const int gi = en[((pa + (b >> 2)) * N + pb + ((b >> 1) & 1)) * N + pc + (b & 1)];
    w8[b] = gi < nowned ? wv[gi] : 0.0;
    z8[b] = gi < nowned ? zv[gi] : 0.0;
  }
  const double base[3] = {corrected ? 1.0 / hsub[pa] : 1.0, corrected ? 1.0 / hsub[pb] : 1.0,
                          corrected ? 1.0 / hsub[pc] : 1.0};
  double sig[3] = {1.0, 1.0, 1.0};
  if (cscale)
    for (int a = 0; a < 3; a++) sig[a] = (double)cscale[3 * tid + a];
  double Xv[8][3];
  for (int t = 0; t < 8; t++)
    for (int d = 0; d < 3; d++) Xv[t][d] = vxyz[3 * (long)ev[t] + d];
  const double i1[2][2] = {{0.78867513459481287, 0.21132486540518713}, {0.21132486540518713, 0.78867513459481287}};
  const double d1[2][2] = {{-1., 1.}, {-1., 1.}};
  double acc[3] = {0, 0, 0};
  for (int q = 0; q < 8; q++) {
    const int ii = q >> 2, jj = (q >> 1) & 1, kk = q & 1;
    const int qi = 2 * pa + ii, qj = 2 * pb + jj, qk = 2 * pc + kk;
    double J[3][3], Ji[3][3], cj[3][3];
    jac_at(Xv, zq[qi], zq[qj], zq[qk], J);
    const double det = invert3(J, Ji);
    double kq = 1.0;
    if (kv) {
      const double f0[2] = {0.5 * (1 - zq[qi]), 0.5 * (1 + zq[qi])};
      const double f1[2] = {0.5 * (1 - zq[qj]), 0.5 * (1 + zq[qj])};
      const double f2[2] = {0.5 * (1 - zq[qk]), 0.5 * (1 + zq[qk])};
      kq = 0;
      for (int t = 0; t < 8; t++) kq += kv[ev[t]] * f0[t >> 2] * f1[(t >> 1) & 1] * f2[t & 1];
    }
    if (corrected) {
      for (int r = 0; r < 3; r++)
        for (int d = 0; d < 3; d++) cj[r][d] = Ji[r][d];
    } else {  // element-natural point pidx*8+q, as the assembly
      const int qe = pidx * 8 + q;
      const int ei = qe / (NQ * NQ), ej = (qe / NQ) % NQ, ek = qe % NQ;
      double J2[3][3];
      jac_at(Xv, zq[ei], zq[ej], zq[ek], J2);
      invert3(J2, cj);
    }
    const double W = wq[qi] * wq[qj] * wq[qk] * det * kq;
    // reference derivative sums along each direction: sum_b dd_a[b] v_b
    double rw[3] = {0, 0, 0}, rz[3] = {0, 0, 0};
    for (int b = 0; b < 8; b++) {
      const int a0 = b >> 2, a1 = (b >> 1) & 1, a2 = b & 1;
      const double dd[3] = {d1[ii][a0] * i1[jj][a1] * i1[kk][a2], i1[ii][a0] * d1[jj][a1] * i1[kk][a2],
                            i1[ii][a0] * i1[jj][a1] * d1[kk][a2]};
      for (int a = 0; a < 3; a++) {
        rw[a] += dd[a] * w8[b];
        rz[a] += dd[a] * z8[b];
      }
    }
    double Uw[3][3], Uz[3][3], Dw[3] = {0, 0, 0}, Dz[3] = {0, 0, 0};
    for (int a = 0; a < 3; a++)
      for (int d = 0; d < 3; d++) {
        Uw[a][d] = base[a] * rw[a] * cj[a][d];
        Uz[a][d] = base[a] * rz[a] * cj[a][d];
        Dw[d] += sig[a] * Uw[a][d];
        Dz[d] += sig[a] * Uz[a][d];
      }
    for (int a = 0; a < 3; a++)
      acc[a] += W * (Uw[a][0] * Dz[0] + Uw[a][1] * Dz[1] + Uw[a][2] * Dz[2] + Dw[0] * Uz[a][0] + Dw[1] * Uz[a][1] +
                     Dw[2] * Uz[a][2]);
  }
  for (int a = 0; a < 3; a++) g[3 * tid + a] += (float)(coef * acc[a]);
}

}  // namespace

// Accessors the operator file uses (defined there).
extern "C" dErr dohpEllipGetQ1Context_Private(dohpEllip op, dohpFS *fs, dohpEllipParam *prm, const double **kvert,
                                              const double **state, int *q1mode);
const float *dohpEllipGetQ1CellScales_Private(dohpEllip op);

// CSR pattern of the 27-point stencil on a box of rows in lexicographic order, built on the device: row (i, j, k)
// couples to its box neighbours in increasing index, and its offset has a closed form (the count of entries of the
// rows before it, a product of per-direction counts 1 + [i > 0] + [i < n - 1]).
__device__ __forceinline__ long box_prefix(int i, int n) { return i + (i > 0 ? i - 1 : 0) + (i < n - 1 ? i : n - 1); }

__global__ void box_pattern_kernel(int nx, int ny, int nz, int *__restrict__ rowptr, int *__restrict__ col)
{
  const long n = (long)nx * ny * nz;
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int k = (int)(r % nz), j = (int)((r / nz) % ny), i = (int)(r / ((long)ny * nz));
  const long cj = 1 + (j > 0) + (j < ny - 1), ck = 1 + (k > 0) + (k < nz - 1), ci = 1 + (i > 0) + (i < nx - 1);
  const long Sy = 3L * ny - 2, Sz = 3L * nz - 2;
  const long off = box_prefix(i, nx) * Sy * Sz + ci * box_prefix(j, ny) * Sz + ci * cj * box_prefix(k, nz);
  rowptr[r] = (int)off;
  if (r == n - 1) rowptr[n] = (int)((3L * nx - 2) * Sy * Sz);
  long m = off;
  for (int di = -1; di <= 1; di++)
    for (int dj = -1; dj <= 1; dj++)
      for (int dk = -1; dk <= 1; dk++) {
        const int a = i + di, b = j + dj, c = k + dk;
        if (a < 0 || b < 0 || c < 0 || a >= nx || b >= ny || c >= nz) continue;
        col[m++] = (int)(((long)a * ny + b) * nz + c);
      }
}

static dErr build_pattern_box(dohpFS fs, _p_dohpMat *A, bool *done)
{
  const long L0 = (long)fs->p * fs->M[0] + 1, L1 = (long)fs->p * fs->M[1] + 1, L2 = (long)fs->p * fs->M[2] + 1;
  const int m = fs->dirichlet_mask;
  const long nx = L0 - ((m & 1) ? 2 : 0), ny = L1 - ((m & 1) ? 2 : 0), nz = L2 - ((m & 2) ? 1 : 0) - ((m & 4) ? 1 : 0);
  *done = false;
  if (nx <= 0 || ny <= 0 || nz <= 0 || nx * ny * nz != fs->nowned) return 0;
  const long nnz = (3 * nx - 2) * (3 * ny - 2) * (3 * nz - 2);
  if (nnz > 2147483647L) return 0;  // 32-bit offsets
  A->nrows = fs->nowned;
  A->nnz = nnz;
  DOHP_CALL(A->rowptr.alloc(A->nrows + 1));
  DOHP_CALL(A->col.alloc(nnz));
  DOHP_CALL(A->val.alloc(nnz));
  DOHP_CALL(A->val32.alloc(nnz));
  hipLaunchKernelGGL(box_pattern_kernel, dim3((unsigned)((A->nrows + 255) / 256)), dim3(256), 0, nullptr, (int)nx,
                     (int)ny, (int)nz, A->rowptr.ptr, A->col.ptr);
  DOHP_HIP(hipGetLastError());
  DOHP_HIP(hipDeviceSynchronize());
  *done = true;
  return 0;
}

// Pattern of the Q1 matrix on the GLL sub-cell lattice: row of every owned node, columns of the
// owned nodes among its 27 lattice neighbours (dFSCreateMatrix preallocates 27 per row,
// src/fs/interface/fs.c:770-813).
static dErr build_pattern(dohpFS fs, _p_dohpMat *A)
{
  bool done = false;
  DOHP_CALL(build_pattern_box(fs, A, &done));
  if (done) return 0;
  const long L0 = (long)fs->p * fs->M[0] + 1, L1 = (long)fs->p * fs->M[1] + 1, L2 = (long)fs->p * fs->M[2] + 1;
  // invert the closure numbering: lattice index of each closure node
  std::vector<long> c2l(fs->nclosure);
  const int n = fs->n, n3 = n * n * n;
  for (int e = 0; e < fs->nelem; e++) {
    const long e0 = e / ((long)fs->M[1] * fs->M[2]), e1 = (e / fs->M[2]) % fs->M[1], e2 = e % fs->M[2];
    for (int a = 0; a < n; a++)
      for (int b = 0; b < n; b++)
        for (int c = 0; c < n; c++)
          c2l[fs->elem_nodes[(long)e * n3 + (a * n + b) * n + c]] =
              ((e0 * fs->p + a) * L1 + e1 * fs->p + b) * L2 + e2 * fs->p + c;
  }
  std::vector<int> l2c(L0 * L1 * L2);
  for (long c = 0; c < fs->nclosure; c++) l2c[c2l[c]] = (int)c;
  A->nrows = fs->nowned;
  A->h_rowptr.assign(fs->nowned + 1, 0);
  std::vector<int> cols;
  cols.reserve(fs->nowned * 27);
  for (long r = 0; r < fs->nowned; r++) {
    const long l = c2l[r], I = l / (L1 * L2), J = (l / L2) % L1, K = l % L2;
    int buf[27], m = 0;
    for (int di = -1; di <= 1; di++)
      for (int dj = -1; dj <= 1; dj++)
        for (int dk = -1; dk <= 1; dk++) {
          const long a = I + di, b = J + dj, c = K + dk;
          if (a < 0 || b < 0 || c < 0 || a >= L0 || b >= L1 || c >= L2) continue;
          const int cc = l2c[(a * L1 + b) * L2 + c];
          if (cc < fs->nowned) buf[m++] = cc;
        }
    std::sort(buf, buf + m);
    cols.insert(cols.end(), buf, buf + m);
    A->h_rowptr[r + 1] = (int)cols.size();
  }
  A->nnz = (long)cols.size();
  DOHP_CALL(A->rowptr.upload(A->h_rowptr.data(), A->h_rowptr.size()));
  DOHP_CALL(A->col.upload(cols.data(), cols.size()));
  DOHP_CALL(A->val.alloc(A->nnz));
  DOHP_CALL(A->val32.alloc(A->nnz));
  return 0;
}

extern "C" dErr dohpEllipAssembleQ1(dohpEllip op, const double *u, dohpMat *out, void *stream)
{
  DOHP_CHECK(op && out, dohpERR_ARG, "null argument");
  dohpFS fs;
  dohpEllipParam prm;
  const double *kvert, *state;
  int q1mode = 0;
  DOHP_CALL(dohpEllipGetQ1Context_Private(op, &fs, &prm, &kvert, &state, &q1mode));
  DOHP_CHECK(fs->p <= kMaxP, dohpERR_ARG, "Q1 assembly supports degree <= 10");
  DOHP_CHECK(fs->is_block(), dohpERR_ARG, "the embedded-Q1 matrix needs a block space (dFSCreateBlock)");
  const float *cscale = dohpEllipGetQ1CellScales_Private(op);
  hipStream_t st = (hipStream_t)stream;
  if (u) DOHP_CALL(dohpEllipFunctionState_Private(op, u, st));
  _p_dohpMat *A = *out;
  if (A) A->mg_stale = true;  // new values: the cached hierarchy is recomputed, in its own buffers, when next used
  if (!A) {
    A = new _p_dohpMat;
    dErr err = build_pattern(fs, A);
    if (err) {
      delete A;
      return err;
    }
    *out = A;
  }
  Q1Tab tb;
  tb.p = fs->p;
  tb.n = fs->n;
  tb.corrected = q1mode;
  A->corrected = q1mode;
  A->scaled = cscale != nullptr;
  {
    double zg[kMaxN], wg[kMaxN];
    gauss_lobatto(fs->n, zg, wg);
    for (int i = 0; i < fs->p; i++) tb.hsub[i] = zg[i + 1] - zg[i];
  }
  int nq;
  DOHP_CALL(dohpTensorRule(dQUADRATURE_METHOD_SPARSE, dGAUSS_LEGENDRE, 2 * fs->p, &nq, tb.zq, tb.wq));
  tb.nq = nq;
  std::vector<double> Bs(nq * fs->n), Ds(nq * fs->n);
  DOHP_CALL(dohpTensorBasis(fs->n, nq, tb.zq, Bs.data(), Ds.data(), nullptr, nullptr));
  for (int q = 0; q < nq; q++)
    for (int a = 0; a < fs->n; a++) {
      tb.Bs[q][a] = Bs[q * fs->n + a];
      tb.Ds[q][a] = Ds[q * fs->n + a];
    }
  Q1Phys ph{prm.p, prm.epsilon * prm.epsilon, prm.lambda, kvert != nullptr, fs->M[1], fs->M[2],
            fs->p * fs->M[1] + 1, fs->p * fs->M[2] + 1};
  const long nt = (long)fs->nelem * fs->p * fs->p * fs->p;
  // On a block space the owned nodes form a box and the matrix is a 27-point stencil: the sub-cells add into the
  // stencil form directly (at computed slots, colour by colour), and the CSR values are read back from it.
  bool box = false;
  DOHP_CALL(dohpMatPrepareStencil_Private(A, fs, &box, st));
  if (box) {
    hipLaunchKernelGGL(zero_f64_kernel, dim3((27 * A->nrows + 255) / 256), dim3(256), 0, st, A->st64.ptr,
                       27 * A->nrows);
  } else {
    hipLaunchKernelGGL(zero_f64_kernel, dim3((A->nnz + 255) / 256), dim3(256), 0, st, A->val.ptr, A->nnz);
  }
  {  // 1-D tables for the kernel: zq, wq (2p each), hsub (p), Bs, Ds (2p x n each)
    const int nq2 = tb.nq, pp = tb.p, nn = tb.n;
    std::vector<double> tab;
    tab.insert(tab.end(), tb.zq, tb.zq + nq2);
    tab.insert(tab.end(), tb.wq, tb.wq + nq2);
    tab.insert(tab.end(), tb.hsub, tb.hsub + pp);
    for (int q = 0; q < nq2; q++) tab.insert(tab.end(), tb.Bs[q], tb.Bs[q] + nn);
    for (int q = 0; q < nq2; q++) tab.insert(tab.end(), tb.Ds[q], tb.Ds[q] + nn);
    if (!A->q1tab.holds(tab.size())) DOHP_CALL(A->q1tab.alloc(tab.size()));
    DOHP_HIP(hipMemcpyAsync(A->q1tab.ptr, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice, st));
    DOHP_HIP(hipStreamSynchronize(st));  // the host vector goes out of scope
  }
  (void)nt;
  double *stp = box ? A->st64.ptr : nullptr;
  for (int color = 0; color < 8; color++) {
  CellColor cc;
  for (int d = 0; d < 3; d++) {
    cc.c[d] = (color >> (2 - d)) & 1;
    const int G = fs->M[d] * fs->p;  // sub-cells per direction
    cc.cnt[d] = G > cc.c[d] ? (G - cc.c[d] + 1) / 2 : 0;
  }
  if (cc.n() == 0) continue;
  const dim3 grid((unsigned)((cc.n() + 127) / 128));
  switch (fs->n) {
#define DOHP_CASE(NN)                                                                                              \
  case NN:                                                                                                         \
    hipLaunchKernelGGL(q1_cell_kernel<NN>, grid, dim3(128), 0, st, fs->nelem, fs->d_elem_nodes.ptr,               \
                       fs->d_elem_verts.ptr, fs->d_vcoords.ptr, kvert, state, fs->nowned, A->rowptr.ptr, A->col.ptr, \
                       A->val.ptr, stp, tb.corrected, A->q1tab.ptr, ph, cscale, cc);                              \
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
  default: set_error("embedded-Q1 assembly supports degree 1..10"); return dohpERR_ARG;
  }
  DOHP_HIP(hipGetLastError());
  }
  if (box) {  // the multigrid and the FP64 product read the stencil form; the CSR values follow on demand
    A->csr_stale = true;
    return 0;
  }
  hipLaunchKernelGGL(to_f32_kernel, dim3((A->nnz + 255) / 256), dim3(256), 0, st, A->val.ptr, A->val32.ptr, A->nnz);
  DOHP_HIP(hipGetLastError());
  return 0;
}

extern "C" dErr dohpMatGetSizes(dohpMat A, long *nrows, long *nnz)
{
  DOHP_CHECK(A, dohpERR_ARG, "null matrix");
  if (nrows) *nrows = A->nrows;
  if (nnz) *nnz = A->nnz;
  return 0;
}

dErr dohpMatSyncCSR_Private(_p_dohpMat *A, hipStream_t st)
{
  if (!A->csr_stale) return 0;
  DOHP_CALL(dohpMatStencilToCSR_Private(A, st));
  A->csr_stale = false;
  return 0;
}

extern "C" dErr dohpMatGetCSR(dohpMat A, int *rowptr, int *col, double *val)
{
  DOHP_CHECK(A, dohpERR_ARG, "null matrix");
  DOHP_HIP(hipDeviceSynchronize());
  DOHP_CALL(dohpMatSyncCSR_Private(A, nullptr));
  DOHP_HIP(hipDeviceSynchronize());
  if (rowptr) DOHP_HIP(hipMemcpy(rowptr, A->rowptr.ptr, sizeof(int) * (A->nrows + 1), hipMemcpyDeviceToHost));
  if (col) DOHP_HIP(hipMemcpy(col, A->col.ptr, sizeof(int) * A->nnz, hipMemcpyDeviceToHost));
  if (val) DOHP_HIP(hipMemcpy(val, A->val.ptr, sizeof(double) * A->nnz, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" dErr dohpMatMult(dohpMat A, const double *x, double *y, void *stream)
{
  DOHP_CHECK(A && x && y, dohpERR_ARG, "null argument");
  if (A->st64.ptr) return dohpStencilMult_Private(A, x, y, (hipStream_t)stream);
  hipLaunchKernelGGL(csr_spmv_kernel<double>, dim3((A->nrows + 255) / 256), dim3(256), 0, (hipStream_t)stream, A->nrows,
                     A->rowptr.ptr, A->col.ptr, A->val.ptr, x, y);
  DOHP_HIP(hipGetLastError());
  return 0;
}

extern "C" dErr dohpMatMultF32(dohpMat A, const float *x, float *y, void *stream)
{
  DOHP_CHECK(A && x && y, dohpERR_ARG, "null argument");
  if (A->st32.ptr) {
    DOHP_CALL(dohpMatSyncCSR_Private(A, (hipStream_t)stream));
    return dohpStencilMultF32_Private(A, x, y, (hipStream_t)stream);
  }
  hipLaunchKernelGGL(csr_spmv_kernel<float>, dim3((A->nrows + 255) / 256), dim3(256), 0, (hipStream_t)stream, A->nrows,
                     A->rowptr.ptr, A->col.ptr, A->val32.ptr, x, y);
  DOHP_HIP(hipGetLastError());
  return 0;
}

extern "C" dErr dohpMatGetDiagonal(dohpMat A, double *d, void *stream)
{
  DOHP_CHECK(A && d, dohpERR_ARG, "null argument");
  if (A->st64.ptr)  // the stencil form holds the diagonal as its middle plane (the CSR values are read from it)
    hipLaunchKernelGGL(copy_f64_kernel, dim3((A->nrows + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       A->st64.ptr + 13 * A->nrows, d, A->nrows);
  else
    hipLaunchKernelGGL(csr_diag_kernel, dim3((A->nrows + 255) / 256), dim3(256), 0, (hipStream_t)stream, A->nrows,
                       A->rowptr.ptr, A->col.ptr, A->val.ptr, d);
  DOHP_HIP(hipGetLastError());
  return 0;
}

extern "C" dErr dohpMatDestroy(dohpMat *A)
{
  if (A && *A) {
    delete *A;
    *A = nullptr;
  }
  return 0;
}

// ------------------------------------------------------ learned sub-cell scales of the embedded-Q1 matrix

namespace {

template <typename F>
dErr q1_dispatch(int n, F f)
{
  switch (n) {
  case 2: f(std::integral_constant<int, 2>()); break;
  case 3: f(std::integral_constant<int, 3>()); break;
  case 4: f(std::integral_constant<int, 4>()); break;
  case 5: f(std::integral_constant<int, 5>()); break;
  case 6: f(std::integral_constant<int, 6>()); break;
  case 7: f(std::integral_constant<int, 7>()); break;
  case 8: f(std::integral_constant<int, 8>()); break;
  case 9: f(std::integral_constant<int, 9>()); break;
  case 10: f(std::integral_constant<int, 10>()); break;
  case 11: f(std::integral_constant<int, 11>()); break;
  default: set_error("embedded-Q1 assembly supports degree 1..10"); return dohpERR_ARG;
  }
  DOHP_HIP(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" dErr dohpQ1CellCount(dohpEllip op, long *ncell)
{
  DOHP_CHECK(op && ncell, dohpERR_ARG, "null argument");
  dohpFS fs;
  dohpEllipParam prm;
  const double *kv, *stt;
  DOHP_CALL(dohpEllipGetQ1Context_Private(op, &fs, &prm, &kv, &stt, nullptr));
  *ncell = (long)fs->nelem * fs->p * fs->p * fs->p;
  return 0;
}

extern "C" dErr dohpQ1CellFeatures(dohpEllip op, dohpMat q1, float *X, void *stream)
{
  DOHP_CHECK(op && q1 && X, dohpERR_ARG, "null argument");
  DOHP_CHECK(q1->q1tab.ptr, dohpERR_ARG, "the embedded-Q1 matrix has not been assembled");
  dohpFS fs;
  dohpEllipParam prm;
  const double *kv, *stt;
  DOHP_CALL(dohpEllipGetQ1Context_Private(op, &fs, &prm, &kv, &stt, nullptr));
  const long nt = (long)fs->nelem * fs->p * fs->p * fs->p;
  const dim3 grid((unsigned)((nt + 127) / 128));
  hipStream_t st = (hipStream_t)stream;
  return q1_dispatch(fs->n, [&](auto NN) {
    hipLaunchKernelGGL(q1_cell_features_kernel<decltype(NN)::value>, grid, dim3(128), 0, st, fs->nelem,
                       fs->d_elem_verts.ptr, fs->d_vcoords.ptr, kv, q1->q1tab.ptr, X);
  });
}

extern "C" dErr dohpQ1CellScaleGradient(dohpEllip op, dohpMat q1, const double *w, const double *z, double coef,
                                        float *g, void *stream)
{
  DOHP_CHECK(op && q1 && w && z && g, dohpERR_ARG, "null argument");
  DOHP_CHECK(q1->q1tab.ptr, dohpERR_ARG, "the embedded-Q1 matrix has not been assembled");
  dohpFS fs;
  dohpEllipParam prm;
  const double *kv, *stt;
  DOHP_CALL(dohpEllipGetQ1Context_Private(op, &fs, &prm, &kv, &stt, nullptr));
  DOHP_CHECK(prm.p == 2.0 && prm.lambda == 0.0, dohpERR_ARG,
             "the sub-cell scale gradient is for the linear operator (ellip p = 2, lambda = 0)");
  const float *cscale = dohpEllipGetQ1CellScales_Private(op);
  const long nt = (long)fs->nelem * fs->p * fs->p * fs->p;
  const dim3 grid((unsigned)((nt + 127) / 128));
  hipStream_t st = (hipStream_t)stream;
  return q1_dispatch(fs->n, [&](auto NN) {
    hipLaunchKernelGGL(q1_scale_grad_kernel<decltype(NN)::value>, grid, dim3(128), 0, st, fs->nelem,
                       fs->d_elem_nodes.ptr, fs->d_elem_verts.ptr, fs->d_vcoords.ptr, kv, fs->nowned, q1->corrected,
                       q1->q1tab.ptr, cscale, w, z, coef, g);
  });
}

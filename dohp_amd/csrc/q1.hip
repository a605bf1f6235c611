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
// columns are dropped, src/fs/interface/fsbuild.c:31-34).  One thread per sub-cell.
#include <algorithm>
#include <cmath>
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

__global__ void q1_assemble_kernel(int nelem, const int *__restrict__ enodes, const int *__restrict__ everts,
                                   const double *__restrict__ vxyz, const double *__restrict__ kv,
                                   const double *__restrict__ ustate, long nowned, const int *__restrict__ rowptr,
                                   const int *__restrict__ colidx, double *__restrict__ val, double *__restrict__ st,
                                   Q1Tab tb, Q1Phys ph)
{
  const int p = tb.p, n = tb.n, nq = tb.nq, np3 = p * p * p;
  const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= (long)nelem * np3) return;
  const long e = tid / np3;
  const int pidx = (int)(tid - e * np3);
  const int pa = pidx / (p * p), pb = (pidx / p) % p, pc = pidx % p;
  double X[8][3], kvx[8];
  for (int t = 0; t < 8; t++) {
    const long vi = everts[e * 8 + t];
    for (int d = 0; d < 3; d++) X[t][d] = vxyz[3 * vi + d];
    kvx[t] = ph.has_k ? kv[vi] : 1.0;
  }
  const int *en = &enodes[e * n * n * n];
  const long e2 = e % ph.M2, e1 = (e / ph.M2) % ph.M1, e0 = e / ((long)ph.M1 * ph.M2);
  // u and reference gradient at the sub-cell's 8 patch points, sum-factorized:
  // contract c for the two kk, then b for the two jj, then a for the two ii.
  double Tz[kMaxP + 1][kMaxP + 1][2][2];  // [a][b][kk][B|D]
  for (int a = 0; a < n; a++)
    for (int b = 0; b < n; b++) {
      double s[2][2] = {{0, 0}, {0, 0}};
      for (int c = 0; c < n; c++) {
        const double uv = ustate[(((int)e0 * p + a) * ph.Ly + (int)e1 * p + b) * ph.Lz + (int)e2 * p + c];
        for (int kk = 0; kk < 2; kk++) {
          s[kk][0] += tb.Bs[2 * pc + kk][c] * uv;
          s[kk][1] += tb.Ds[2 * pc + kk][c] * uv;
        }
      }
      for (int kk = 0; kk < 2; kk++) {
        Tz[a][b][kk][0] = s[kk][0];
        Tz[a][b][kk][1] = s[kk][1];
      }
    }
  double Ty[kMaxP + 1][2][2][3];  // [a][jj][kk][BB | DB | BD]
  for (int a = 0; a < n; a++)
    for (int jj = 0; jj < 2; jj++)
      for (int kk = 0; kk < 2; kk++) {
        double bb = 0, db = 0, bd = 0;
        for (int b = 0; b < n; b++) {
          const double B = tb.Bs[2 * pb + jj][b], D = tb.Ds[2 * pb + jj][b];
          bb += B * Tz[a][b][kk][0];
          db += D * Tz[a][b][kk][0];
          bd += B * Tz[a][b][kk][1];
        }
        Ty[a][jj][kk][0] = bb;
        Ty[a][jj][kk][1] = db;
        Ty[a][jj][kk][2] = bd;
      }
  double uq[8], gq[8][3], jwq[8], kq[8];
  double cj[8][3][3];  // inverse Jacobian used by the sub-cell tables (element-natural index quirk)
  for (int ii = 0; ii < 2; ii++)
    for (int jj = 0; jj < 2; jj++)
      for (int kk = 0; kk < 2; kk++) {
        const int q = (ii * 2 + jj) * 2 + kk;
        double u = 0, g0 = 0, g1 = 0, g2 = 0;
        for (int a = 0; a < n; a++) {
          const double B = tb.Bs[2 * pa + ii][a], D = tb.Ds[2 * pa + ii][a];
          u += B * Ty[a][jj][kk][0];
          g0 += D * Ty[a][jj][kk][0];
          g1 += B * Ty[a][jj][kk][1];
          g2 += B * Ty[a][jj][kk][2];
        }
        const int qi = 2 * pa + ii, qj = 2 * pb + jj, qk = 2 * pc + kk;
        double J[3][3], Ji[3][3];
        jac_at(X, tb.zq[qi], tb.zq[qj], tb.zq[qk], J);
        const double det = invert3(J, Ji);
        jwq[q] = tb.wq[qi] * tb.wq[qj] * tb.wq[qk] * det;
        uq[q] = u;
        // physical gradient: grad u[e] = sum_k ghat_k jinv[k][e]
        for (int d = 0; d < 3; d++) gq[q][d] = g0 * Ji[0][d] + g1 * Ji[1][d] + g2 * Ji[2][d];
        {
          const double f0[2] = {0.5 * (1 - tb.zq[qi]), 0.5 * (1 + tb.zq[qi])};
          const double f1[2] = {0.5 * (1 - tb.zq[qj]), 0.5 * (1 + tb.zq[qj])};
          const double f2[2] = {0.5 * (1 - tb.zq[qk]), 0.5 * (1 + tb.zq[qk])};
          double kk2 = 0;
          for (int t = 0; t < 8; t++) kk2 += kvx[t] * f0[t >> 2] * f1[(t >> 1) & 1] * f2[t & 1];
          kq[q] = kk2;
        }
        if (tb.corrected) {  // the inverse Jacobian at the point itself
          for (int r = 0; r < 3; r++)
            for (int d = 0; d < 3; d++) cj[q][r][d] = Ji[r][d];
        } else {  // element-natural point pidx*8+q, decoded on the (2p)^3 sparse lattice
          const int qe = pidx * 8 + q;
          const int ei = qe / (nq * nq), ej = (qe / nq) % nq, ek = qe % nq;
          double J2[3][3];
          jac_at(X, tb.zq[ei], tb.zq[ej], tb.zq[ek], J2);
          invert3(J2, cj[q]);
        }
      }
  // sub-cell tables and the 8x8 block (ellip.c:507-520)
  const double i1[2][2] = {{0.78867513459481287, 0.21132486540518713}, {0.21132486540518713, 0.78867513459481287}};
  const double d1[2][2] = {{-1., 1.}, {-1., 1.}};
  // per-direction derivative scale of the sub-cell's hat functions: 1 in the reference's arithmetic,
  // 1/h (the sub-interval width) when corrected
  const double sx = tb.corrected ? 1.0 / tb.hsub[pa] : 1.0, sy = tb.corrected ? 1.0 / tb.hsub[pb] : 1.0,
               sz = tb.corrected ? 1.0 / tb.hsub[pc] : 1.0;
  double K[8][8];
  for (int r = 0; r < 8; r++)
    for (int c = 0; c < 8; c++) K[r][c] = 0.0;
  for (int q = 0; q < 8; q++) {
    const int ii = q >> 2, jj = (q >> 1) & 1, kk = q & 1;
    // EllipPointwiseComputeStore (ellip.c:333-343) with the coefficient factor
    const double *Du = gq[q];
    const double gamma = 0.5 * (Du[0] * Du[0] + Du[1] * Du[1] + Du[2] * Du[2]);
    const double espg = ph.eps2 + gamma;
    double eta = pow(espg, (ph.p - 2) / 2);
    const double smdc = sqrt(-((ph.p - 2) / 2) * eta / espg) * sqrt(kq[q]);
    eta *= kq[q];
    const double smd[3] = {smdc * Du[0], smdc * Du[1], smdc * Du[2]};
    const double lexp = ph.lambda != 0 ? ph.lambda * exp(uq[q]) : 0.0;
    const double w = jwq[q];
    double I8[8], D8[8][3];
    for (int b = 0; b < 8; b++) {
      const int a0 = b >> 2, a1 = (b >> 1) & 1, a2 = b & 1;
      const double dX = sx * d1[ii][a0] * i1[jj][a1] * i1[kk][a2];
      const double dY = sy * i1[ii][a0] * d1[jj][a1] * i1[kk][a2];
      const double dZ = sz * i1[ii][a0] * i1[jj][a1] * d1[kk][a2];
      I8[b] = i1[ii][a0] * i1[jj][a1] * i1[kk][a2];
      for (int d = 0; d < 3; d++) D8[b][d] = dX * cj[q][0][d] + dY * cj[q][1][d] + dZ * cj[q][2][d];
    }
    for (int c = 0; c < 8; c++) {
      // EllipPointwiseJacobian (ellip.c:369-377) on basis function c
      const double dotw = (smd[0] * D8[c][0] + smd[1] * D8[c][1] + smd[2] * D8[c][2]) * w;
      const double etaw = eta * w;
      const double v = -w * lexp * I8[c];
      double dv[3];
      for (int d = 0; d < 3; d++) dv[d] = etaw * D8[c][d] - dotw * smd[d];
      for (int r = 0; r < 8; r++) K[r][c] += I8[r] * v + D8[r][0] * dv[0] + D8[r][1] * dv[1] + D8[r][2] * dv[2];
    }
  }
  int gidx[8];
  for (int b = 0; b < 8; b++) {
    const int a0 = b >> 2, a1 = (b >> 1) & 1, a2 = b & 1;
    gidx[b] = en[((pa + a0) * n + pb + a1) * n + pc + a2];
  }
  for (int r = 0; r < 8; r++) {
    const int gr = gidx[r];
    if (gr >= nowned) continue;
    if (st) {  // stencil form [27][rows]: the entry's slot is the lattice offset from the row node, no search
      for (int c = 0; c < 8; c++) {
        const int gc = gidx[c];
        if (gc >= nowned) continue;
        const int sl = ((c >> 2) - (r >> 2) + 1) * 9 + (((c >> 1) & 1) - ((r >> 1) & 1) + 1) * 3 + ((c & 1) - (r & 1) + 1);
        atomicAdd(&st[sl * nowned + gr], K[r][c]);
      }
    } else {
      const int lo = rowptr[gr], hi = rowptr[gr + 1];
      for (int c = 0; c < 8; c++) {
        const int gc = gidx[c];
        if (gc >= nowned) continue;
        atomicAdd(&val[find_col(colidx, lo, hi, gc)], K[r][c]);
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

__global__ void zero_f64_kernel(double *__restrict__ a, long n)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

}  // namespace

// Accessors the operator file uses (defined there).
extern "C" dErr dohpEllipGetQ1Context_Private(dohpEllip op, dohpFS *fs, dohpEllipParam *prm, const double **kvert,
                                              const double **state, int *q1mode);

// Pattern of the Q1 matrix on the GLL sub-cell lattice: row of every owned node, columns of the
// owned nodes among its 27 lattice neighbours (dFSCreateMatrix preallocates 27 per row,
// src/fs/interface/fs.c:770-813).
static dErr build_pattern(dohpFS fs, _p_dohpMat *A)
{
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
  hipStream_t st = (hipStream_t)stream;
  if (u) DOHP_CALL(dohpEllipFunctionState_Private(op, u, st));
  _p_dohpMat *A = *out;
  if (A) dohpMGDestroy(&A->mg);  // new values: the cached hierarchy is stale
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
  // stencil form directly (coalesced atomics at computed slots), and the CSR values are read back from it.
  bool box = false;
  DOHP_CALL(dohpMatPrepareStencil_Private(A, fs, &box, st));
  if (box) {
    hipLaunchKernelGGL(zero_f64_kernel, dim3((27 * A->nrows + 255) / 256), dim3(256), 0, st, A->st64.ptr,
                       27 * A->nrows);
  } else {
    hipLaunchKernelGGL(zero_f64_kernel, dim3((A->nnz + 255) / 256), dim3(256), 0, st, A->val.ptr, A->nnz);
  }
  hipLaunchKernelGGL(q1_assemble_kernel, dim3((nt + 127) / 128), dim3(128), 0, st, fs->nelem, fs->d_elem_nodes.ptr,
                     fs->d_elem_verts.ptr, fs->d_vcoords.ptr, kvert, state, fs->nowned, A->rowptr.ptr, A->col.ptr,
                     A->val.ptr, box ? A->st64.ptr : nullptr, tb, ph);
  DOHP_HIP(hipGetLastError());
  if (box) return dohpMatStencilToCSR_Private(A, st);
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

extern "C" dErr dohpMatGetCSR(dohpMat A, int *rowptr, int *col, double *val)
{
  DOHP_CHECK(A, dohpERR_ARG, "null matrix");
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
  if (A->st32.ptr) return dohpStencilMultF32_Private(A, x, y, (hipStream_t)stream);
  hipLaunchKernelGGL(csr_spmv_kernel<float>, dim3((A->nrows + 255) / 256), dim3(256), 0, (hipStream_t)stream, A->nrows,
                     A->rowptr.ptr, A->col.ptr, A->val32.ptr, x, y);
  DOHP_HIP(hipGetLastError());
  return 0;
}

extern "C" dErr dohpMatGetDiagonal(dohpMat A, double *d, void *stream)
{
  DOHP_CHECK(A && d, dohpERR_ARG, "null argument");
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

/*
 * dohp_oracle.c -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * CPU restatement of the arithmetic of dohp's matrix-free hp element operator.
 * It is the parity checker for the HIP path and the timed CPU baseline; the
 * product never links it.  Loop structure and the reference's quirks are kept
 * on purpose (three full contractions per gradient, +-1 sub-cell derivatives,
 * element-ordered inverse Jacobians indexed in patch order during Q1 assembly).
 */
#include "dohp_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846264338327950288
#define ORC_MAXN 16

/* When nonzero, the Q1 assembly reads the inverse Jacobian in patch-natural order (the
 * documented meaning of dEFSGetExplicitSparse's cjinv) instead of the element-natural
 * array the reference iterator actually passes.  Tests use it to show that the golden
 * refouts distinguish the two. */
static int orc_q1_cjinv_patch_order = 0;
void orc_set_q1_cjinv_patch_order(int flag) { orc_q1_cjinv_patch_order = flag; }

/* ------------------------------------------------------------------------- */
/* Jacobi polynomials (polylib.c:800-940 jacobfd/jacobd): three-term recurrence,
 * derivative d/dz P^{a,b}_n = (n+a+b+1)/2 P^{a+1,b+1}_{n-1}.                  */
static double jac_val(int n, double a, double b, double z)
{
  if (n == 0) return 1.0;
  double p0 = 1.0, p1 = 0.5 * (a - b + (a + b + 2.0) * z);
  for (int k = 1; k < n; k++) {
    const double s = 2.0 * k + a + b;
    const double a1 = 2.0 * (k + 1) * (k + a + b + 1) * s;
    const double a2 = (s + 1) * (a * a - b * b);
    const double a3 = s * (s + 1) * (s + 2);
    const double a4 = 2.0 * (k + a) * (k + b) * (s + 2);
    const double p2 = ((a2 + a3 * z) * p1 - a4 * p0) / a1;
    p0 = p1;
    p1 = p2;
  }
  return p1;
}
static double jac_der(int n, double a, double b, double z)
{
  if (n == 0) return 0.0;
  return 0.5 * (a + b + n + 1) * jac_val(n - 1, a + 1, b + 1, z);
}

/* Zeros of P^{a,b}_n, ascending: Newton with deflation from Chebyshev guesses
 * (the published algorithm of polylib.c:955-981). */
static void jac_zeros(int n, double a, double b, double *z)
{
  double rlast = 0;
  for (int k = 0; k < n; k++) {
    double r = -cos((2.0 * k + 1.0) * ORC_PI / (2.0 * n));
    if (k) r = 0.5 * (r + rlast);
    for (int it = 0; it < 100; it++) {
      double sum = 0;
      for (int i = 0; i < k; i++) sum += 1.0 / (r - z[i]);
      const double p = jac_val(n, a, b, r), dp = jac_der(n, a, b, r);
      const double delr = -p / (dp - sum * p);
      r += delr;
      if (fabs(delr) < 1e-16) break;
    }
    z[k] = r;
    rlast = r;
  }
  for (int i = 1; i < n; i++) /* insertion sort, ascending */
    for (int j = i; j > 0 && z[j - 1] > z[j]; j--) { double t = z[j]; z[j] = z[j - 1]; z[j - 1] = t; }
}

/* Gauss-Legendre nodes/weights (polylib.c:194 zwgj with alpha=beta=0). */
int orc_gauss_legendre(int np, double *z, double *w)
{
  if (np < 1) return 1;
  if (np == 1) { z[0] = 0; w[0] = 2; return 0; }
  jac_zeros(np, 0, 0, z);
  for (int i = 0; i < np; i++) {
    const double d = jac_der(np, 0, 0, z[i]);
    w[i] = 2.0 / (d * d * (1.0 - z[i] * z[i]));
  }
  return 0;
}

/* Gauss-Lobatto-Legendre nodes/weights (polylib.c:272-295 zwglj). */
int orc_gauss_lobatto(int np, double *z, double *w)
{
  if (np < 1) return 1;
  if (np == 1) { z[0] = 0; w[0] = 2; return 0; }
  z[0] = -1;
  z[np - 1] = 1;
  jac_zeros(np - 2, 1, 1, z + 1);
  for (int i = 0; i < np; i++) {
    const double pv = jac_val(np - 1, 0, 0, z[i]);
    w[i] = 2.0 / ((double)(np - 1) * np * pv * pv);
  }
  return 0;
}

/* Lagrange interpolant through GLL points (polylib.c:632-660 hglj, :741 Imglj):
 * h_j(z) = (1-z^2) P'_{np-1}(z) / [ d/dz((1-z^2)P'_{np-1})(z_j) (z - z_j) ], written with P^{1,1}_{np-2}. */
int orc_lagrange_interp_gll(int np, const double *zgll, int mz, const double *zm, double *im)
{
  const double eps = 100 * DBL_EPSILON;
  for (int i = 0; i < mz; i++) {
    const double z = zm[i];
    for (int j = 0; j < np; j++) {
      const double zj = zgll[j], dz = z - zj;
      double h;
      if (fabs(dz) < eps) h = 1.0;
      else {
        const double pj = jac_val(np - 2, 1, 1, zj), dpj = jac_der(np - 2, 1, 1, zj);
        const double denom = (1.0 - zj * zj) * dpj - 2.0 * zj * pj;
        h = (1.0 - z * z) * jac_val(np - 2, 1, 1, z) / (denom * dz);
      }
      im[i * np + j] = h;
    }
  }
  return 0;
}

/* Collocation derivative matrix on GLL points (polylib.c:452-510 Dglj), alpha=beta=0. */
int orc_deriv_gll(int np, const double *z, double *D)
{
  double pd[ORC_MAXN];
  if (np > ORC_MAXN) return 1;
  pd[0] = 2.0 * ((np % 2) ? -1.0 : 1.0) * (np - 1);
  for (int i = 1; i < np - 1; i++) pd[i] = jac_der(np - 2, 1, 1, z[i]) * (1.0 - z[i] * z[i]);
  pd[np - 1] = -2.0 * (np - 1);
  for (int i = 0; i < np; i++)
    for (int j = 0; j < np; j++) {
      if (i != j) D[i * np + j] = pd[i] / (pd[j] * (z[i] - z[j]));
      else if (i == 0) D[i * np + j] = -(double)(np - 1) * np / 4.0;
      else if (i == np - 1) D[i * np + j] = (double)(np - 1) * np / 4.0;
      else D[i * np + j] = 0.0;
    }
  return 0;
}

/* TensorBasisCreate (tensor.c:22-109): GLL nodal basis evaluated at rule points. */
int orc_tensor_basis(int P, int Q, const double *rcoord, double *interp, double *deriv,
                     double *interpT, double *derivT, double *node, double *weight)
{
  double cD[ORC_MAXN * ORC_MAXN], nd[ORC_MAXN], wt[ORC_MAXN];
  if (P < 2 || P > ORC_MAXN || Q < 1) return 1;
  orc_gauss_lobatto(P, nd, wt);
  orc_lagrange_interp_gll(P, nd, Q, rcoord, interp);
  orc_deriv_gll(P, nd, cD);
  for (int i = 0; i < Q; i++)
    for (int j = 0; j < P; j++) {
      double s = 0;
      for (int k = 0; k < P; k++) s += interp[i * P + k] * cD[k * P + j];
      deriv[i * P + j] = s;
    }
  for (int i = 0; i < Q; i++)
    for (int j = 0; j < P; j++) {
      if (interpT) interpT[j * Q + i] = interp[i * P + j];
      if (derivT) derivT[j * Q + i] = deriv[i * P + j];
    }
  if (node) memcpy(node, nd, P * sizeof(double));
  if (weight) memcpy(weight, wt, P * sizeof(double));
  return 0;
}

/* TensorGetRule (tensorquad.c:51-118) with two_point_private (:32-49). */
int orc_rule_1d(int method, int family, int order, int *size, double *coord, double *weight)
{
  int n;
  if (order < 0 || order >= 100) return 1;
  switch (method) {
  case ORC_QUAD_SPARSE: {
    n = order > 0 ? order : 1;
    *size = n;
    if (n == 1) { coord[0] = 0; weight[0] = 2; return 0; }
    const int nel = n / 2;
    const double x0 = family == ORC_GAUSS_LOBATTO ? -1.0 : -0.57735026918962573;
    const double x1 = -x0;
    double iv[ORC_MAXN * 8], un[ORC_MAXN * 8];
    if (nel + 1 > ORC_MAXN * 8) return 1;
    orc_gauss_lobatto(nel + 1, iv, un);
    for (int i = 0; i < nel; i++) {
      const double h = iv[i + 1] - iv[i];
      coord[2 * i + 0] = iv[i] + (x0 + 1) * h / 2.;
      coord[2 * i + 1] = iv[i] + (x1 + 1) * h / 2.;
      weight[2 * i + 0] = h / 2.;
      weight[2 * i + 1] = h / 2.;
    }
    return 0;
  }
  case ORC_QUAD_FAST:
    n = family == ORC_GAUSS_LOBATTO ? 2 + order / 2 : 1 + order / 2;
    break;
  case ORC_QUAD_SELF:
    if (family != ORC_GAUSS_LOBATTO) return 1;
    n = 1 + order / 2;
    break;
  default: return 1;
  }
  *size = n;
  if (n == 1) { coord[0] = 0; weight[0] = 2; return 0; }
  if (family == ORC_GAUSS_LOBATTO) return orc_gauss_lobatto(n, coord, weight);
  return orc_gauss_legendre(n, coord, weight);
}

/* dRulePatchSetup_Tensor (tensorquad.c:120-169), isotropic rule.  The reported patch
 * weight uses weight[jelem] for the third factor, as the reference does (:161). */
int orc_rule_patches(int Q1d, const double *w1d, int sparse, int *npatches, int *patchsize,
                     int *ind, double *pweight)
{
  int P1, N1;
  if (sparse) {
    P1 = Q1d > 1 ? 2 : 1;
    N1 = Q1d / P1;
    if (N1 * P1 != Q1d) return 1;
  } else {
    N1 = 1;
    P1 = Q1d;
  }
  const int NN = N1 * N1 * N1, PP = P1 * P1 * P1;
  *npatches = NN;
  *patchsize = PP;
  for (int i = 0; i < N1; i++)
    for (int j = 0; j < N1; j++)
      for (int k = 0; k < N1; k++)
        for (int ii = 0; ii < P1; ii++)
          for (int jj = 0; jj < P1; jj++)
            for (int kk = 0; kk < P1; kk++) {
              const int patch = (i * N1 + j) * N1 + k, pidx = (ii * P1 + jj) * P1 + kk;
              const int ie = i * P1 + ii, je = j * P1 + jj, ke = k * P1 + kk;
              if (ind) ind[patch * PP + pidx] = (ie * Q1d + je) * Q1d + ke;
              if (pweight) pweight[patch * PP + pidx] = w1d[ie] * w1d[je] * w1d[je];
              (void)ke;
            }
  return 0;
}

/* ------------------------------------------------------------------------- */
/* TensorMult_Hex_nounroll (inlinetmulthex.h:19-100): three sequential 1-D passes. */
int orc_tensor_mult_hex(int D, const int P[3], const int Q[3], const double *A[3],
                        const double *in, double *out, int imode)
{
  const int na = Q[0] * P[1] * P[2] * D, nb = Q[0] * Q[1] * P[2] * D;
  double amem[na], bmem[nb];
  memset(amem, 0, sizeof(amem));
  memset(bmem, 0, sizeof(bmem));
  if (imode == ORC_INSERT_VALUES) memset(out, 0, sizeof(double) * Q[0] * Q[1] * Q[2] * D);
  else if (imode != ORC_ADD_VALUES) return 1;
  const double *Ax = A[0], *Ay = A[1], *Az = A[2];
  for (int l = 0; l < Q[0]; l++) {
    const int JKD = P[1] * P[2] * D;
    double *aa = &amem[l * JKD];
    for (int i = 0; i < P[0]; i++) {
      const double *ff = &in[i * JKD];
      const double ax = Ax[l * P[0] + i];
      for (int jkd = 0; jkd < JKD; jkd++) aa[jkd] += ax * ff[jkd];
    }
  }
  for (int i = 0; i < Q[0]; i++) {
    const int JKD = P[1] * P[2] * D, KD = P[2] * D;
    double *aa = &amem[i * JKD];
    for (int l = 0; l < Q[1]; l++) {
      double *bb = &bmem[(i * Q[1] + l) * KD];
      for (int j = 0; j < P[1]; j++) {
        const double *aaa = &aa[j * KD];
        const double ay = Ay[l * P[1] + j];
        for (int kd = 0; kd < KD; kd++) bb[kd] += ay * aaa[kd];
      }
    }
  }
  if (D == 1) {
    for (int ij = 0; ij < Q[0] * Q[1]; ij++) {
      const double *bb = &bmem[ij * P[2]];
      double *cc = &out[ij * Q[2]];
      for (int l = 0; l < Q[2]; l++) {
        double s = 0;
        for (int k = 0; k < P[2]; k++) s += Az[l * P[2] + k] * bb[k];
        cc[l] += s;
      }
    }
  } else {
    for (int ij = 0; ij < Q[0] * Q[1]; ij++) {
      const double *bb = &bmem[ij * P[2] * D];
      for (int l = 0; l < Q[2]; l++) {
        double *cc = &out[(ij * Q[2] + l) * D];
        int kd = 0;
        for (int k = 0; k < P[2]; k++) {
          const double az = Az[l * P[2] + k];
          for (int d = 0; d < D; d++, kd++) cc[d] += az * bb[kd];
        }
      }
    }
  }
  return 0;
}

/* TensorRuleMapping / Transpose / NoMapping (efstopo.c:518-623). */
static int rule_mapping(int Q, const double *jinv, int D, const double *in, double *out, int imode)
{
  for (int i = 0; i < Q; i++)
    for (int d = 0; d < D; d++)
      for (int e = 0; e < 3; e++) {
        double v;
        if (!jinv) v = in[(e * Q + i) * D + d];
        else
          v = in[(0 * Q + i) * D + d] * jinv[i * 9 + 0 * 3 + e] + in[(1 * Q + i) * D + d] * jinv[i * 9 + 1 * 3 + e] +
              in[(2 * Q + i) * D + d] * jinv[i * 9 + 2 * 3 + e];
        if (imode == ORC_INSERT_VALUES) out[(i * D + d) * 3 + e] = v;
        else out[(i * D + d) * 3 + e] += v;
      }
  return 0;
}
static int rule_mapping_transpose(int Q, const double *jinv, int D, const double *in, double *out, int imode)
{
  if (!jinv) return 1;
  for (int i = 0; i < Q; i++)
    for (int d = 0; d < D; d++)
      for (int e = 0; e < 3; e++) {
        const double *u = &in[(i * D + d) * 3];
        const double v = u[0] * jinv[i * 9 + e * 3 + 0] + u[1] * jinv[i * 9 + e * 3 + 1] + u[2] * jinv[i * 9 + e * 3 + 2];
        if (imode == ORC_INSERT_VALUES) out[(e * Q + i) * D + d] = v;
        else out[(e * Q + i) * D + d] += v;
      }
  return 0;
}

/* One isotropic tensor basis on a rule: the cached TensorBasis of tensor.c:470-501. */
typedef struct {
  int P, Q;
  double B[ORC_MAXN * ORC_MAXN * 4], Dm[ORC_MAXN * ORC_MAXN * 4], BT[ORC_MAXN * ORC_MAXN * 4], DT[ORC_MAXN * ORC_MAXN * 4];
} orc_basis;

static int basis_setup(orc_basis *b, int P, int Q, const double *rcoord)
{
  if (P > ORC_MAXN || P * Q > ORC_MAXN * ORC_MAXN * 4) return 1;
  b->P = P;
  b->Q = Q;
  return orc_tensor_basis(P, Q, rcoord, b->B, b->Dm, b->BT, b->DT, NULL, NULL);
}

/* dEFSApply_Tensor_Hex (efstopo.c:368-423) for an isotropic P^3 basis on a Q^3 rule. */
static int efs_apply(const orc_basis *b, const double *jinv, int D, const double *in, double *out, int amode, int imode)
{
  const int P = b->P, Q = b->Q;
  const int PP[3] = {P, P, P}, QQ[3] = {Q, Q, Q};
  const int nq = Q * Q * Q;
  const double *A[3];
  int err = 0;
  switch (amode) {
  case ORC_APPLY_INTERP:
    A[0] = A[1] = A[2] = b->B;
    return orc_tensor_mult_hex(D, PP, QQ, A, in, out, imode);
  case ORC_APPLY_INTERP_TRANSPOSE:
    A[0] = A[1] = A[2] = b->BT;
    return orc_tensor_mult_hex(D, QQ, PP, A, in, out, imode);
  case ORC_APPLY_GRAD: {
    double df[3 * nq * D];
    A[0] = b->Dm; A[1] = b->B; A[2] = b->B;
    err |= orc_tensor_mult_hex(D, PP, QQ, A, in, df, ORC_INSERT_VALUES);
    A[0] = b->B; A[1] = b->Dm; A[2] = b->B;
    err |= orc_tensor_mult_hex(D, PP, QQ, A, in, df + nq * D, ORC_INSERT_VALUES);
    A[0] = b->B; A[1] = b->B; A[2] = b->Dm;
    err |= orc_tensor_mult_hex(D, PP, QQ, A, in, df + 2 * nq * D, ORC_INSERT_VALUES);
    err |= rule_mapping(nq, jinv, D, df, out, imode);
    return err;
  }
  case ORC_APPLY_GRAD_TRANSPOSE: {
    double df[3 * nq * D];
    err |= rule_mapping_transpose(nq, jinv, D, in, df, ORC_INSERT_VALUES);
    if (imode == ORC_INSERT_VALUES) memset(out, 0, sizeof(double) * P * P * P * D);
    else if (imode != ORC_ADD_VALUES) err = 1;
    A[0] = b->DT; A[1] = b->BT; A[2] = b->BT;
    err |= orc_tensor_mult_hex(D, QQ, PP, A, df, out, ORC_ADD_VALUES);
    A[0] = b->BT; A[1] = b->DT; A[2] = b->BT;
    err |= orc_tensor_mult_hex(D, QQ, PP, A, df + nq * D, out, ORC_ADD_VALUES);
    A[0] = b->BT; A[1] = b->BT; A[2] = b->DT;
    err |= orc_tensor_mult_hex(D, QQ, PP, A, df + 2 * nq * D, out, ORC_ADD_VALUES);
    return err;
  }
  default: return 1;
  }
}

int orc_efs_apply_hex(int P, int Q, const double *rcoord, const double *jinv, int D,
                      const double *in, double *out, int amode, int imode)
{
  orc_basis b;
  if (basis_setup(&b, P, Q, rcoord)) return 1;
  return efs_apply(&b, jinv, D, in, out, amode, imode);
}

/* dGeomInvert3 (include/dohpgeom.h:57-75): cofactor inverse and determinant. */
int orc_geom_invert3(const double *a, double *b, double *det)
{
  const double b0 = (a[1 * 3 + 1] * a[2 * 3 + 2] - a[2 * 3 + 1] * a[1 * 3 + 2]);
  const double b3 = -(a[1 * 3 + 0] * a[2 * 3 + 2] - a[2 * 3 + 0] * a[1 * 3 + 2]);
  const double b6 = (a[1 * 3 + 0] * a[2 * 3 + 1] - a[2 * 3 + 0] * a[1 * 3 + 1]);
  const double ldet = a[0] * b0 + a[1] * b3 + a[2] * b6;
  const double idet = 1.0 / ldet;
  b[0] = idet * b0;
  b[1] = -idet * (a[0 * 3 + 1] * a[2 * 3 + 2] - a[2 * 3 + 1] * a[0 * 3 + 2]);
  b[2] = idet * (a[0 * 3 + 1] * a[1 * 3 + 2] - a[1 * 3 + 1] * a[0 * 3 + 2]);
  b[3] = idet * b3;
  b[4] = idet * (a[0 * 3 + 0] * a[2 * 3 + 2] - a[2 * 3 + 0] * a[0 * 3 + 2]);
  b[5] = -idet * (a[0 * 3 + 0] * a[1 * 3 + 2] - a[1 * 3 + 0] * a[0 * 3 + 2]);
  b[6] = idet * b6;
  b[7] = -idet * (a[0 * 3 + 0] * a[2 * 3 + 1] - a[2 * 3 + 0] * a[0 * 3 + 1]);
  b[8] = idet * (a[0 * 3 + 0] * a[1 * 3 + 1] - a[1 * 3 + 0] * a[0 * 3 + 1]);
  *det = ldet;
  return 0;
}

/* dRuleComputePhysical (jacobi.c:383-414): jw = w*det, error on det < 0.  w holds the
 * tensor-product node weights (jacobi.c:334-369). Returns 2 on a negative Jacobian. */
int orc_rule_compute_physical(int n, const double *w, const double *jac, double *jinv, double *jw)
{
  for (int i = 0; i < n; i++) {
    double det;
    orc_geom_invert3(jac + i * 9, jinv + i * 9, &det);
    if (det < 0.) return 2;
    jw[i] = w[i] * det;
  }
  return 0;
}

/* Explicit element tables by unit applies (jacobi.c:530-570, dEFSGetExplicit). */
static int explicit_full(int P, int Q1d, const double *rcoord, const double *jinv, double *einterp, double *ederiv)
{
  const int np = P * P * P, nq = Q1d * Q1d * Q1d;
  double *tmp = calloc(np, sizeof(double)), *b = malloc(sizeof(double) * nq), *d = malloc(sizeof(double) * nq * 3);
  int err = 0;
  for (int i = 0; i < np; i++) {
    memset(tmp, 0, sizeof(double) * np);
    tmp[i] = 1;
    err |= orc_efs_apply_hex(P, Q1d, rcoord, jinv, 1, tmp, b, ORC_APPLY_INTERP, ORC_INSERT_VALUES);
    err |= orc_efs_apply_hex(P, Q1d, rcoord, jinv, 1, tmp, d, ORC_APPLY_GRAD, ORC_INSERT_VALUES);
    for (int q = 0; q < nq; q++) {
      einterp[q * np + i] = b[q];
      for (int l = 0; l < 3; l++) ederiv[(q * np + i) * 3 + l] = d[q * 3 + l];
    }
  }
  free(tmp); free(b); free(d);
  return err;
}

/* dEFSGetExplicitSparse_Tensor_Hex (efstopo.c:660-717) with the exact fallback
 * dEFSGetExplicitSparse_Basic (jacobi.c:611-643).  Sub-cell derivatives are +-1 in
 * element reference coordinates; cjinv is read at patch-natural offsets. */
int orc_explicit_sparse_hex(int P, int Q1d, const double *rcoord, int npatches, int Q,
                            const int *qidx, const double *cjinv, int eoffset, int *Pout,
                            int *eidx, double *interp, double *deriv)
{
  const int px = P - 1, py = P - 1, pz = P - 1, PP = 8;
  const double i1[2][2] = {{0.78867513459481287, 0.21132486540518713}, {0.21132486540518713, 0.78867513459481287}};
  const double d1[2][2] = {{-1., 1.}, {-1., 1.}};
  if (Q != 8 || px * py * pz != npatches) {
    const int np = P * P * P, nq = Q1d * Q1d * Q1d;
    double *ecj = NULL, *ei = malloc(sizeof(double) * nq * np), *ed = malloc(sizeof(double) * nq * np * 3);
    if (cjinv) {
      ecj = malloc(sizeof(double) * npatches * Q * 9);
      for (int i = 0; i < npatches * Q; i++)
        for (int j = 0; j < 9; j++) ecj[i * 9 + j] = cjinv[qidx[i] * 9 + j];
    }
    int err = explicit_full(P, Q1d, rcoord, ecj, ei, ed);
    *Pout = np;
    for (int i = 0; i < npatches; i++) {
      for (int k = 0; k < np; k++) eidx[i * np + k] = eoffset + k;
      for (int j = 0; j < Q; j++) {
        const int pq = i * Q + j, eq = qidx[pq];
        for (int k = 0; k < np; k++) {
          interp[pq * np + k] = ei[eq * np + k];
          for (int l = 0; l < 3; l++) deriv[(pq * np + k) * 3 + l] = ed[(eq * np + k) * 3 + l];
        }
      }
    }
    free(ecj); free(ei); free(ed);
    return err;
  }
  *Pout = PP;
  for (int i = 0; i < px; i++)
    for (int j = 0; j < py; j++)
      for (int k = 0; k < pz; k++) {
        const int pidx = (i * py + j) * pz + k;
        for (int a = 0; a < 2; a++)
          for (int b = 0; b < 2; b++)
            for (int c = 0; c < 2; c++)
              eidx[pidx * PP + (a * 2 + b) * 2 + c] = eoffset + ((i + a) * (py + 1) + j + b) * (pz + 1) + k + c;
        for (int ii = 0; ii < 2; ii++)
          for (int jj = 0; jj < 2; jj++)
            for (int kk = 0; kk < 2; kk++) {
              const int pq = (ii * 2 + jj) * 2 + kk;
              static const double eye[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
              const double *X = cjinv ? &cjinv[(pidx * 8 + pq) * 9] : eye;
              if (qidx[pidx * Q + pq] != ((i * 2 + ii) * py * 2 + (j * 2 + jj)) * pz * 2 + (k * 2 + kk)) return 3;
              for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2; b++)
                  for (int c = 0; c < 2; c++) {
                    const int bidx = (a * 2 + b) * 2 + c, off = (pidx * Q + pq) * PP + bidx;
                    const double dX = d1[ii][a] * i1[jj][b] * i1[kk][c];
                    const double dY = i1[ii][a] * d1[jj][b] * i1[kk][c];
                    const double dZ = i1[ii][a] * i1[jj][b] * d1[kk][c];
                    interp[off] = i1[ii][a] * i1[jj][b] * i1[kk][c];
                    deriv[off * 3 + 0] = dX * X[0] + dY * X[3] + dZ * X[6];
                    deriv[off * 3 + 1] = dX * X[1] + dY * X[4] + dZ * X[7];
                    deriv[off * 3 + 2] = dX * X[2] + dY * X[5] + dZ * X[8];
                  }
            }
      }
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Exact solutions (ellip.c:42-117).  The forcing is f = -div(eta grad u) with
 * eta = (eps^2 + |grad u|^2/2)^((p-2)/2), evaluated from u's gradient and Hessian:
 *   f = -eta lap(u) - (p-2)/2 (eps^2+gamma)^((p-4)/2) grad(u)^T H grad(u).
 * Exact 4 uses the Poisson forcing 3 pi^2 u regardless of p, as the reference does. */
static void exact_uh(const orc_problem *pb, const double X[3], double *u, double du[3], double H[9])
{
  const double a = pb->ea, b = pb->eb, c = pb->ec, x = X[0], y = X[1], z = X[2];
  memset(H, 0, 9 * sizeof(double));
  switch (pb->exact) {
  case 0: {
    const double cx = cos(a * x), sx = sin(a * x), ey = exp(b * y), cz = cos(c * z), sz = sin(c * z);
    *u = cx * ey * sz;
    du[0] = -a * sx * ey * sz; du[1] = b * cx * ey * sz; du[2] = c * cx * ey * cz;
    H[0] = -a * a * cx * ey * sz; H[4] = b * b * cx * ey * sz; H[8] = -c * c * cx * ey * sz;
    H[1] = H[3] = -a * b * sx * ey * sz;
    H[2] = H[6] = -a * c * sx * ey * cz;
    H[5] = H[7] = b * c * cx * ey * cz;
  } break;
  case 1:
    *u = a * x * x + b * y * y + c * (1 - z * z);
    du[0] = 2 * a * x; du[1] = 2 * b * y; du[2] = -2 * c * z;
    H[0] = 2 * a; H[4] = 2 * b; H[8] = -2 * c;
    break;
  case 2:
    *u = a * x + b * y + c * z;
    du[0] = a; du[1] = b; du[2] = c;
    break;
  case 3:
    *u = a * x * x * x + b * y * y * z + c * (1 - z * z) * x;
    du[0] = 3 * a * x * x + c * (1 - z * z); du[1] = 2 * b * y * z; du[2] = b * y * y - 2 * c * z * x;
    H[0] = 6 * a * x; H[4] = 2 * b * z; H[8] = -2 * c * x;
    H[1] = H[3] = 0; H[2] = H[6] = -2 * c * z; H[5] = H[7] = 2 * b * y;
    break;
  case 4: {
    const double s0 = sin(ORC_PI * x), s1 = sin(ORC_PI * y), s2 = sin(ORC_PI * z);
    const double c0 = cos(ORC_PI * x), c1 = cos(ORC_PI * y), c2 = cos(ORC_PI * z), pi = ORC_PI;
    *u = s0 * s1 * s2;
    du[0] = pi * c0 * s1 * s2; du[1] = pi * s0 * c1 * s2; du[2] = pi * s0 * s1 * c2;
    H[0] = H[4] = H[8] = -pi * pi * s0 * s1 * s2;
    H[1] = H[3] = pi * pi * c0 * c1 * s2; H[2] = H[6] = pi * pi * c0 * s1 * c2; H[5] = H[7] = pi * pi * s0 * c1 * c2;
  } break;
  default:
    *u = 0; du[0] = du[1] = du[2] = 0;
  }
}

int orc_exact_solution(const orc_problem *pb, const double x[3], double *u, double du[3])
{
  double H[9];
  exact_uh(pb, x, u, du, H);
  return 0;
}

int orc_forcing(const orc_problem *pb, const double x[3], double *f)
{
  double u, du[3], H[9];
  if (pb->exact < 0) { *f = 0; return 0; }
  exact_uh(pb, x, &u, du, H);
  if (pb->exact == 4) { *f = 3 * ORC_PI * ORC_PI * u; return 0; }
  const double e2g = pb->eps * pb->eps + 0.5 * (du[0] * du[0] + du[1] * du[1] + du[2] * du[2]);
  const double eta = pow(e2g, (pb->ep - 2) / 2);
  double quad = 0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) quad += du[i] * H[i * 3 + j] * du[j];
  *f = -eta * (H[0] + H[4] + H[8]) - (pb->ep - 2) / 2 * pow(e2g, (pb->ep - 4) / 2) * quad;
  return 0;
}

/* EllipStore and pointwise kernels (ellip.c:119-123, 333-377).  kf scales the
 * coefficient (variable-coefficient problems); kf = 1 gives the reference. */
typedef struct { double eta, smd[3], lexp; } orc_store;
static void compute_store(const orc_problem *pb, double u, const double Du[3], double kf, orc_store *st)
{
  const double gamma = 0.5 * (Du[0] * Du[0] + Du[1] * Du[1] + Du[2] * Du[2]);
  const double espg = pb->eps * pb->eps + gamma, p = pb->ep;
  st->eta = pow(espg, (p - 2) / 2);
  const double smd = sqrt(-((p - 2) / 2) * st->eta / espg);
  for (int i = 0; i < 3; i++) st->smd[i] = smd * Du[i];
  st->lexp = pb->lambda != 0 ? pb->lambda * exp(u) : 0;
  if (kf != 1.0) {
    const double sk = sqrt(kf);
    st->eta *= kf;
    for (int i = 0; i < 3; i++) st->smd[i] *= sk;
  }
}
static void pointwise_jacobian(const orc_store *st, double w, double u, const double Du[3], double *v, double Dv[3])
{
  const double dotw = (st->smd[0] * Du[0] + st->smd[1] * Du[1] + st->smd[2] * Du[2]) * w;
  const double etaw = st->eta * w;
  *v = -w * st->lexp * u;
  for (int i = 0; i < 3; i++) Dv[i] = etaw * Du[i] - dotw * st->smd[i];
}

/* ------------------------------------------------------------------------- */
/* Per-element context of dRulesetIteratorGetPatchApplied (fsrulesetit.c:517-589). */
typedef struct {
  int n, Q1d, nq;
  double rc[4 * ORC_MAXN], rw[4 * ORC_MAXN];
  double *w;     /* tensor node weights [nq] */
  int sparse, npatches, patchsize;
  int *patchind;
  orc_basis bc, bu;   /* coordinate (P=2) and solution (P=n) bases on this rule */
} orc_rule;

static int rule_setup(orc_rule *r, int p, int method)
{
  memset(r, 0, sizeof(*r));
  r->n = p + 1;
  const int fam = method == ORC_QUAD_SELF ? ORC_GAUSS_LOBATTO : ORC_GAUSS_LEGENDRE;
  if (orc_rule_1d(method, fam, 2 * p, &r->Q1d, r->rc, r->rw)) return 1;
  r->nq = r->Q1d * r->Q1d * r->Q1d;
  r->w = malloc(sizeof(double) * r->nq);
  for (int i = 0; i < r->Q1d; i++)
    for (int j = 0; j < r->Q1d; j++)
      for (int k = 0; k < r->Q1d; k++) r->w[(i * r->Q1d + j) * r->Q1d + k] = r->rw[i] * r->rw[j] * r->rw[k];
  r->sparse = method == ORC_QUAD_SPARSE;
  r->patchind = malloc(sizeof(int) * r->nq);
  if (orc_rule_patches(r->Q1d, r->rw, r->sparse, &r->npatches, &r->patchsize, r->patchind, NULL)) return 1;
  if (basis_setup(&r->bc, 2, r->Q1d, r->rc)) return 1;
  if (basis_setup(&r->bu, r->n, r->Q1d, r->rc)) return 1;
  return 0;
}
static void rule_free(orc_rule *r) { free(r->w); free(r->patchind); }

typedef struct {
  double *xq, *Jq, *cjinv, *jw, *uq, *duq, *v, *dv, *ye, *kq;
} orc_work;
static void work_alloc(orc_work *wk, int nq, int np)
{
  wk->xq = malloc(sizeof(double) * nq * 3);
  wk->Jq = malloc(sizeof(double) * nq * 9);
  wk->cjinv = malloc(sizeof(double) * nq * 9);
  wk->jw = malloc(sizeof(double) * nq);
  wk->uq = malloc(sizeof(double) * nq);
  wk->duq = malloc(sizeof(double) * nq * 3);
  wk->v = malloc(sizeof(double) * nq);
  wk->dv = malloc(sizeof(double) * nq * 3);
  wk->ye = malloc(sizeof(double) * np);
  wk->kq = malloc(sizeof(double) * nq);
}
static void work_free(orc_work *wk)
{
  free(wk->xq); free(wk->Jq); free(wk->cjinv); free(wk->jw); free(wk->uq);
  free(wk->duq); free(wk->v); free(wk->dv); free(wk->ye); free(wk->kq);
}

/* Geometry link: INTERP and GRAD(jinv=NULL) of the Q1 coordinates, then dRuleComputePhysical. */
static int element_geometry(const orc_problem *pb, const orc_rule *r, int e, orc_work *wk)
{
  const double *vx = &pb->elem_vx[(long)e * 24];
  int err = efs_apply(&r->bc, NULL, 3, vx, wk->xq, ORC_APPLY_INTERP, ORC_INSERT_VALUES);
  err |= efs_apply(&r->bc, NULL, 3, vx, wk->Jq, ORC_APPLY_GRAD, ORC_INSERT_VALUES);
  if (err) return err;
  if (pb->kvert) err |= efs_apply(&r->bc, NULL, 1, &pb->kvert[(long)e * 8], wk->kq, ORC_APPLY_INTERP, ORC_INSERT_VALUES);
  else for (int q = 0; q < r->nq; q++) wk->kq[q] = 1.0;
  return err | orc_rule_compute_physical(r->nq, r->w, wk->Jq, wk->cjinv, wk->jw);
}

static void gather(const orc_problem *pb, int e, const double *xc, double *xe)
{
  const int n = pb->p + 1, np = n * n * n;
  const int *idx = &pb->elem_nodes[(long)e * np];
  for (int i = 0; i < np; i++) xe[i] = xc[idx[i]];
}

/* Deterministic E^T scatter: elements first write their expanded slices, then each
 * closure node sums its expanded entries in element order (fs.c:481-500). */
typedef struct { long *off; long *ent; } orc_transpose;
static void transpose_build(const orc_problem *pb, orc_transpose *t)
{
  const int n = pb->p + 1, np = n * n * n;
  const long nexp = (long)pb->nelem * np;
  t->off = calloc(pb->nclosure + 1, sizeof(long));
  t->ent = malloc(sizeof(long) * nexp);
  for (long i = 0; i < nexp; i++) t->off[pb->elem_nodes[i] + 1]++;
  for (long i = 0; i < pb->nclosure; i++) t->off[i + 1] += t->off[i];
  long *fill = malloc(sizeof(long) * pb->nclosure);
  memcpy(fill, t->off, sizeof(long) * pb->nclosure);
  for (long i = 0; i < nexp; i++) t->ent[fill[pb->elem_nodes[i]]++] = i;
  free(fill);
}
static void transpose_free(orc_transpose *t) { free(t->off); free(t->ent); }

static void set_threads(const orc_problem *pb)
{
#ifdef _OPENMP
  if (pb->nthreads > 0) omp_set_num_threads(pb->nthreads);
#else
  (void)pb;
#endif
}

/* mode 0: F(u) (EllipFunction, ellip.c:407-440); mode 1: J(u) x (EllipShellMatMult,
 * ellip.c:442-472) with the store recomputed from u at each point (the reference reads
 * the store that the last F(u) wrote, ellip.c:311,462 -- the same numbers). */
static int ellip_apply(const orc_problem *pb, int mode, const double *uc, const double *xc, double *yc)
{
  const int n = pb->p + 1, np = n * n * n;
  orc_rule r;
  if (rule_setup(&r, pb->p, pb->fmethod)) { rule_free(&r); return 1; }
  if (r.npatches != 1) { rule_free(&r); return 4; }
  double *yexp = malloc(sizeof(double) * (long)pb->nelem * np);
  int err = 0;
  set_threads(pb);
#pragma omp parallel reduction(| : err)
  {
    orc_work wk;
    work_alloc(&wk, r.nq, np);
    double *ue = malloc(sizeof(double) * np), *xe = malloc(sizeof(double) * np);
    double *xq2 = malloc(sizeof(double) * r.nq), *dxq2 = malloc(sizeof(double) * r.nq * 3);
#pragma omp for schedule(static)
    for (int e = 0; e < pb->nelem; e++) {
      int ee = element_geometry(pb, &r, e, &wk);
      gather(pb, e, uc, ue);
      ee |= efs_apply(&r.bu, NULL, 1, ue, wk.uq, ORC_APPLY_INTERP, ORC_INSERT_VALUES);
      ee |= efs_apply(&r.bu, wk.cjinv, 1, ue, wk.duq, ORC_APPLY_GRAD, ORC_INSERT_VALUES);
      double *uuq = wk.uq, *dduq = wk.duq;
      if (mode == 1) {
        gather(pb, e, xc, xe);
        ee |= efs_apply(&r.bu, NULL, 1, xe, xq2, ORC_APPLY_INTERP, ORC_INSERT_VALUES);
        ee |= efs_apply(&r.bu, wk.cjinv, 1, xe, dxq2, ORC_APPLY_GRAD, ORC_INSERT_VALUES);
      }
      for (int q = 0; q < r.nq; q++) {
        orc_store st;
        compute_store(pb, uuq[q], &dduq[q * 3], wk.kq[q], &st);
        if (mode == 0) {
          double f;
          orc_forcing(pb, &wk.xq[q * 3], &f);
          wk.v[q] = -wk.jw[q] * f;
          wk.v[q] += -wk.jw[q] * st.lexp;
          for (int i = 0; i < 3; i++) wk.dv[q * 3 + i] = wk.jw[q] * st.eta * dduq[q * 3 + i];
        } else {
          pointwise_jacobian(&st, wk.jw[q], xq2[q], &dxq2[q * 3], &wk.v[q], &wk.dv[q * 3]);
        }
      }
      double *ye = &yexp[(long)e * np];
      ee |= efs_apply(&r.bu, wk.cjinv, 1, wk.v, ye, ORC_APPLY_INTERP_TRANSPOSE, ORC_INSERT_VALUES);
      ee |= efs_apply(&r.bu, wk.cjinv, 1, wk.dv, ye, ORC_APPLY_GRAD_TRANSPOSE, ORC_ADD_VALUES);
      err |= ee;
    }
    free(ue); free(xe); free(xq2); free(dxq2);
    work_free(&wk);
  }
  orc_transpose t;
  transpose_build(pb, &t);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < pb->nclosure; i++) {
    double s = 0;
    for (long k = t.off[i]; k < t.off[i + 1]; k++) s += yexp[t.ent[k]];
    yc[i] = s;
  }
  transpose_free(&t);
  free(yexp);
  rule_free(&r);
  return err;
}

int orc_ellip_function(const orc_problem *pb, const double *uc, double *yc) { return ellip_apply(pb, 0, uc, NULL, yc); }
int orc_ellip_jacobian_apply(const orc_problem *pb, const double *uc, const double *xc, double *yc)
{
  return ellip_apply(pb, 1, uc, xc, yc);
}

/* EllipJacobian (ellip.c:474-536) with dRulesetIteratorGetPatchAssembly (fsrulesetit.c:657-701)
 * and dFSMatSetValuesBlockedExpanded (fs.c:826-930): one dense block per patch, mapped
 * through E, Dirichlet rows and columns dropped. */
long orc_ellip_q1_assemble(const orc_problem *pb, const double *uc, int qmethod, long cap,
                           int *rows, int *cols, double *vals)
{
  const int n = pb->p + 1, np = n * n * n;
  orc_rule r;
  if (rule_setup(&r, pb->p, qmethod)) { rule_free(&r); return -1; }
  const int npatch = r.npatches, Q = r.patchsize;
  const int identity = npatch == 1;
  long count = 0;
  orc_work wk;
  work_alloc(&wk, r.nq, np);
  double *ue = malloc(sizeof(double) * np);
  int Pmax = np;
  int *eidx = malloc(sizeof(int) * npatch * Pmax);
  double *interp = malloc(sizeof(double) * npatch * Q * Pmax), *deriv = malloc(sizeof(double) * npatch * Q * Pmax * 3);
  double *K = malloc(sizeof(double) * Pmax * Pmax);
  double *pu = malloc(sizeof(double) * r.nq), *pdu = malloc(sizeof(double) * r.nq * 3), *pjw = malloc(sizeof(double) * r.nq);
  double *pk = malloc(sizeof(double) * r.nq);
  long ret = 0;
  for (int e = 0; e < pb->nelem && ret == 0; e++) {
    int ee = element_geometry(pb, &r, e, &wk);
    gather(pb, e, uc, ue);
    ee |= efs_apply(&r.bu, NULL, 1, ue, wk.uq, ORC_APPLY_INTERP, ORC_INSERT_VALUES);
    ee |= efs_apply(&r.bu, wk.cjinv, 1, ue, wk.duq, ORC_APPLY_GRAD, ORC_INSERT_VALUES);
    if (ee) { ret = -1; break; }
    /* ValueCacheExtract / ValueCachePhysicalExtract: patch-natural ordering */
    for (int i = 0; i < r.nq; i++) {
      const int s = identity ? i : r.patchind[i];
      pu[i] = wk.uq[s];
      pjw[i] = wk.jw[s];
      pk[i] = wk.kq[s];
      for (int l = 0; l < 3; l++) pdu[i * 3 + l] = wk.duq[s * 3 + l];
    }
    int P;
    /* The solution link receives the element-natural cjinv (fsrulesetit.c:684). */
    const double *cj = wk.cjinv;
    double *pcj = NULL;
    if (orc_q1_cjinv_patch_order && !identity) {
      pcj = malloc(sizeof(double) * r.nq * 9);
      for (int i = 0; i < r.nq; i++)
        for (int l = 0; l < 9; l++) pcj[i * 9 + l] = wk.cjinv[r.patchind[i] * 9 + l];
      cj = pcj;
    }
    const int xe = orc_explicit_sparse_hex(n, r.Q1d, r.rc, npatch, Q, r.patchind, cj, 0, &P, eidx, interp, deriv);
    free(pcj);
    if (xe) { ret = -1; break; }
    for (int pt = 0; pt < npatch; pt++) {
      const double *I = &interp[(long)pt * Q * P], *Dd = &deriv[(long)pt * Q * P * 3];
      const int *rc = &eidx[pt * P];
      memset(K, 0, sizeof(double) * P * P);
      for (int q = 0; q < Q; q++) {
        const int qq = pt * Q + q;
        orc_store st;
        compute_store(pb, pu[qq], &pdu[qq * 3], pk[qq], &st);
        for (int j = 0; j < P; j++) {
          double v, dv[3];
          pointwise_jacobian(&st, pjw[qq], I[q * P + j], &Dd[(q * P + j) * 3], &v, dv);
          for (int i = 0; i < P; i++)
            K[i * P + j] += I[q * P + i] * v + Dd[(q * P + i) * 3 + 0] * dv[0] + Dd[(q * P + i) * 3 + 1] * dv[1] +
                            Dd[(q * P + i) * 3 + 2] * dv[2];
        }
      }
      const int *enodes = &pb->elem_nodes[(long)e * np];
      for (int i = 0; i < P; i++) {
        const int gi = enodes[rc[i]];
        if (gi >= pb->nowned) continue;
        for (int j = 0; j < P; j++) {
          const int gj = enodes[rc[j]];
          if (gj >= pb->nowned) continue;
          if (count < cap) { rows[count] = gi; cols[count] = gj; vals[count] = K[i * P + j]; }
          count++;
        }
      }
    }
  }
  free(ue); free(eidx); free(interp); free(deriv); free(K); free(pu); free(pdu); free(pjw); free(pk);
  work_free(&wk);
  rule_free(&r);
  if (ret) return ret;
  return count <= cap ? count : -count;
}

/* EllipErrorNorms (ellip.c:538-564) with dNormsUpdate/Finish (util.c:119-153). */
int orc_ellip_error_norms(const orc_problem *pb, const double *uc, double en[3], double gn[3])
{
  const int n = pb->p + 1, np = n * n * n;
  orc_rule r;
  if (rule_setup(&r, pb->p, pb->fmethod)) { rule_free(&r); return 1; }
  orc_work wk;
  work_alloc(&wk, r.nq, np);
  double *ue = malloc(sizeof(double) * np);
  int err = 0;
  for (int i = 0; i < 3; i++) en[i] = gn[i] = 0;
  for (int e = 0; e < pb->nelem; e++) {
    err |= element_geometry(pb, &r, e, &wk);
    gather(pb, e, uc, ue);
    err |= efs_apply(&r.bu, NULL, 1, ue, wk.uq, ORC_APPLY_INTERP, ORC_INSERT_VALUES);
    err |= efs_apply(&r.bu, wk.cjinv, 1, ue, wk.duq, ORC_APPLY_GRAD, ORC_INSERT_VALUES);
    for (int q = 0; q < r.nq; q++) {
      double uu, duu[3];
      orc_exact_solution(pb, &wk.xq[q * 3], &uu, duu);
      const double rr = uu - wk.uq[q], jw = wk.jw[q];
      en[0] += fabs(rr) * jw;
      en[1] += rr * rr * jw;
      en[2] = fmax(en[2], fabs(rr));
      const double d0 = duu[0] - wk.duq[q * 3], d1 = duu[1] - wk.duq[q * 3 + 1], d2 = duu[2] - wk.duq[q * 3 + 2];
      const double g2 = d0 * d0 + d1 * d1 + d2 * d2, ga = sqrt(g2);
      gn[0] += ga * jw;
      gn[1] += g2 * jw;
      gn[2] = fmax(gn[2], ga);
    }
  }
  en[1] = sqrt(en[1]);
  gn[1] = sqrt(gn[1]);
  free(ue);
  work_free(&wk);
  rule_free(&r);
  return err;
}

/* Element-local stiffness of independent hexahedra of per-element degree: for each element, the
 * Jacobian action of ellip at p = 2 (EllipShellMatMult, ellip.c:442-472: GRAD, Dv = jw k grad,
 * GRAD_T) on a one-element space with no Dirichlet nodes, i.e. K_e x_e.  Element slices are
 * consecutive, offsets sum (degree+1)^3 in element order (cont.c:372); elements of different
 * degree get their own rules and bases, as the mixed-order dEFS arrays of jacobi/tests/ex1.c:282-309. */
int orc_expanded_poisson_apply(int nelem, const int *degree, const double *elem_vx, const double *kvert,
                               const double *x, double *y, int nthreads)
{
  long *off = malloc(sizeof(long) * (nelem + 1));
  off[0] = 0;
  for (int e = 0; e < nelem; e++) {
    const long n = degree[e] + 1;
    off[e + 1] = off[e] + n * n * n;
  }
  int err = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 8) reduction(| : err)
  for (int e = 0; e < nelem; e++) {
    const int n = degree[e] + 1, np = n * n * n;
    int *id = malloc(sizeof(int) * np);
    double *zero = calloc(np, sizeof(double));
    for (int i = 0; i < np; i++) id[i] = i;
    orc_problem pb;
    memset(&pb, 0, sizeof(pb));
    pb.nelem = 1;
    pb.p = degree[e];
    pb.nclosure = pb.nowned = np;
    pb.elem_nodes = id;
    pb.elem_vx = &elem_vx[(long)e * 24];
    pb.ep = 2.0;
    pb.eps = 1.0;
    pb.fmethod = ORC_QUAD_FAST;
    pb.nthreads = 1;
    pb.kvert = kvert ? &kvert[(long)e * 8] : NULL;
    err |= ellip_apply(&pb, 1, zero, &x[off[e]], &y[off[e]]);
    free(id);
    free(zero);
  }
  free(off);
  return err;
}

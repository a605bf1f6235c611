/*
 * dohp_oracle.h -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * A CPU restatement, in plain C, of the arithmetic of dohp's matrix-free hp
 * element operator (jedbrown/dohp).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker or as the
 * timed CPU baseline.  The product path (dohp_amd/) never links or calls it.
 *
 * Each function cites the reference file:line it restates.  Parity status of
 * each function (what pins it) is listed in oracle/README.md.
 */
#ifndef DOHP_ORACLE_H
#define DOHP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Reference enum values (include/dohpjacobi.h:101-108, PETSc InsertMode). */
enum { ORC_APPLY_INTERP = 40001, ORC_APPLY_INTERP_TRANSPOSE, ORC_APPLY_GRAD, ORC_APPLY_GRAD_TRANSPOSE };
enum { ORC_INSERT_VALUES = 1, ORC_ADD_VALUES = 2 };
/* dQuadratureMethod (include/dohpjacobi.h:143-150) and dGaussFamily (:218). */
enum { ORC_QUAD_FAST = 0, ORC_QUAD_SPARSE = 1, ORC_QUAD_SELF = 2 };
enum { ORC_GAUSS_LEGENDRE = 0, ORC_GAUSS_LOBATTO = 1 };

/* ---- 1-D rules and bases (polylib.c, tensorquad.c, tensor.c) ---- */
int orc_gauss_legendre(int np, double *z, double *w);
int orc_gauss_lobatto(int np, double *z, double *w);
int orc_lagrange_interp_gll(int np, const double *zgll, int mz, const double *zm, double *im);
int orc_deriv_gll(int np, const double *zgll, double *D);
int orc_tensor_basis(int P, int Q, const double *rcoord, double *interp, double *deriv,
                     double *interpT, double *derivT, double *node, double *weight);
int orc_rule_1d(int method, int family, int order, int *size, double *coord, double *weight);
int orc_rule_patches(int Q1d, const double *w1d, int sparse, int *npatches, int *patchsize,
                     int *ind, double *pweight);

/* ---- Element kernels (inlinetmulthex.h, efstopo.c, jacobi.c, dohpgeom.h) ---- */
int orc_tensor_mult_hex(int D, const int P[3], const int Q[3], const double *A[3],
                        const double *in, double *out, int imode);
int orc_efs_apply_hex(int P, int Q, const double *rcoord, const double *jinv, int D,
                      const double *in, double *out, int amode, int imode);
int orc_geom_invert3(const double *a, double *b, double *det);
int orc_rule_compute_physical(int n, const double *w, const double *jac, double *jinv, double *jw);
int orc_explicit_sparse_hex(int P, int Q1d, const double *rcoord, int npatches, int Q,
                            const int *qidx, const double *cjinv, int eoffset, int *Pout,
                            int *eidx, double *interp, double *deriv);

/* ---- Whole-operator restatement (ellip.c + fsrulesetit.c + fs.c) ---- */
typedef struct {
  int nelem;               /* number of hexahedra */
  int p;                   /* polynomial degree, n = p+1 nodes per direction */
  long nclosure;           /* closure nodes: owned first, then Dirichlet */
  long nowned;             /* owned (non-Dirichlet) nodes */
  const int *elem_nodes;   /* [nelem][n^3] closure index of each element node (the E matrix) */
  const double *elem_vx;   /* [nelem][8][3] Q1 vertex coordinates in tensor order */
  double ep, eps, lambda;  /* ellip p, epsilon, Bratu lambda (ellip.c:27-32) */
  int exact;               /* exact-solution choice 0..4 for the forcing (ellip.c:42-117) */
  double ea, eb, ec;       /* exact-solution scale parameters */
  int fmethod;             /* quadrature method for residual / matrix-free (default FAST) */
  int nthreads;            /* OpenMP threads (0 = runtime default) */
  const double *kvert;     /* optional [nelem][8] coefficient at vertices (tensor order), or NULL */
} orc_problem;

int orc_exact_solution(const orc_problem *pb, const double x[3], double *u, double du[3]);
int orc_forcing(const orc_problem *pb, const double x[3], double *f);

/* F(u): uc is the closure vector (owned | Dirichlet), yc the closure output (overwritten). */
int orc_ellip_function(const orc_problem *pb, const double *uc, double *yc);
/* J(u) x: linearization state uc (inhomogeneous closure), direction xc (homogeneous closure). */
int orc_ellip_jacobian_apply(const orc_problem *pb, const double *uc, const double *xc, double *yc);
/* Embedded-Q1 matrix (ellip.c:474-536) as COO on owned nodes.  Returns the number of entries
 * written, or -needed if cap is too small.  qmethod SPARSE gives the reference default. */
long orc_ellip_q1_assemble(const orc_problem *pb, const double *uc, int qmethod, long cap,
                           int *rows, int *cols, double *vals);
void orc_set_q1_cjinv_patch_order(int flag);
/* Error norms (util.c:110-168 via ellip.c:538-564). */
int orc_ellip_error_norms(const orc_problem *pb, const double *uc, double enorm[3], double gnorm[3]);
/* Element-local Poisson stiffness K_e x_e of independent elements of per-element degree
 * (ellip.c:442-472 at p = 2 on one-element spaces; slices at offsets sum (degree+1)^3).
 * elem_vx [nelem][8][3] tensor order, kvert [nelem][8] or NULL. */
int orc_expanded_poisson_apply(int nelem, const int *degree, const double *elem_vx, const double *kvert,
                               const double *x, double *y, int nthreads);

#ifdef __cplusplus
}
#endif
#endif

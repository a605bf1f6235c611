/*
 * dohp_amd.h -- C ABI of the MI355X matrix-free hp element operator.
 *
 * Drop-in boundary for dohp's operator path (jedbrown/dohp).  Every entry point
 * returns a dErr (0 = success, nonzero = error; dohpGetLastError gives the text),
 * never aborts, and takes plain pointers and sizes.  Pointers named *_dev are
 * device (HBM) pointers; *_host are host pointers.  `stream` is a hipStream_t
 * (NULL = the default stream).  Callers own every array they pass in.
 *
 * Layouts follow the reference (Appendix conventions of dohp):
 *   element node index (i*n + j)*n + k, k fastest;  quadrature index likewise;
 *   block size D interleaved fastest;  GRAD output [Q][D][3];
 *   jinv[q][k][e] = dX_k/dx_e;  closure vectors are [owned | Dirichlet].
 */
#ifndef DOHP_AMD_H
#define DOHP_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int dErr;

/* dApplyMode (include/dohpjacobi.h:101-108) and PETSc InsertMode values. */
enum { dAPPLY_INTERP = 40001, dAPPLY_INTERP_TRANSPOSE, dAPPLY_GRAD, dAPPLY_GRAD_TRANSPOSE };
enum { dINSERT_VALUES = 1, dADD_VALUES = 2 };
/* dFSHomogeneousMode (include/dohpfs.h:49). */
enum { dFS_HOMOGENEOUS = 0, dFS_INHOMOGENEOUS = 1 };
/* dQuadratureMethod (include/dohpjacobi.h:143-150) and dGaussFamily (:218). */
enum { dQUADRATURE_METHOD_FAST = 0, dQUADRATURE_METHOD_SPARSE = 1, dQUADRATURE_METHOD_SELF = 2 };
enum { dGAUSS_LEGENDRE = 0, dGAUSS_LOBATTO = 1 };
/* Precision of an operator application. */
enum { dohpPREC_F64 = 0, dohpPREC_F32 = 1 };
/* Error codes beyond generic failure. */
/* dohpERR_NAN: a non-finite value met on the device (input vector, state, result); dohpERR_OOM: a device
 * allocation failed (the object stays usable: a later call allocates again). */
enum { dohpERR_GENERIC = 1, dohpERR_ARG = 62, dohpERR_NEGATIVE_JACOBIAN = 77, dohpERR_NAN = 78, dohpERR_OOM = 98,
       dohpERR_HIP = 99 };

/* Copies the message of the last error on this thread into buf. */
void dohpGetLastError(char *buf, size_t len);

/* ------------------------------------------------------------------ dJacobi */
/* 1-D rule of TensorGetRule (src/jacobi/quadrature/impls/tensor/tensorquad.c:51-118):
 * method FAST/SPARSE/SELF, family Gauss/Lobatto, polynomial order.  coord/weight need
 * room for order+4 entries; *size receives the rule size. */
dErr dohpTensorRule(int method, int family, int order, int *size, double *coord_host, double *weight_host);

/* Nodal GLL basis of TensorBasisCreate (src/jacobi/impls/tensor/tensor.c:22-109):
 * interp[Q][P] = l_j(rcoord_q), deriv = interp * D_gll, plus GLL nodes and weights. */
dErr dohpTensorBasis(int P, int Q, const double *rcoord_host, double *interp_host, double *deriv_host,
                     double *node_host, double *weight_host);

/* --------------------------------------------------------------------- dEFS */
/* Batched dEFSApply on isotropic hexahedra (replaces dEFSApply, src/jacobi/interface/jacobi.c:489-506,
 * dispatched to dEFSApply_Tensor_Hex, src/jacobi/impls/tensor/efstopo.c:368-423).
 * nelem elements, each with in/out slices of the per-mode sizes (INTERP: P^3*D -> Q^3*D,
 * GRAD: P^3*D -> Q^3*D*3, transposes the reverse), consecutive per element.  jinv_dev is
 * [nelem][Q^3][9] or NULL (GRAD only: layout swap, efstopo.c:518-546).  rcoord is the 1-D rule. */
dErr dEFSApplyHexBatched(int P, int Q, const double *rcoord_host, int nelem, const double *jinv_dev, int D,
                         const double *in_dev, double *out_dev, int amode, int imode, void *stream);

/* No counterpart in the reference.  dEFSApplyHexBatched without any device allocation or wait, for callers
 * that cannot allocate (code under HIP graph capture, tight loops applying elements the way the reference's
 * per-element dEFSApply call sites do).  Same arguments, results and errors as dEFSApplyHexBatched (it shares its
 * kernel); P, Q <= 12.  The B and D tables (2 Q P <= 288 doubles) are staged in a 2.3 KB device array of the
 * library, one per GPU, held until the process ends: calls that overlap on one GPU (several streams or host
 * threads) race on it. */
dErr dEFSApplyHexBatchedNoAlloc(int P, int Q, const double *rcoord_host, int nelem, const double *jinv_dev, int D,
                                const double *in_dev, double *out_dev, int amode, int imode, void *stream);

/* Batched dRuleComputePhysical (include/dohpjacobi.h:193, src/jacobi/interface/jacobi.c:383-414) for a
 * tensor rule of Q points per direction, 1-D weights weight_host[Q], point order (i Q + j) Q + k
 * (jacobi.c:351): for each of nelem elements and each point q, jinv[q] = jac[q]^-1 (3x3 row-major
 * cofactor inverse, dGeomInvert3, include/dohpgeom.h:57-75) and jw[q] = w_i w_j w_k det jac[q].
 * jac_dev and jinv_dev are [nelem][Q^3][9], jw_dev [nelem][Q^3].  A negative determinant returns
 * dohpERR_NEGATIVE_JACOBIAN (jacobi.c:397); a non-finite one returns dohpERR_NAN (the reference
 * carries it on).  Waits for the stream (the device-detected flag is read back). */
dErr dRuleComputePhysicalBatched(int Q, const double *weight_host, int nelem, const double *jac_dev,
                                 double *jinv_dev, double *jw_dev, void *stream);

/* ---------------------------------------------------------------------- dFS */
typedef struct _p_dohpFS *dohpFS;

/* Continuous degree-p space on an M0 x M1 x M2 hexahedral block (replaces dFSCreate +
 * dFSSetMesh/dFSSetDegree/dFSRegisterBoundary/dFSBuildSpace, src/fs/interface/fs.c:11-110,423,
 * src/fs/impls/cont/cont.c:258-385).  vcoords_host: [(M0+1)(M1+1)(M2+1)][3] vertex coordinates in
 * block-generator order (src/fs/mesh/interface/genblock.c:206-227).  dirichlet_mask: bit 1 = walls
 * (set 100), 2 = top (200), 4 = bottom (300).  Closure numbering: owned nodes first, then Dirichlet. */
dErr dFSCreateBlock(int M0, int M1, int M2, int degree, int dirichlet_mask, const double *vcoords_host,
                    dohpFS *fs);
/* Continuous degree-p space on any conforming hexahedral mesh, from explicit maps (the conforming case
 * of dFSBuildSpace_Cont, src/fs/impls/cont/cont.c:258-385, whose element assembly matrix E,
 * cont.c:200-256, is then a gather by index).  elem_nodes_host: [nelem][n^3] closure index of node
 * (a, b, c) at (a n + b) n + c, n = degree + 1 (owned nodes 0..nowned-1, Dirichlet nodes nowned..nclosure-1,
 * every closure node in some element); elem_verts_host: [nelem][8] Q1 vertex of corner (a, b, c) at
 * (a 2 + b) 2 + c; vcoords_host: [nvert][3].  The operator's kernels read the maps (the block form
 * evaluates them); the embedded-Q1 matrix, its multigrid and the persistent launch shapes need a block. */
dErr dFSCreateFromElements(int degree, int nelem, long nowned, long nclosure, const int *elem_nodes_host,
                           long nvert, const double *vcoords_host, const int *elem_verts_host, dohpFS *fs);
dErr dFSDestroy(dohpFS *fs);
/* Sizes: owned (global) nodes, closure nodes, expanded nodes (nelem * n^3), elements. */
dErr dFSGetSizes(dohpFS fs, long *nowned, long *nclosure, long *nexpanded, int *nelem);
/* Closure nodal coordinates averaged over elements (src/fs/interface/fsgeom.c:72-109), [nclosure][3]. */
dErr dFSGetNodalCoordinatesClosure(dohpFS fs, double *coords_host);
/* Cache Dirichlet values (replaces dFSInhomogeneousDirichletCommit, src/fs/interface/fs.c:523-537);
 * dvals_host holds the (nclosure - nowned) Dirichlet entries of the closure. */
dErr dFSInhomogeneousDirichletCommit(dohpFS fs, const double *dvals_host);
/* Gather/scatter of src/fs/interface/fs.c:543-640 on device vectors (FP64).  g is the global
 * (owned) vector, x the expanded vector [nelem][n^3]. */
dErr dFSGlobalToExpanded(dohpFS fs, const double *g_dev, double *x_dev, int hmode, int imode, void *stream);
dErr dFSExpandedToGlobal(dohpFS fs, const double *x_dev, double *g_dev, int hmode, int imode, void *stream);
/* Element -> closure node map [nelem][n^3] (the E matrix, src/fs/impls/cont/cont.c:200-249). */
dErr dFSGetElementNodes(dohpFS fs, int *elem_nodes_host);

/* ------------------------------------------------------------ ellip operator */
typedef struct {
  double p;        /* p of the p-Laplacian (ellip.c:27-32, -ellip_p) */
  double epsilon;  /* regularization (-ellip_eps) */
  double lambda;   /* Bratu nonlinearity (-ellip_lam) */
  int exact;       /* forcing of exact solution 0..4 (ellip.c:42-117, -exact); -1 = f = 0 */
  double a, b, c;  /* exact-solution parameters (-exact_a/b/c) */
} dohpEllipParam;

typedef struct _p_dohpEllip *dohpEllip;

/* Operator context: the MATSHELL context of ellip (src/fs/tests/ellip.c:697-699).  Validates the
 * geometry: a nonpositive Jacobian determinant at any quadrature point returns
 * dohpERR_NEGATIVE_JACOBIAN (src/jacobi/interface/jacobi.c:397). */
dErr dohpEllipCreate(dohpFS fs, const dohpEllipParam *prm, dohpEllip *op);
dErr dohpEllipDestroy(dohpEllip *op);
/* Vertex coefficient field k (length = number of vertices, genblock order), multiplying the
 * diffusion: -div(k eta grad u).  NULL restores k = 1. */
dErr dohpEllipSetCoefficient(dohpEllip op, const double *kvert_host);
/* Residual F(u) (EllipFunction, src/fs/tests/ellip.c:407-440) with inhomogeneous Dirichlet values;
 * records u as the linearization state of later Jacobian applications (the store of ellip.c:311).
 * With checks on (dohpEllipSetChecks, default) the residual and the MatMult entry points wait for their
 * stream and return dohpERR_NAN when the element kernels met a non-finite contribution (a NaN or Inf in
 * the input vector or the state). */
dErr dohpEllipFunction(dohpEllip op, const double *u_dev, double *f_dev, void *stream);
/* y = J(u) x, homogeneous Dirichlet (EllipShellMatMult, src/fs/tests/ellip.c:442-472): the MATSHELL
 * MATOP_MULT body.  FP64 and FP32 variants; the FP32 one keeps geometry in FP64. */
dErr dohpEllipShellMatMult(dohpEllip op, const double *x_dev, double *y_dev, void *stream);
dErr dohpEllipShellMatMultF32(dohpEllip op, const float *x_dev, float *y_dev, void *stream);
/* y += J(u) x: MATOP_MULT_ADD with v2 = v3 for the same MATSHELL (ellip.c:442-472 sets MATOP_MULT only,
 * so PETSc forms MatMultAdd there as MatMult into a work vector plus VecAXPY; this adds in the scatter).
 * Needs the atomic scatter (dohpERR_ARG under the colour scatter); checked like MatMult. */
dErr dohpEllipShellMatMultAdd(dohpEllip op, const double *x_dev, double *y_dev, void *stream);
/* 1 (default): the residual and MatMult entry points wait for their stream and report non-finite values
 * (dohpERR_NAN); 0: they return at once, unchecked (asynchronous use; the library's solvers apply the
 * operator unchecked and test the FP64 residual norm instead, returning dohpERR_NAN if it is not finite). */
dErr dohpEllipSetChecks(dohpEllip op, int mode);
/* Scatter strategy of the element kernels: 0 = eight-colour deterministic (bitwise reproducible),
 * 1 = hardware atomics (default; summation order varies at rounding level). */
dErr dohpEllipSetScatter(dohpEllip op, int mode);
/* How y = J x (and the linear operator) are applied.  1: the first application after each residual
 * (after creation, or after dohpEllipSetCoefficient, for the linear operator) stores per quadrature
 * point the symmetric 3x3 tensor G with d_m = sum_n G_mn dx/dX_n (weights, inverse Jacobians, eta and
 * the rank-one term folded in) and, when lambda != 0, the Bratu factor: 48 bytes per point in FP64
 * (56 with lambda), plus half that once an FP32 application asks for it; applications then read it
 * and recompute no geometry.  This is the reference's residual-to-Jacobian stash (struct EllipStore,
 * src/fs/tests/ellip.c:119-123, 311, 460-467) with the geometry folded in.  0: recompute the state's
 * gradient and the geometry at every application (no storage; the memory is returned).  2 (default):
 * the first application in each precision times both forms (each with its tuned launch shape) and
 * keeps the faster per application (the recomputed one when the data does not fit in device memory);
 * the memory is returned when neither precision keeps the stored form.  That first application waits
 * for its stream (it times launches), as the launch-shape tuner's first application of a mode does.  At 64^3 hexahedra, p = 6 the nonlinear Jacobian keeps the stored form and the linear operator
 * (p = 2, lambda = 0) recomputes. */
dErr dohpEllipSetStorage(dohpEllip op, int mode);
/* The form that applications in a precision use (1 = stored quadrature data, 0 = recomputed, -1 = with
 * storage 2, not timed yet) and, with storage 2, the timing behind the choice: ms[0] recomputed, ms[1]
 * from stored data (milliseconds per application; 0 until timed). */
dErr dohpEllipGetStorage(dohpEllip op, int fp32, int *stored, float *ms);
/* Launch shape the autotuner chose on the first application of a mode (0 = linear, 1 = residual,
 * 2 = Jacobian action recomputed from the state, 3 = application from stored quadrature data, linear
 * or Jacobian) and precision: -1 = not tuned yet, 0 = several elements per four-wavefront
 * workgroup, 1 = one-wavefront workgroups, 2 = persistent one-wavefront workgroups with the next
 * batch's loads in flight, 3 = persistent four-wavefront workgroups, 4 = four-wavefront workgroups with
 * the one-dimensional contractions as 16x16x4 MFMA products in the kernel's precision (degrees 3..7;
 * no counterpart in the reference, whose TensorMult_Hex loops, src/jacobi/impls/tensor/inlinetmulthex.h,
 * this shape replaces; timed by the tuner in FP32 only); ms[5] receives the measured milliseconds of each
 * (0 for shapes not timed). */
dErr dohpEllipGetTuning(dohpEllip op, int mode, int fp32, int *shape, float *ms);
/* Fixes the launch shape of a mode and precision (-1 = tune again on the next application). */
dErr dohpEllipSetShape(dohpEllip op, int mode, int fp32, int shape);

/* ------------------------------------------- batched mixed-order Poisson operator */
typedef struct _p_dohpBatch *dohpBatch;

/* A batch of linear Poisson operators -div(k grad u) applied together in one persistent launch
 * (elements bucketed by order, highest first).  Two kinds of segment share one packed vector:
 *  - block spaces: y = A x on the owned nodes of a dFSCreateBlock space, A the homogeneous-
 *    Dirichlet stiffness (EllipShellMatMult at ellip p = 2, src/fs/tests/ellip.c:442-472);
 *  - element sets: independent hexahedra of per-element degree 1..10 with their own n^3 GLL nodes,
 *    y_e = K_e x_e (dEFSApply GRAD -> jw k grad -> GRAD_T per element,
 *    src/jacobi/impls/tensor/efstopo.c:368-423, jacobi.c:383-414), the element-local form of the
 *    mixed-order dEFS arrays of src/jacobi/tests/ex1.c:282-309 (tensor.c:417 rejects continuity
 *    across order jumps).  Element slices follow in element order (offsets sum n_e^3,
 *    src/fs/impls/cont/cont.c:372).
 * Block spaces come first in the packed vector, element sets after them, each in the order added;
 * dohpBatchGetSegment gives the offsets.  The segments' data are copied at Add time. */
dErr dohpBatchCreate(dohpBatch *batch);
dErr dohpBatchDestroy(dohpBatch *batch);
/* Adds the space's operator; kvert_host: k at the space's vertices (NULL = 1). */
dErr dohpBatchAddSpace(dohpBatch batch, dohpFS fs, const double *kvert_host, int *segment);
/* Adds nelem hexahedra: degree_host[nelem] in 1..10, elem_vx_host [nelem][8][3] vertex coordinates
 * in tensor order (src/fs/interface/fsgeom.c:151), kvert_host [nelem][8] or NULL (k = 1).
 * A nonpositive Jacobian determinant at a Gauss point returns dohpERR_NEGATIVE_JACOBIAN. */
dErr dohpBatchAddElements(dohpBatch batch, int nelem, const int *degree_host, const double *elem_vx_host,
                          const double *kvert_host, int *segment);
/* Freezes the batch (implicit in the calls below): packed layout, order buckets, tables. */
dErr dohpBatchSetUp(dohpBatch batch);
dErr dohpBatchGetSizes(dohpBatch batch, long *n, int *nsegments);
dErr dohpBatchGetSegment(dohpBatch batch, int segment, long *offset, long *length);
/* y = A x over every segment of the packed vectors in one persistent launch.  The first call in a
 * precision times the two tiles of each order present (elements packed over a four-wavefront
 * workgroup, or per wavefront) on y and keeps the faster.  FP32 keeps the geometry in FP64. */
dErr dohpBatchMult(dohpBatch batch, const double *x_dev, double *y_dev, void *stream);
dErr dohpBatchMultF32(dohpBatch batch, const float *x_dev, float *y_dev, void *stream);
/* MFMA contractions for FP32 applies from degree 4 up (no reference counterpart): passes = 0 keeps the
 * FP32 VALU contractions (default); 1 offers the tuner a tile whose contractions are 16x16x16 BF16
 * MFMAs with FP32 accumulation (relative error ~1e-3); 3 splits both operands into two BF16 parts and
 * sums three MFMAs per tile (~1e-5).  Changing it retunes the FP32 tiles on the next application. */
dErr dohpBatchSetMFMA(dohpBatch batch, int passes);
/* Fixes the tile of a degree and precision instead of timing it (-1 = tune again); an unavailable tile
 * is an error at the next application. */
dErr dohpBatchSetTile(dohpBatch batch, int degree, int fp32, int tile);
/* Tile chosen for a degree (-1 = none yet, 0 = workgroup-packed, 1 = wavefront-packed, 2 = MFMA) and
 * the milliseconds measured for each, ms[3] (-1 = not measured). */
dErr dohpBatchGetTuning(dohpBatch batch, int degree, int fp32, int *tile, float *ms);

/* -------------------------------------------------- embedded-Q1 preconditioner */
typedef struct _p_dohpMat *dohpMat;

/* Assemble the embedded-Q1 matrix on the GLL sub-cells (EllipJacobian, src/fs/tests/ellip.c:474-536,
 * with dRulesetIteratorGetPatchAssembly, src/fs/interface/fsrulesetit.c:657-701, and
 * dEFSGetExplicitSparse_Tensor_Hex, src/jacobi/impls/tensor/efstopo.c:660-717) at the state u
 * (owned vector, Dirichlet values from the cache; NULL = the state of the last function
 * evaluation) into a device CSR over owned nodes.  *mat == NULL creates the matrix (pattern:
 * 27 lattice neighbours, dFSCreateMatrix, src/fs/interface/fs.c:770-813); otherwise it is
 * refilled in place (SAME_NONZERO_PATTERN). */
dErr dohpEllipAssembleQ1(dohpEllip op, const double *u_dev, dohpMat *mat, void *stream);
/* Arithmetic of later Q1 assemblies (no reference counterpart): 0 = the reference's (default; sub-cell
 * derivatives +-1 per element-reference unit, efstopo.c:665-706, and the element-natural inverse
 * Jacobian index of fsrulesetit.c:684), 1 = corrected: derivatives +-1/h of the GLL sub-interval and the
 * inverse Jacobian at each point, i.e. the Q1 finite-element stiffness of the GLL sub-mesh. */
dErr dohpEllipSetQ1Mode(dohpEllip op, int mode);
/* Learned sub-cell derivative scales of later embedded-Q1 assemblies (no counterpart in the reference): the
 * sub-cell derivatives along the element's three reference directions (+-1 in the reference, efstopo.c:665-666)
 * are multiplied by scales_dev[cell][3], cell = element * p^3 + (a p + b) p + c for GLL sub-cell (a, b, c).
 * The array is the caller's and is read by every later assembly until replaced; NULL = none. */
dErr dohpEllipSetQ1CellScales(dohpEllip op, const float *scales_dev);
/* Number of GLL sub-cells, nelem * p^3. */
dErr dohpQ1CellCount(dohpEllip op, long *ncell);
/* Features of every sub-cell for the learned scales, X [ncell][dohpMLPNumInputsKind(1)]: log widths of its three
 * reference sub-intervals, its centre (element-reference coordinates), log physical lengths of the reference
 * directions at the centre minus their mean, log det J minus their sum, log k at the centre minus the element's
 * mean log k.  q1 must have been assembled by this operator (its 1-D tables are used). */
dErr dohpQ1CellFeatures(dohpEllip op, dohpMat q1, float *X_dev, void *stream);
/* g[cell][a] += coef * w^T (dK / dsigma_{cell,a}) z: gradient of the bilinear form of the linear operator's
 * (ellip p = 2, lambda = 0) embedded-Q1 matrix in the sub-cell scales set on op; w, z over the owned nodes. */
dErr dohpQ1CellScaleGradient(dohpEllip op, dohpMat q1, const double *w_dev, const double *z_dev, double coef,
                             float *g_dev, void *stream);
dErr dohpMatGetSizes(dohpMat A, long *nrows, long *nnz);
/* Copies the CSR arrays to host buffers (rowptr: nrows+1, col/val: nnz); any may be NULL. */
dErr dohpMatGetCSR(dohpMat A, int *rowptr_host, int *col_host, double *val_host);
/* y = A x in FP64, and in FP32 with the values rounded once at assembly (MatMult). */
dErr dohpMatMult(dohpMat A, const double *x_dev, double *y_dev, void *stream);
dErr dohpMatMultF32(dohpMat A, const float *x_dev, float *y_dev, void *stream);
dErr dohpMatGetDiagonal(dohpMat A, double *d_dev, void *stream);
dErr dohpMatDestroy(dohpMat *A);

/* ------------------------------------------------------------ Q1 multigrid */
typedef struct _p_dohpMG *dohpMG;

/* Geometric multigrid on the embedded-Q1 matrix of a block space (the multilevel preconditioner
 * ellip's driver takes from PETSc for its Q1 matrix, src/fs/tests/ellip.c:594-652,701-747).
 * Levels halve the owned-node box, Galerkin coarse operators P^T A P, Jacobi-Chebyshev smoothing
 * of the given degree on [0.1, 1.1] * lmax(D^-1 A), explicit inverse on a coarsest level of at
 * most max_coarse nodes (0 = 512).  The V-cycle is symmetric (usable inside CG).  One apply at a
 * time per object: the levels' work vectors are shared. */
dErr dohpMGCreate(dohpMat q1, int smooth_degree, int max_coarse, dohpMG *mg, void *stream);
/* Recompute every level of mg from q1 refilled with new values on the same box (a Newton step's re-assembly), in
 * the hierarchy's own buffers: no allocation (no counterpart in the reference, which rebuilds its PC). */
dErr dohpMGRefresh(dohpMG mg, dohpMat q1, void *stream);
/* V-cycles as HIP graphs (no counterpart in the reference): with on = 1 (the default) the second consecutive
 * application to the same vectors captures the V-cycle and later ones replay it (the coarse levels' kernels are
 * shorter than their launches); 0 launches every kernel eagerly. */
dErr dohpMGSetGraphs(dohpMG mg, int on);
dErr dohpMGGetInfo(dohpMG mg, int *nlevels, long *ncoarse, double *lmax_fine);
/* Box dimensions and 27-point stencil [27][n] (FP64) of level `level` (0 = the Q1 matrix). */
dErr dohpMGGetLevel(dohpMG mg, int level, int *box, double *stencil_host);
dErr dohpMGApply(dohpMG mg, const double *r_dev, double *z_dev, void *stream);
dErr dohpMGApplyF32(dohpMG mg, const float *r_dev, float *z_dev, void *stream);
/* r = b - A_l x on level `level` (b may be NULL: r = -A_l x), the residual the V-cycle's smoothers
 * and restriction use (in FP32 with the level's FP32 coefficients); the counterpart of PETSc's
 * MatResidual on the level operator.  r must not alias b or x. */
dErr dohpMGResidual(dohpMG mg, int level, const double *b_dev, const double *x_dev, double *r_dev, void *stream);
dErr dohpMGResidualF32(dohpMG mg, int level, const float *b_dev, const float *x_dev, float *r_dev, void *stream);
/* The FP32 V-cycle's coefficients (no counterpart in the reference, whose preconditioner is FP64): fp16 = 1 (the
 * default) reads the symmetrically scaled off-diagonals of every smoothed level in FP16 and a diagonal that keeps the
 * FP64 row sums; fp16 = 0 reads FP32 copies.  dohpMGResidualF16 is the residual on the FP16 coefficients (levels
 * above the coarsest). */
dErr dohpMGSetCoefficients(dohpMG mg, int fp16);
dErr dohpMGResidualF16(dohpMG mg, int level, const float *b_dev, const float *x_dev, float *r_dev, void *stream);
dErr dohpMGDestroy(dohpMG *mg);

/* ------------------------------------------- learned correction of the Q1 preconditioner */
/* No counterpart in the reference.  A 3-layer perceptron (32 -> 64 -> 64 -> 8, SiLU, bfloat16
 * weights, FP32 accumulation) per disjoint 2x2x2 patch of the owned-node box maps patch features
 * (the Q1 diagonal's log pattern and each node's position inside its element) to log-weights of
 * the patch nodes, s = exp(clamp(o, -4, 4) / 2) (1 off the patches).  The multigrid output is
 * corrected to z = S M(S r): linear, symmetric, positive definite, so CG stays valid.  Training
 * minimizes the relative energy error ||e - a z||_J / ||e||_J (a one scale shared by a set of
 * pairs) over residual/error pairs (r, e) of the operator J, with the closed-form gradient
 * dE/ds = -2a [(J d) .* M(S r) + r .* M(S J d)] / ||e||_J^2, d = e - a z.  Parameters are one
 * FP32 vector [W1 64x32 | b1 | W2 64x64 | b2 | W3 8x64 | b3] (row-major, dohpMLPNumParams()
 * entries); the network uses their bfloat16 values. */
typedef struct _p_dohpMLP *dohpMLP;
int dohpMLPNumParams(void);
int dohpMLPNumInputs(void);
/* params_host NULL = all zeros (all weights 1: the plain preconditioner). */
dErr dohpMLPCreate(const float *params_host, dohpMLP *mlp);
/* Network kinds: 0 = the patch node weights above; 1 = the sub-cell scales of the embedded-Q1 matrix, the same
 * 3-layer perceptron (11 -> 64 -> 64 -> 3, SiLU, bfloat16 weights, FP32 accumulation) mapping the features of a
 * GLL sub-cell (dohpQ1CellFeatures) to sigma = exp(clamp(o, -4, 4)) per reference direction.  Zero parameters
 * give sigma = 1: the plain (reference) matrix.  Sizes per kind (-1 for an unknown kind): */
int dohpMLPNumParamsKind(int kind);
int dohpMLPNumInputsKind(int kind);
int dohpMLPNumOutputsKind(int kind);
dErr dohpMLPCreateKind(int kind, const float *params_host, dohpMLP *mlp);
/* Sub-cell scales sigma [ncell][3] of a kind-1 network from the features X [ncell][inputs]. */
dErr dohpMLPCellScales(dohpMLP mlp, long ncell, const float *X_dev, float *scales_dev, void *stream);
/* One Adam step (beta 0.9/0.999) of a kind-1 network from dL/dsigma [ncell][3] (backpropagated through the
 * exponential, the clamp and the network). */
dErr dohpMLPCellStep(dohpMLP mlp, long ncell, const float *X_dev, const float *gsig_dev, float lr, void *stream);
/* FP32 master parameters and the bfloat16 values the network uses (either may be NULL). */
dErr dohpMLPGetParams(dohpMLP mlp, float *master_host, float *bf16_values_host);
dErr dohpMLPDestroy(dohpMLP *mlp);
dErr dohpMLPNumPatches(dohpMat q1, long *npatch);
/* Patch features X [npatch][32] of the Q1 matrix of space fs (dinv: its inverse diagonal). */
dErr dohpMLPFeatures(dohpMat q1, dohpFS fs, const double *dinv_dev, float *X_dev, void *stream);
/* Node weights s (FP64, and FP32 if s32_dev is not NULL) over the owned nodes. */
dErr dohpMLPWeights(dohpMLP mlp, dohpMat q1, const float *X_dev, double *s_dev, float *s32_dev, void *stream);
/* Mean-squared-error gradient of the raw outputs over the first B rows of X against T [B][8]
 * (checks the training kernels). */
dErr dohpMLPGradient(dohpMLP mlp, int B, const float *X_dev, const float *T_dev, float *grad_host, double *loss,
                     void *stream);
/* Energy error E = ||e - a z||_J^2 / ||e||_J^2 of z = S M(S r) for the pair (r, e) at node
 * weights s, a = alpha if alpha > 0 else the pair's optimal scale, and its gradient dE/ds into
 * gs_dev (may be NULL).  terms (may be NULL) receives (z, r), (z, J z), (e, J e), from which the
 * optimal scale shared by several pairs follows. */
dErr dohpMLPEnergyWeightGradient(dohpEllip op, dohpMG mg, long n, const double *s_dev, const double *r_dev,
                                 const double *e_dev, double alpha, double *gs_dev, double *energy, double *terms,
                                 void *stream);
/* One Adam step (beta 0.9/0.999) from the gradient gs_dev = dL/ds of a loss L in the node
 * weights (backpropagated through the network). */
dErr dohpMLPWeightStep(dohpMLP mlp, dohpMat q1, const float *X_dev, const double *gs_dev, float lr, void *stream);
/* One Adam step (beta 0.9/0.999) on E for the pair (r, e) with scale alpha (as above); energy
 * receives E before the step. */
dErr dohpMLPEnergyStep(dohpMLP mlp, dohpEllip op, dohpMat q1, dohpMG mg, const float *X_dev, const double *r_dev,
                       const double *e_dev, double alpha, float lr, double *energy, void *stream);

/* ------------------------------------------------------------------ solvers */
typedef struct {
  int mixed;          /* 0: plain FP64 PCG; 1: FP32 inner PCG inside FP64 defect correction;
                         2: FP64 PCG with the preconditioner applied in FP32 */
  double rtol;        /* relative residual target ||b - A x|| <= rtol ||b|| (FP64 residual) */
  int max_it;         /* PCG iteration limit (per inner solve when mixed) */
  double inner_rtol;  /* inner FP32 PCG tolerance (mixed) */
  int max_outer;      /* defect-correction sweeps limit (mixed) */
  int cheb_degree;    /* degree of the Jacobi-Chebyshev iteration on the Q1 matrix */
  double cheb_ratio;  /* Chebyshev interval [lmax/ratio, lmax] of D^{-1} A_q1 */
  int eig_its;        /* power iterations for lmax */
  int pc;             /* 0: Jacobi-Chebyshev on the Q1 matrix; 1: multigrid V-cycle on it (dohpMG) */
  int mg_degree;      /* Chebyshev degree of the multigrid smoother (pre and post).  0: 1 on a corrected
                         embedded-Q1 matrix (dohpEllipSetQ1Mode 1; 48^3, p = 5: 0.08 s against 0.14 s at
                         degree 2) and on one assembled with learned sub-cell scales (48^3, p = 6: 42 iterations
                         against 51), 2 on the reference arithmetic's, where degree 1 with an FP32
                         preconditioner can stall (48^3, p = 6: 400 iterations without reaching 1e-10, against 100) */
  int fallback_its;   /* learned solve: iterations after which the plain preconditioner takes over (0 = never) */
  int mg_coeffs;      /* coefficients of the FP32 V-cycle (mixed = 1, 2): 0 = FP16 off-diagonals with a
                         row-sum-preserving FP32 diagonal (dohpMGSetCoefficients), 1 = FP32 */
} dohpSolveOpts;

typedef struct {
  int outer_its, inner_its;
  double bnorm, rnorm, lmax;
  double setup_seconds, seconds;
  int fallback;       /* learned solve: 1 if the plain preconditioner took over */
} dohpSolveStats;

/* Solve J(u) x = b (the linear operator of the ellip context; x holds the initial guess) by CG
 * preconditioned with the embedded-Q1 matrix under Jacobi-Chebyshev smoothing -- the KSP/PC
 * role of ellip's driver (src/fs/tests/ellip.c:594-652,701-747).  Krylov dot products and the
 * outer residual accumulate in FP64.  With pc = 1 the multigrid hierarchy is built by the first
 * solve and kept with q1 until q1 is refilled (stats->setup_seconds includes it when built). */
dErr dohpEllipSolve(dohpEllip op, dohpMat q1, const double *b_dev, double *x_dev, const dohpSolveOpts *opts,
                    dohpSolveStats *stats, void *stream);
/* The same with the multigrid preconditioner corrected by the learned node weights (no
 * counterpart in the reference): z = S M(S r).  After opts->fallback_its iterations without
 * convergence the plain multigrid preconditioner takes over (CG restarts from the current
 * iterate). */
dErr dohpEllipSolveLearned(dohpEllip op, dohpMat q1, dohpMLP mlp, const double *b_dev, double *x_dev,
                           const dohpSolveOpts *opts, dohpSolveStats *stats, void *stream);
/* The same CG with the multigrid preconditioner on q1 (e.g. the embedded-Q1 matrix assembled with learned sub-cell
 * scales) until opts->fallback_its iterations have passed without convergence; then the multigrid of q1_fallback
 * (the plain matrix) takes over and CG restarts from the current iterate (no counterpart in the reference). */
dErr dohpEllipSolveFallback(dohpEllip op, dohpMat q1, dohpMat q1_fallback, const double *b_dev, double *x_dev,
                            const dohpSolveOpts *opts, dohpSolveStats *stats, void *stream);
/* Plain FP64 solve that also records the iterate after every `every`-th iteration into
 * snap_dev [nsnap][nowned] (the residual/error pairs of the learned correction's training). */
dErr dohpEllipSolveRecord(dohpEllip op, dohpMat q1, const double *b_dev, double *x_dev, const dohpSolveOpts *opts,
                          dohpSolveStats *stats, double *snap_dev, int nsnap, int every, int *nsnapped, void *stream);

#ifdef __cplusplus
}
#endif
#endif

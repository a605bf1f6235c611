"""ctypes binding of the dohp_amd HIP library (host-side mirror of the C ABI).

The compute path is the HIP library dohp_amd/libdohp_amd.so (built from
dohp_amd/csrc by __graft_entry__.build()).  This module only binds its C ABI
(include/dohp_amd.h) with ctypes and offers thin objects shaped like dohp's
dFS / dEFS / MATSHELL interfaces.  Device arrays are passed as raw pointers;
tests and benchmarks allocate them with torch.  There is no CPU fallback: if the
library is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.environ.get("DOHP_AMD_LIB", os.path.join(_HERE, "libdohp_amd.so"))

APPLY_INTERP, APPLY_INTERP_TRANSPOSE, APPLY_GRAD, APPLY_GRAD_TRANSPOSE = 40001, 40002, 40003, 40004
INSERT_VALUES, ADD_VALUES = 1, 2
HOMOGENEOUS, INHOMOGENEOUS = 0, 1
QUAD_FAST, QUAD_SPARSE, QUAD_SELF = 0, 1, 2
GAUSS_LEGENDRE, GAUSS_LOBATTO = 0, 1
ERR_ARG, ERR_NEGATIVE_JACOBIAN, ERR_NAN, ERR_OOM, ERR_HIP = 62, 77, 78, 98, 99

if not os.path.exists(LIBPATH):
    raise ImportError(f"dohp_amd: HIP library not built ({LIBPATH}); run __graft_entry__.build()")
_lib = ctypes.CDLL(LIBPATH)

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_long = ctypes.c_long


class EllipParam(ctypes.Structure):
    _fields_ = [("p", ctypes.c_double), ("epsilon", ctypes.c_double), ("lam", ctypes.c_double),
                ("exact", ctypes.c_int), ("a", ctypes.c_double), ("b", ctypes.c_double), ("c", ctypes.c_double)]


def _proto(name, *args):
    f = getattr(_lib, name)
    f.argtypes = list(args)
    f.restype = ctypes.c_int
    return f


_proto("dohpTensorRule", ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _dp, _dp)
_proto("dohpTensorBasis", ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp)
_proto("dEFSApplyHexBatched", ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp,
       ctypes.c_int, ctypes.c_int, _vp)
_proto("dEFSApplyHexBatchedNoAlloc", ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp,
       ctypes.c_int, ctypes.c_int, _vp)
_proto("dRuleComputePhysicalBatched", ctypes.c_int, _dp, ctypes.c_int, _vp, _vp, _vp, _vp)
_proto("dFSCreateBlock", ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.POINTER(_vp))
_proto("dFSCreateFromElements", ctypes.c_int, ctypes.c_int, _long, _long, _ip, _long, _dp, _ip,
       ctypes.POINTER(_vp))
_proto("dFSDestroy", ctypes.POINTER(_vp))
_proto("dFSGetSizes", _vp, ctypes.POINTER(_long), ctypes.POINTER(_long), ctypes.POINTER(_long), _ip)
_proto("dFSGetNodalCoordinatesClosure", _vp, _dp)
_proto("dFSInhomogeneousDirichletCommit", _vp, _dp)
_proto("dFSGlobalToExpanded", _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp)
_proto("dFSExpandedToGlobal", _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp)
_proto("dFSGetElementNodes", _vp, _ip)
_proto("dohpEllipCreate", _vp, ctypes.POINTER(EllipParam), ctypes.POINTER(_vp))
_proto("dohpEllipDestroy", ctypes.POINTER(_vp))
_proto("dohpEllipSetCoefficient", _vp, _dp)
_proto("dohpEllipFunction", _vp, _vp, _vp, _vp)
_proto("dohpEllipShellMatMult", _vp, _vp, _vp, _vp)
_proto("dohpEllipShellMatMultF32", _vp, _vp, _vp, _vp)
_proto("dohpEllipShellMatMultAdd", _vp, _vp, _vp, _vp)
_proto("dohpEllipSetScatter", _vp, ctypes.c_int)
_proto("dohpEllipSetChecks", _vp, ctypes.c_int)
_proto("dohpEllipGetTuning", _vp, ctypes.c_int, ctypes.c_int, _vp, _vp)
_proto("dohpEllipSetShape", _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int)
_proto("dohpEllipAssembleQ1", _vp, _vp, ctypes.POINTER(_vp), _vp)
_proto("dohpEllipSetQ1Mode", _vp, ctypes.c_int)
_proto("dohpEllipSetStorage", _vp, ctypes.c_int)
_proto("dohpEllipGetStorage", _vp, ctypes.c_int, _vp, _vp)
_proto("dohpMatGetSizes", _vp, ctypes.POINTER(_long), ctypes.POINTER(_long))
_proto("dohpMatGetCSR", _vp, _ip, _ip, _dp)
_proto("dohpMatMult", _vp, _vp, _vp, _vp)
_proto("dohpMatMultF32", _vp, _vp, _vp, _vp)
_proto("dohpMatGetDiagonal", _vp, _vp, _vp)
_proto("dohpMatDestroy", ctypes.POINTER(_vp))


class MG:
    """Multigrid V-cycle on the embedded-Q1 matrix (dohpMG)."""

    def __init__(self, q1, smooth_degree=2, max_coarse=0, stream=None):
        self.h = ctypes.c_void_p()
        check(_lib.dohpMGCreate(q1.h, int(smooth_degree), int(max_coarse), ctypes.byref(self.h), _stream(stream)))
        nl, nc, lm = ctypes.c_int(), ctypes.c_long(), ctypes.c_double()
        check(_lib.dohpMGGetInfo(self.h, ctypes.byref(nl), ctypes.byref(nc), ctypes.byref(lm)))
        self.nlevels, self.ncoarse, self.lmax = nl.value, nc.value, lm.value

    def level(self, l):
        box = (ctypes.c_int * 3)()
        check(_lib.dohpMGGetLevel(self.h, int(l), box, None))
        n = box[0] * box[1] * box[2]
        st = np.zeros(27 * n)
        check(_lib.dohpMGGetLevel(self.h, int(l), box, st.ctypes.data_as(ctypes.c_void_p)))
        return tuple(box), st.reshape(27, n)

    def refresh(self, q1, stream=None):
        """Recompute every level from q1 refilled on the same box (no allocation)."""
        check(_lib.dohpMGRefresh(self.h, q1.h, _stream(stream)))
        nl, nc, lm = ctypes.c_int(), ctypes.c_long(), ctypes.c_double()
        check(_lib.dohpMGGetInfo(self.h, ctypes.byref(nl), ctypes.byref(nc), ctypes.byref(lm)))
        self.lmax = lm.value

    def set_graphs(self, on):
        """Replay V-cycles applied to the same vectors as HIP graphs (True, the default) or launch eagerly."""
        check(_lib.dohpMGSetGraphs(self.h, int(bool(on))))

    def apply(self, r, z, stream=None):
        check(_lib.dohpMGApply(self.h, _ptr(r), _ptr(z), _stream(stream)))

    def apply_f32(self, r, z, stream=None):
        check(_lib.dohpMGApplyF32(self.h, _ptr(r), _ptr(z), _stream(stream)))

    def residual(self, level, b, x, r, stream=None, fp16=False):
        """r = b - A_level x (b None: r = -A x); FP32 tensors use the level's FP32 coefficients, or with fp16 its FP16
        ones (the FP32 V-cycle's default)."""
        f = _lib.dohpMGResidualF32 if x.element_size() == 4 else _lib.dohpMGResidual
        if fp16:
            f = _lib.dohpMGResidualF16
        check(f(self.h, int(level), None if b is None else _ptr(b), _ptr(x), _ptr(r), _stream(stream)))

    def set_coefficients(self, fp16):
        """Coefficients of the FP32 V-cycle: FP16 (True, the default) or FP32."""
        check(_lib.dohpMGSetCoefficients(self.h, int(bool(fp16))))

    def __del__(self):
        if _lib is not None and getattr(self, "h", None):
            _lib.dohpMGDestroy(ctypes.byref(self.h))


class SolveOpts(ctypes.Structure):
    _fields_ = [("mixed", ctypes.c_int), ("rtol", ctypes.c_double), ("max_it", ctypes.c_int),
                ("inner_rtol", ctypes.c_double), ("max_outer", ctypes.c_int), ("cheb_degree", ctypes.c_int),
                ("cheb_ratio", ctypes.c_double), ("eig_its", ctypes.c_int), ("pc", ctypes.c_int),
                ("mg_degree", ctypes.c_int), ("fallback_its", ctypes.c_int), ("mg_coeffs", ctypes.c_int)]


class SolveStats(ctypes.Structure):
    _fields_ = [("outer_its", ctypes.c_int), ("inner_its", ctypes.c_int), ("bnorm", ctypes.c_double),
                ("rnorm", ctypes.c_double), ("lmax", ctypes.c_double), ("setup_seconds", ctypes.c_double),
                ("seconds", ctypes.c_double), ("fallback", ctypes.c_int)]


_proto("dohpMGCreate", _vp, ctypes.c_int, ctypes.c_int, _vp, _vp)
_proto("dohpMGGetInfo", _vp, _vp, _vp, _vp)
_proto("dohpMGGetLevel", _vp, ctypes.c_int, _vp, _vp)
_proto("dohpMGApply", _vp, _vp, _vp, _vp)
_proto("dohpMGRefresh", _vp, _vp, _vp)
_proto("dohpMGSetGraphs", _vp, ctypes.c_int)
_proto("dohpMGApplyF32", _vp, _vp, _vp, _vp)
_proto("dohpMGResidual", _vp, ctypes.c_int, _vp, _vp, _vp, _vp)
_proto("dohpMGResidualF32", _vp, ctypes.c_int, _vp, _vp, _vp, _vp)
_proto("dohpMGResidualF16", _vp, ctypes.c_int, _vp, _vp, _vp, _vp)
_proto("dohpMGSetCoefficients", _vp, ctypes.c_int)
_proto("dohpMGDestroy", _vp)
_proto("dohpEllipSolve", _vp, _vp, _vp, _vp, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveStats), _vp)
_proto("dohpEllipSolveLearned", _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveStats), _vp)
_proto("dohpEllipSolveRecord", _vp, _vp, _vp, _vp, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveStats), _vp,
       ctypes.c_int, ctypes.c_int, _vp, _vp)
_lib.dohpMLPNumParams.argtypes = []
_lib.dohpMLPNumParams.restype = ctypes.c_int
_lib.dohpMLPNumInputs.argtypes = []
_lib.dohpMLPNumInputs.restype = ctypes.c_int
_proto("dohpMLPCreate", _vp, _vp)
_proto("dohpMLPGetParams", _vp, _vp, _vp)
_proto("dohpMLPDestroy", _vp)
_proto("dohpMLPNumPatches", _vp, _vp)
_proto("dohpMLPFeatures", _vp, _vp, _vp, _vp, _vp)
_proto("dohpMLPWeights", _vp, _vp, _vp, _vp, _vp, _vp)
_proto("dohpMLPGradient", _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp)
_proto("dohpMLPEnergyWeightGradient", _vp, _vp, ctypes.c_long, _vp, _vp, _vp, ctypes.c_double, _vp, _vp, _vp, _vp)
_proto("dohpMLPWeightStep", _vp, _vp, _vp, _vp, ctypes.c_float, _vp)
_proto("dohpMLPEnergyStep", _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_double, ctypes.c_float, _vp, _vp)
for _f in ("dohpMLPNumParamsKind", "dohpMLPNumInputsKind", "dohpMLPNumOutputsKind"):
    getattr(_lib, _f).argtypes = [ctypes.c_int]
    getattr(_lib, _f).restype = ctypes.c_int
_proto("dohpMLPCreateKind", ctypes.c_int, _vp, _vp)
_proto("dohpMLPCellScales", _vp, ctypes.c_long, _vp, _vp, _vp)
_proto("dohpMLPCellStep", _vp, ctypes.c_long, _vp, _vp, ctypes.c_float, _vp)
_proto("dohpEllipSetQ1CellScales", _vp, _vp)
_proto("dohpQ1CellCount", _vp, _vp)
_proto("dohpQ1CellFeatures", _vp, _vp, _vp, _vp)
_proto("dohpQ1CellScaleGradient", _vp, _vp, _vp, _vp, ctypes.c_double, _vp, _vp)
_proto("dohpEllipSolveFallback", _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveStats), _vp)


class MLP:
    """Learned node weights of the Q1 multigrid preconditioner, z = S M(S r) (dohpMLP)."""

    nparam = _lib.dohpMLPNumParams()
    ninputs = _lib.dohpMLPNumInputs()

    def __init__(self, params=None):
        self.h = ctypes.c_void_p()
        p = None
        if params is not None:
            self._p = np.ascontiguousarray(params, dtype=np.float32)
            assert self._p.size == self.nparam
            p = self._p.ctypes.data_as(ctypes.c_void_p)
        check(_lib.dohpMLPCreate(p, ctypes.byref(self.h)))

    def params(self):
        m, w = np.zeros(self.nparam, np.float32), np.zeros(self.nparam, np.float32)
        check(_lib.dohpMLPGetParams(self.h, m.ctypes.data_as(ctypes.c_void_p), w.ctypes.data_as(ctypes.c_void_p)))
        return m, w

    @staticmethod
    def num_patches(q1):
        n = ctypes.c_long()
        check(_lib.dohpMLPNumPatches(q1.h, ctypes.byref(n)))
        return n.value

    @staticmethod
    def features(q1, fs, dinv, X, stream=None):
        check(_lib.dohpMLPFeatures(q1.h, fs.h, _ptr(dinv), _ptr(X), _stream(stream)))

    def weights(self, q1, X, s, s32=None, stream=None):
        check(_lib.dohpMLPWeights(self.h, q1.h, _ptr(X), _ptr(s), _ptr(s32) if s32 is not None else None,
                                  _stream(stream)))

    def gradient(self, X, T, B=None, stream=None):
        """MSE gradient of the raw outputs (checks the training kernels)."""
        B = int(B or X.shape[0])
        g = np.zeros(self.nparam, np.float32)
        loss = ctypes.c_double()
        check(_lib.dohpMLPGradient(self.h, B, _ptr(X), _ptr(T), g.ctypes.data_as(ctypes.c_void_p), ctypes.byref(loss),
                                   _stream(stream)))
        return loss.value, g

    @staticmethod
    def energy_weight_gradient(op, mg, s, r, e, alpha=0.0, gs=None, stream=None):
        """(E, (z.r, z.Jz, e.Je)) for z = S M(S r); gs receives dE/ds if given."""
        E = ctypes.c_double()
        terms = (ctypes.c_double * 3)()
        check(_lib.dohpMLPEnergyWeightGradient(op.h, mg.h, int(s.shape[0]), _ptr(s), _ptr(r), _ptr(e),
                                               ctypes.c_double(alpha), _ptr(gs) if gs is not None else None,
                                               ctypes.byref(E), terms, _stream(stream)))
        return E.value, tuple(terms)

    def weight_step(self, q1, X, gs, lr, stream=None):
        """Adam step from dL/ds at the node weights."""
        check(_lib.dohpMLPWeightStep(self.h, q1.h, _ptr(X), _ptr(gs), ctypes.c_float(lr), _stream(stream)))

    def energy_step(self, op, q1, mg, X, r, e, lr, alpha=0.0, stream=None):
        E = ctypes.c_double()
        check(_lib.dohpMLPEnergyStep(self.h, op.h, q1.h, mg.h, _ptr(X), _ptr(r), _ptr(e), ctypes.c_double(alpha),
                                     ctypes.c_float(lr), ctypes.byref(E), _stream(stream)))
        return E.value

    def __del__(self):
        if _lib is not None and getattr(self, "h", None):
            _lib.dohpMLPDestroy(ctypes.byref(self.h))


class CellMLP:
    """Learned sub-cell derivative scales of the embedded-Q1 matrix (dohpMLP of kind 1): the features of each GLL
    sub-cell map to sigma = exp(clamp(o)) per reference direction."""

    kind = 1
    nparam = _lib.dohpMLPNumParamsKind(1)
    ninputs = _lib.dohpMLPNumInputsKind(1)
    noutputs = _lib.dohpMLPNumOutputsKind(1)

    def __init__(self, params=None):
        self.h = ctypes.c_void_p()
        p = None
        if params is not None:
            self._p = np.ascontiguousarray(params, dtype=np.float32)
            assert self._p.size == self.nparam
            p = self._p.ctypes.data_as(ctypes.c_void_p)
        check(_lib.dohpMLPCreateKind(1, p, ctypes.byref(self.h)))

    def params(self):
        m, w = np.zeros(self.nparam, np.float32), np.zeros(self.nparam, np.float32)
        check(_lib.dohpMLPGetParams(self.h, m.ctypes.data_as(ctypes.c_void_p), w.ctypes.data_as(ctypes.c_void_p)))
        return m, w

    def gradient(self, X, T, B=None, stream=None):
        """MSE gradient of the raw outputs (checks the training kernels)."""
        B = int(B or X.shape[0])
        g = np.zeros(self.nparam, np.float32)
        loss = ctypes.c_double()
        check(_lib.dohpMLPGradient(self.h, B, _ptr(X), _ptr(T), g.ctypes.data_as(ctypes.c_void_p), ctypes.byref(loss),
                                   _stream(stream)))
        return loss.value, g

    def scales(self, X, sig, stream=None):
        check(_lib.dohpMLPCellScales(self.h, int(X.shape[0]), _ptr(X), _ptr(sig), _stream(stream)))

    def step(self, X, gsig, lr, stream=None):
        check(_lib.dohpMLPCellStep(self.h, int(X.shape[0]), _ptr(X), _ptr(gsig), ctypes.c_float(lr), _stream(stream)))

    def __del__(self):
        if _lib is not None and getattr(self, "h", None):
            _lib.dohpMLPDestroy(ctypes.byref(self.h))


class DohpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dohp_amd error {code}: {msg}")
        self.code = code


def check(rc):
    if rc:
        buf = ctypes.create_string_buffer(512)
        _lib.dohpGetLastError(buf, 512)
        raise DohpError(rc, buf.value.decode())
    return rc


def lib():
    return _lib


def _ptr(a):
    """Device or host pointer of a torch tensor (or None)."""
    if a is None:
        return None
    return ctypes.c_void_p(a.data_ptr())


def _np(a, dtype=np.float64):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a, a.ctypes.data_as(_dp if dtype == np.float64 else _ip)


def _stream(stream):
    if stream is None:
        return None
    return ctypes.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))


# ----------------------------------------------------------------- dJacobi
def tensor_rule(method, family, order):
    size = ctypes.c_int()
    z, w = np.zeros(order + 4), np.zeros(order + 4)
    check(_lib.dohpTensorRule(method, family, order, ctypes.byref(size), z.ctypes.data_as(_dp), w.ctypes.data_as(_dp)))
    return z[:size.value].copy(), w[:size.value].copy()


def tensor_basis(P, rcoord):
    rc, rcp = _np(rcoord)
    Q = rc.size
    B, D, node, weight = np.zeros(Q * P), np.zeros(Q * P), np.zeros(P), np.zeros(P)
    check(_lib.dohpTensorBasis(P, Q, rcp, B.ctypes.data_as(_dp), D.ctypes.data_as(_dp), node.ctypes.data_as(_dp),
                               weight.ctypes.data_as(_dp)))
    return B.reshape(Q, P), D.reshape(Q, P), node, weight


def efs_apply_hex_batched(P, rcoord, nelem, jinv, D, x, y, amode, imode=INSERT_VALUES, stream=None):
    """dEFSApply over a batch of elements; jinv/x/y are device tensors (float64)."""
    rc, rcp = _np(rcoord)
    check(_lib.dEFSApplyHexBatched(P, rc.size, rcp, nelem, _ptr(jinv), D, _ptr(x), _ptr(y), amode, imode,
                                   _stream(stream)))


def efs_apply_hex_batched_noalloc(P, rcoord, nelem, jinv, D, x, y, amode, imode=INSERT_VALUES, stream=None):
    """dEFSApplyHexBatched without device allocation (tables staged in the library's per-GPU array); returns
    without waiting for the stream."""
    rc, rcp = _np(rcoord)
    check(_lib.dEFSApplyHexBatchedNoAlloc(P, rc.size, rcp, nelem, _ptr(jinv), D, _ptr(x), _ptr(y), amode, imode,
                                          _stream(stream)))


def rule_compute_physical_batched(weight1d, nelem, jac, jinv, jw, stream=None):
    """dRuleComputePhysical over nelem elements of a tensor rule with 1-D weights weight1d: jac, jinv
    [nelem][Q^3][9] and jw [nelem][Q^3] are device tensors (float64).  Raises DohpError with code
    ERR_NEGATIVE_JACOBIAN or ERR_NAN when the device finds a negative or non-finite determinant."""
    w, wp = _np(weight1d)
    check(_lib.dRuleComputePhysicalBatched(w.size, wp, int(nelem), _ptr(jac), _ptr(jinv), _ptr(jw), _stream(stream)))


# ---------------------------------------------------------------------- dFS
class FS:
    """Continuous degree-p space on a structured hexahedral block (dFS)."""

    def __init__(self, M, degree, vcoords, dirichlet_mask=7):
        self.M = tuple(int(m) for m in M)
        self.degree = int(degree)
        v, vp = _np(vcoords)
        h = _vp()
        check(_lib.dFSCreateBlock(self.M[0], self.M[1], self.M[2], self.degree, int(dirichlet_mask), vp,
                                  ctypes.byref(h)))
        self.h = h
        a, b, c, d = _long(), _long(), _long(), ctypes.c_int()
        check(_lib.dFSGetSizes(h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d)))
        self.nowned, self.nclosure, self.nexpanded, self.nelem = a.value, b.value, c.value, d.value
        self.nvert = v.shape[0] if v.ndim == 2 else v.size // 3

    @classmethod
    def from_elements(cls, degree, elem_nodes, nowned, nclosure, vcoords, elem_verts):
        """Space on any conforming hexahedral mesh from its maps (dFSCreateFromElements): elem_nodes
        [nelem][n^3] closure indices (owned first), elem_verts [nelem][8] vertex indices in tensor order,
        vcoords [nvert][3]."""
        self = cls.__new__(cls)
        self.M = None
        self.degree = int(degree)
        E = np.ascontiguousarray(elem_nodes, dtype=np.int32)
        Vt = np.ascontiguousarray(elem_verts, dtype=np.int32)
        v, vp = _np(vcoords)
        nvert = v.shape[0] if v.ndim == 2 else v.size // 3
        h = _vp()
        check(_lib.dFSCreateFromElements(self.degree, int(E.shape[0]), int(nowned), int(nclosure),
                                         E.ctypes.data_as(_ip), int(nvert), vp, Vt.ctypes.data_as(_ip),
                                         ctypes.byref(h)))
        self.h = h
        a, b, c, d = _long(), _long(), _long(), ctypes.c_int()
        check(_lib.dFSGetSizes(h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d)))
        self.nowned, self.nclosure, self.nexpanded, self.nelem = a.value, b.value, c.value, d.value
        self.nvert = nvert
        return self

    def nodal_coordinates(self):
        X = np.zeros((self.nclosure, 3))
        check(_lib.dFSGetNodalCoordinatesClosure(self.h, X.ctypes.data_as(_dp)))
        return X

    def element_nodes(self):
        n = self.degree + 1
        E = np.zeros((self.nelem, n ** 3), dtype=np.int32)
        check(_lib.dFSGetElementNodes(self.h, E.ctypes.data_as(_ip)))
        return E

    def dirichlet_commit(self, dvals):
        d, dp = _np(dvals)
        assert d.size == self.nclosure - self.nowned
        check(_lib.dFSInhomogeneousDirichletCommit(self.h, dp))

    def global_to_expanded(self, g, x, hmode=INHOMOGENEOUS, imode=INSERT_VALUES, stream=None):
        check(_lib.dFSGlobalToExpanded(self.h, _ptr(g), _ptr(x), hmode, imode, _stream(stream)))

    def expanded_to_global(self, x, g, hmode=INHOMOGENEOUS, imode=INSERT_VALUES, stream=None):
        check(_lib.dFSExpandedToGlobal(self.h, _ptr(x), _ptr(g), hmode, imode, _stream(stream)))

    def __del__(self):
        if _lib is not None and getattr(self, "h", None):
            _lib.dFSDestroy(ctypes.byref(self.h))
            self.h = None


# --------------------------------------------------------------- CSR matrix
class Mat:
    """Device CSR matrix (the assembled embedded-Q1 preconditioning matrix)."""

    def __init__(self):
        self.h = _vp()
        self.nrows = self.nnz = 0

    def _sizes(self):
        a, b = _long(), _long()
        check(_lib.dohpMatGetSizes(self.h, ctypes.byref(a), ctypes.byref(b)))
        self.nrows, self.nnz = a.value, b.value

    def csr(self):
        rp = np.zeros(self.nrows + 1, dtype=np.int32)
        col = np.zeros(self.nnz, dtype=np.int32)
        val = np.zeros(self.nnz)
        check(_lib.dohpMatGetCSR(self.h, rp.ctypes.data_as(_ip), col.ctypes.data_as(_ip), val.ctypes.data_as(_dp)))
        return rp, col, val

    def to_scipy(self):
        import scipy.sparse as sp
        rp, col, val = self.csr()
        return sp.csr_matrix((val, col, rp), shape=(self.nrows, self.nrows))

    def mult(self, x, y, stream=None):
        check(_lib.dohpMatMult(self.h, _ptr(x), _ptr(y), _stream(stream)))

    def mult_f32(self, x, y, stream=None):
        check(_lib.dohpMatMultF32(self.h, _ptr(x), _ptr(y), _stream(stream)))

    def diagonal(self, d, stream=None):
        check(_lib.dohpMatGetDiagonal(self.h, _ptr(d), _stream(stream)))

    def __del__(self):
        if _lib is not None and getattr(self, "h", None):
            _lib.dohpMatDestroy(ctypes.byref(self.h))
            self.h = None


# ------------------------------------------------------------------- ellip
class Ellip:
    """ellip's residual and MATSHELL Jacobian action on the GPU (src/fs/tests/ellip.c:407-472)."""

    def __init__(self, fs: FS, p=2.0, eps=1.0, lam=0.0, exact=-1, a=1.0, b=1.0, c=1.0, kvert=None):
        self.fs = fs
        self.prm = EllipParam(float(p), float(eps), float(lam), int(exact), float(a), float(b), float(c))
        h = _vp()
        check(_lib.dohpEllipCreate(fs.h, ctypes.byref(self.prm), ctypes.byref(h)))
        self.h = h
        if kvert is not None:
            self.set_coefficient(kvert)

    def set_coefficient(self, kvert):
        if kvert is None:
            check(_lib.dohpEllipSetCoefficient(self.h, None))
        else:
            k, kp = _np(kvert)
            check(_lib.dohpEllipSetCoefficient(self.h, kp))

    _TUNE_MODES = {"linear": 0, "residual": 1, "jacobian": 2, "stored": 3}

    def tuning(self, mode, fp32=False):
        """(chosen shape, ms of [four-wavefront, one-wavefront, persistent one-wavefront,
        persistent four-wavefront, four-wavefront on the matrix cores (degrees 3..7)]) of a mode: "linear"
        and "jacobian" recompute the geometry (storage 0), "stored" applies either operator from its
        quadrature data (storage 1)."""
        sh, ms = ctypes.c_int(), (ctypes.c_float * 5)()
        check(_lib.dohpEllipGetTuning(self.h, self._TUNE_MODES[mode], int(fp32), ctypes.byref(sh), ms))
        return sh.value, [round(v, 4) for v in ms]

    def set_shape(self, mode, fp32, shape):
        """Fix the launch shape of a mode ("linear", "residual", "jacobian", "stored") and precision
        (-1 = tune)."""
        check(_lib.dohpEllipSetShape(self.h, self._TUNE_MODES[mode], int(fp32), int(shape)))

    def set_storage(self, mode):
        """1: applications read per-point quadrature data formed once per linearization (48 bytes
        per point in FP64, 56 with lambda; half again for FP32 applications); 0: recompute from the
        state and the geometry at every application, no storage; 2 (default): the faster of the two,
        timed on the first application in each precision."""
        check(_lib.dohpEllipSetStorage(self.h, int(mode)))

    def storage(self, fp32=False):
        """(form in use: 1 stored, 0 recomputed, -1 not timed yet; [ms recomputed, ms stored] of the timing
        that chose it, zeros unless storage 2 timed it)."""
        st, ms = ctypes.c_int(), (ctypes.c_float * 2)()
        check(_lib.dohpEllipGetStorage(self.h, int(fp32), ctypes.byref(st), ms))
        return st.value, [round(v, 4) for v in ms]

    def set_scatter(self, mode):
        check(_lib.dohpEllipSetScatter(self.h, int(mode)))

    def set_checks(self, mode):
        """1 (default): function and mat_mult* wait for their stream and raise DohpError(ERR_NAN) when a
        non-finite value met the element kernels; 0: asynchronous, unchecked."""
        check(_lib.dohpEllipSetChecks(self.h, int(mode)))

    def set_q1_cell_scales(self, sig):
        """Learned sub-cell derivative scales [ncell][3] (device float32 tensor, kept alive here) for later Q1
        assemblies; None = none."""
        self._cscale = sig
        check(_lib.dohpEllipSetQ1CellScales(self.h, _ptr(sig)))

    def q1_cell_count(self):
        n = ctypes.c_long()
        check(_lib.dohpQ1CellCount(self.h, ctypes.byref(n)))
        return n.value

    def q1_cell_features(self, q1, X, stream=None):
        check(_lib.dohpQ1CellFeatures(self.h, q1.h, _ptr(X), _stream(stream)))

    def q1_cell_scale_gradient(self, q1, w, z, coef, g, stream=None):
        """g[cell][a] += coef * w^T dK/dsigma_{cell,a} z (the scales set on this operator)."""
        check(_lib.dohpQ1CellScaleGradient(self.h, q1.h, _ptr(w), _ptr(z), ctypes.c_double(coef), _ptr(g),
                                           _stream(stream)))

    def set_q1_mode(self, mode):
        """Later embedded-Q1 assemblies: 0 = the reference's arithmetic, 1 = corrected (GLL sub-mesh Q1 FEM)."""
        check(_lib.dohpEllipSetQ1Mode(self.h, int(mode)))

    def function(self, u, f, stream=None):
        check(_lib.dohpEllipFunction(self.h, _ptr(u), _ptr(f), _stream(stream)))

    def mat_mult(self, x, y, stream=None):
        check(_lib.dohpEllipShellMatMult(self.h, _ptr(x), _ptr(y), _stream(stream)))

    def mat_mult_f32(self, x, y, stream=None):
        check(_lib.dohpEllipShellMatMultF32(self.h, _ptr(x), _ptr(y), _stream(stream)))

    def mat_mult_add(self, x, y, stream=None):
        """y += J x: the element kernel alone (atomic scatter)."""
        check(_lib.dohpEllipShellMatMultAdd(self.h, _ptr(x), _ptr(y), _stream(stream)))

    def assemble_q1(self, u=None, mat=None, stream=None):
        """Embedded-Q1 matrix at state u (EllipJacobian); reuses `mat` (same pattern) if given."""
        m = mat if mat is not None else Mat()
        check(_lib.dohpEllipAssembleQ1(self.h, _ptr(u), ctypes.byref(m.h), _stream(stream)))
        m._sizes()
        return m

    def solve(self, q1, b, x, mixed=False, rtol=1e-10, max_it=2000, inner_rtol=1e-4, max_outer=50, cheb_degree=4,
              cheb_ratio=30.0, eig_its=20, pc="mg", mg_degree=0, mlp=None, fallback_its=0, record=None,
              fallback_q1=None, mg_coeffs="fp16", stream=None):
        """CG on the linear operator, preconditioned through the embedded-Q1 matrix: pc="mg" is a
        multigrid V-cycle with Jacobi-Chebyshev smoothing, pc="cheb" Jacobi-Chebyshev alone.
        mixed: 0 plain FP64, 1 FP32 inner PCG inside FP64 defect correction, 2 FP64 PCG with the
        preconditioner in FP32.  mg_degree: smoothing degree, 0 = 1 on a corrected Q1 matrix or one with learned
        sub-cell scales, 2 otherwise.  mlp: learned node weights around the V-cycle (plain fallback after
        fallback_its iterations).  fallback_q1: q1 is a learned embedded-Q1 matrix and the multigrid of fallback_q1
        takes over after fallback_its iterations.  mg_coeffs: "fp16" or "fp32" coefficients of the FP32 V-cycle.
        record=(snap, every): plain FP64 solve keeping iterates."""
        o = SolveOpts(int(mixed), float(rtol), int(max_it), float(inner_rtol), int(max_outer), int(cheb_degree),
                      float(cheb_ratio), int(eig_its), {"cheb": 0, "mg": 1}[pc], int(mg_degree), int(fallback_its),
                      {"fp16": 0, "fp32": 1}[mg_coeffs])
        st = SolveStats()
        out = {}
        if fallback_q1 is not None:
            check(_lib.dohpEllipSolveFallback(self.h, q1.h, fallback_q1.h, _ptr(b), _ptr(x), ctypes.byref(o),
                                              ctypes.byref(st), _stream(stream)))
        elif mlp is not None:
            check(_lib.dohpEllipSolveLearned(self.h, q1.h, mlp.h, _ptr(b), _ptr(x), ctypes.byref(o), ctypes.byref(st),
                                             _stream(stream)))
        elif record is not None:
            snap, every = record
            ns = ctypes.c_int()
            check(_lib.dohpEllipSolveRecord(self.h, q1.h, _ptr(b), _ptr(x), ctypes.byref(o), ctypes.byref(st),
                                            _ptr(snap), int(snap.shape[0]), int(every), ctypes.byref(ns),
                                            _stream(stream)))
            out["nsnap"] = ns.value
        else:
            check(_lib.dohpEllipSolve(self.h, q1.h, _ptr(b), _ptr(x), ctypes.byref(o), ctypes.byref(st),
                                      _stream(stream)))
        out.update({f: getattr(st, f) for f, _ in SolveStats._fields_})
        return out

    def __del__(self):
        if _lib is not None and getattr(self, "h", None):
            _lib.dohpEllipDestroy(ctypes.byref(self.h))
            self.h = None


# ------------------------------------------------ batched mixed-order Poisson operator
_proto("dohpBatchCreate", ctypes.POINTER(_vp))
_proto("dohpBatchDestroy", ctypes.POINTER(_vp))
_proto("dohpBatchAddSpace", _vp, _vp, _dp, _ip)
_proto("dohpBatchAddElements", _vp, ctypes.c_int, _ip, _dp, _dp, _ip)
_proto("dohpBatchSetUp", _vp)
_proto("dohpBatchGetSizes", _vp, ctypes.POINTER(_long), _ip)
_proto("dohpBatchGetSegment", _vp, ctypes.c_int, ctypes.POINTER(_long), ctypes.POINTER(_long))
_proto("dohpBatchMult", _vp, _vp, _vp, _vp)
_proto("dohpBatchMultF32", _vp, _vp, _vp, _vp)
_proto("dohpBatchGetTuning", _vp, ctypes.c_int, ctypes.c_int, _ip, _vp)
_proto("dohpBatchSetMFMA", _vp, ctypes.c_int)
_proto("dohpBatchSetTile", _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int)


class Batch:
    """Linear Poisson operators of block spaces and of independent mixed-order elements, applied
    together in one persistent launch (dohpBatch).  Vectors are packed: block spaces first, element
    sets after them; segment(i) gives (offset, length) of segment i."""

    def __init__(self):
        self.h = _vp()
        check(_lib.dohpBatchCreate(ctypes.byref(self.h)))

    def add_space(self, fs, kvert=None):
        seg = ctypes.c_int()
        kp = None
        if kvert is not None:
            self._k = np.ascontiguousarray(kvert, dtype=np.float64)
            kp = self._k.ctypes.data_as(_dp)
        check(_lib.dohpBatchAddSpace(self.h, fs.h, kp, ctypes.byref(seg)))
        return seg.value

    def add_elements(self, degrees, elem_vx, kvert=None):
        """degrees [nelem] in 1..10, elem_vx [nelem][8][3] (tensor vertex order), kvert [nelem][8]."""
        d = np.ascontiguousarray(degrees, dtype=np.int32)
        v = np.ascontiguousarray(elem_vx, dtype=np.float64)
        assert v.size == 24 * d.size
        kp = None
        if kvert is not None:
            k = np.ascontiguousarray(kvert, dtype=np.float64)
            assert k.size == 8 * d.size
            kp = k.ctypes.data_as(_dp)
        seg = ctypes.c_int()
        check(_lib.dohpBatchAddElements(self.h, int(d.size), d.ctypes.data_as(_ip), v.ctypes.data_as(_dp), kp,
                                        ctypes.byref(seg)))
        return seg.value

    def setup(self):
        check(_lib.dohpBatchSetUp(self.h))

    @property
    def size(self):
        n, s = _long(), ctypes.c_int()
        check(_lib.dohpBatchGetSizes(self.h, ctypes.byref(n), ctypes.byref(s)))
        return n.value

    def segment(self, i):
        o, n = _long(), _long()
        check(_lib.dohpBatchGetSegment(self.h, int(i), ctypes.byref(o), ctypes.byref(n)))
        return o.value, n.value

    def mult(self, x, y, stream=None):
        check(_lib.dohpBatchMult(self.h, _ptr(x), _ptr(y), _stream(stream)))

    def mult_f32(self, x, y, stream=None):
        check(_lib.dohpBatchMultF32(self.h, _ptr(x), _ptr(y), _stream(stream)))

    def set_mfma(self, passes):
        """FP32 MFMA contraction tiles from degree 4 up: 0 off, 1 BF16, 3 split BF16 (three MFMAs)."""
        check(_lib.dohpBatchSetMFMA(self.h, int(passes)))

    def set_tile(self, degree, fp32, tile):
        """Fix the tile of a degree and precision (-1 = tune)."""
        check(_lib.dohpBatchSetTile(self.h, int(degree), int(fp32), int(tile)))

    def tuning(self, degree, fp32=False):
        """(tile: -1 none, 0 workgroup-packed, 1 wavefront-packed, 2 MFMA; [ms of each])."""
        t, ms = ctypes.c_int(), (ctypes.c_float * 3)()
        check(_lib.dohpBatchGetTuning(self.h, int(degree), int(fp32), ctypes.byref(t), ms))
        return t.value, [round(v, 4) for v in ms]

    def __del__(self):
        if _lib is not None and getattr(self, "h", None):
            _lib.dohpBatchDestroy(ctypes.byref(self.h))
            self.h = None

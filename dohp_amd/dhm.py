"""ctypes binding of the DHM field-file library (dohp_amd/libdohp_dhm.so, include/dohp_dhm.h): the on-disk format of
dohp's viewer (src/viewer/dhm.c, src/fs/impls/cont/contviewdhm.c) for block function spaces and closure vectors.
Host arrays only; the library is host code linked with HDF5."""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.environ.get("DOHP_DHM_LIB", os.path.join(_HERE, "libdohp_dhm.so"))
if not os.path.exists(LIBPATH):
    raise ImportError(f"dohp_amd: DHM library not built ({LIBPATH}); run __graft_entry__.build()")
_lib = ctypes.CDLL(LIBPATH)
READ, WRITE, APPEND = 0, 1, 2
_vp, _dp, _ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)


def _proto(name, *args):
    f = getattr(_lib, name)
    f.argtypes = list(args)
    f.restype = ctypes.c_int


_proto("dohpDHMOpen", ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(_vp))
_proto("dohpDHMClose", ctypes.POINTER(_vp))
_proto("dohpDHMSetTime", _vp, ctypes.c_double)
_proto("dohpDHMSetTimeUnits", _vp, ctypes.c_char_p, ctypes.c_double)
_proto("dohpDHMGetSteps", _vp, _ip, _dp, ctypes.c_int)
_proto("dohpDHMSetTimeStep", _vp, ctypes.c_int)
_proto("dohpDHMWriteBlockSpace", _vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
       ctypes.c_int, _dp, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p))
_proto("dohpDHMReadBlockSpace", _vp, ctypes.c_char_p, _ip, _ip, _ip, _ip, _dp)
_proto("dohpDHMWriteVec", _vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_long, ctypes.c_int, _dp, ctypes.c_double)
_proto("dohpDHMReadVec", _vp, ctypes.c_char_p, ctypes.c_long, ctypes.c_int, _dp)
_lib.dohpDHMGetLastError.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
_lib.dohpDHMGetLastError.restype = None


class DHMError(RuntimeError):
    pass


def _check(rc):
    if rc:
        buf = ctypes.create_string_buffer(512)
        _lib.dohpDHMGetLastError(buf, 512)
        raise DHMError(buf.value.decode())


class DHM:
    """A DHM file: spaces, time steps and the closure vectors written in them."""

    def __init__(self, path, mode=READ):
        self.h = _vp()
        _check(_lib.dohpDHMOpen(os.fsencode(path), int(mode), ctypes.byref(self.h)))

    def close(self):
        if getattr(self, "h", None):
            _lib.dohpDHMClose(ctypes.byref(self.h))

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_time(self, t):
        _check(_lib.dohpDHMSetTime(self.h, float(t)))

    def set_time_units(self, units, scale):
        _check(_lib.dohpDHMSetTimeUnits(self.h, units.encode(), float(scale)))

    def steps(self):
        n = ctypes.c_int()
        _check(_lib.dohpDHMGetSteps(self.h, ctypes.byref(n), None, 0))
        t = np.zeros(max(n.value, 1))
        _check(_lib.dohpDHMGetSteps(self.h, ctypes.byref(n), t.ctypes.data_as(_dp), n.value))
        return t[:n.value]

    def set_step(self, s):
        _check(_lib.dohpDHMSetTimeStep(self.h, int(s)))

    def write_block_space(self, name, M, degree, vcoords, dirichlet_mask=7, fields=None):
        v = np.ascontiguousarray(vcoords, dtype=np.float64)
        bs = len(fields) if fields else 1
        arr = (ctypes.c_char_p * bs)(*[f.encode() for f in fields]) if fields else None
        _check(_lib.dohpDHMWriteBlockSpace(self.h, name.encode(), int(M[0]), int(M[1]), int(M[2]), int(degree),
                                           int(dirichlet_mask), v.ctypes.data_as(_dp), bs, arr))

    def read_block_space(self, name, with_coords=True):
        M = (ctypes.c_int * 3)()
        deg, mask, bs = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(_lib.dohpDHMReadBlockSpace(self.h, name.encode(), M, ctypes.byref(deg), ctypes.byref(mask),
                                          ctypes.byref(bs), None))
        out = {"M": tuple(M), "degree": deg.value, "dirichlet_mask": mask.value, "bs": bs.value}
        if with_coords:
            V = np.zeros(((M[0] + 1) * (M[1] + 1) * (M[2] + 1), 3))
            _check(_lib.dohpDHMReadBlockSpace(self.h, name.encode(), M, ctypes.byref(deg), ctypes.byref(mask),
                                              ctypes.byref(bs), V.ctypes.data_as(_dp)))
            out["vcoords"] = V
        return out

    def write_vec(self, fsname, name, x, bs=1, unit_scale=1.0):
        x = np.ascontiguousarray(x, dtype=np.float64)
        _check(_lib.dohpDHMWriteVec(self.h, fsname.encode(), name.encode(), x.size, int(bs), x.ctypes.data_as(_dp),
                                    float(unit_scale)))

    def read_vec(self, name, n, bs=1):
        x = np.zeros(int(n))
        _check(_lib.dohpDHMReadVec(self.h, name.encode(), x.size, int(bs), x.ctypes.data_as(_dp)))
        return x

    def __del__(self):
        self.close()

"""ctypes binding of the optional HDF5 writer (``include/pp_h5.h``,
libpp_h5.so): the /DescrSpatStats groups and the NPCORR datasets of the
reference's results file, and a reader for them."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libpp_h5.so")
KIND_DTYPE = {1: np.uint32, 2: np.uint64, 3: np.float32}
_lib = None


def available() -> bool:
    return os.path.exists(LIB_PATH)


def load():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError(f"{LIB_PATH} is missing (built only where HDF5 is installed)")
        lib = C.CDLL(LIB_PATH)
        P = C.c_void_p
        lib.pph5_create_results.argtypes = [C.c_char_p, C.c_int32]
        lib.pph5_report_npc3d_bincenters.argtypes = [C.c_char_p, C.c_uint64, C.c_float]
        lib.pph5_report_npc3d_histvalues.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p,
                                                     C.c_int32, P, C.c_uint64]
        lib.pph5_read.argtypes = [C.c_char_p, C.c_char_p, P, C.c_uint64,
                                  C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                  C.POINTER(C.c_int32)]
        _lib = lib
    return _lib


def read(fn: str, dset: str):
    """A two-dimensional dataset as a numpy array, or None when absent."""
    lib = load()
    nr, nc, kind = C.c_uint64(), C.c_uint64(), C.c_int32()
    rc = lib.pph5_read(fn.encode(), dset.encode(), None, 0, C.byref(nr), C.byref(nc),
                       C.byref(kind))
    if rc not in (0, -2) or kind.value not in KIND_DTYPE:
        return None
    out = np.empty((nr.value, nc.value), KIND_DTYPE[kind.value])
    rc = lib.pph5_read(fn.encode(), dset.encode(), out.ctypes.data, out.nbytes, C.byref(nr),
                       C.byref(nc), C.byref(kind))
    return out if rc == 0 else None

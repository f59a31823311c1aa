"""ctypes wrapper of oracle/pp_cpu_baseline.cpp (the reference's CPU
algorithm: z-slabs, KD-trees, OpenMP region sweep).  Bench and test
infrastructure only."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libcpu_baseline.so")
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            from oracle.build_cpu_baseline import build

            build()
        lib = C.CDLL(LIB)
        P = C.c_void_p
        lib.pp_cpu_spatstat.argtypes = [
            C.c_int64, P, P, P, C.c_int32, P, P, C.c_float, C.c_float, C.c_float, C.c_uint32,
            C.c_int32, P, C.c_int32, P, P, P, P, P, C.c_int64, P, P,
        ]
        lib.pp_cpu_spatstat.restype = C.c_int
        lib.pp_cpu_ion2surf.argtypes = [C.c_int64, P, C.c_int64, P, C.c_float, C.c_int32,
                                        C.c_int64, P, P]
        lib.pp_cpu_ion2surf.restype = C.c_int
        lib.pp_cpu_npcorr.argtypes = [C.c_int64, P, P, P, C.c_int32, P, P, C.c_float, C.c_float,
                                      C.c_int64, C.c_int32, P, P]
        lib.pp_cpu_npcorr.restype = C.c_int
        _lib = lib
    return _lib


def _mask(types):
    m = np.zeros(32, dtype=np.uint8)
    for t in types:
        m[int(t) >> 3] |= np.uint8(1 << (int(t) & 7))
    return m


def spatstat(xyz, types, ion2surf, tasks, rmin, rincr, rmax, families, kvalues=(),
             nthreads=0, cntnb_cap=65536):
    lib = _load()
    xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
    types = np.ascontiguousarray(types, dtype=np.uint8)
    s2 = np.ascontiguousarray(ion2surf, dtype=np.float32)
    n, T = xyz.shape[0], len(tasks)
    trg = np.zeros((max(T, 1), 32), np.uint8)
    env = np.zeros((max(T, 1), 32), np.uint8)
    for i, (t, e) in enumerate(tasks):
        trg[i], env[i] = _mask(t), _mask(e)
    kd = np.array(sorted((int(k) - 1 for k in kvalues), reverse=True), dtype=np.int64)
    from oracle.oracle import hist_geometry

    nb, _, ok = hist_geometry(rmin, rincr, rmax)
    if not ok:
        raise ValueError("invalid histogram geometry")
    Cn = nb + 2
    out = {"rdf": np.zeros((T, Cn), np.int64), "nn1": np.zeros((T, Cn), np.int64),
           "ripk": np.zeros((T, Cn), np.int64), "mknn": np.zeros((T, len(kd), Cn), np.int64),
           "cntnb": np.zeros((T, cntnb_cap), np.int64)}
    stats = np.zeros(4, np.int64)
    tim = np.zeros(2, np.float64)
    st = lib.pp_cpu_spatstat(n, xyz.ctypes.data, types.ctypes.data, s2.ctypes.data, T,
                             trg.ctypes.data, env.ctypes.data, rmin, rincr, rmax, families,
                             len(kd), kd.ctypes.data, nthreads, out["rdf"].ctypes.data,
                             out["nn1"].ctypes.data, out["ripk"].ctypes.data,
                             out["mknn"].ctypes.data, out["cntnb"].ctypes.data, cntnb_cap,
                             stats.ctypes.data, tim.ctypes.data)
    if st != 0:
        raise ValueError("cpu baseline rejected the arguments")
    keep = {}
    for name, bit in (("rdf", 1), ("nn1", 2), ("ripk", 4), ("mknn", 8), ("cntnb", 32)):
        if families & bit:
            keep[name] = out[name]
    return keep, {"considered": int(stats[0]), "discarded": int(stats[1]),
                  "inball": int(stats[2]), "index_s": float(tim[0]), "compute_s": float(tim[1])}


def ion2surf(xyz, triangles, R, nthreads=0, stride=1):
    """Every stride-th ion's squared hull distance by a triangle-box tree
    query per ion (the reference's surface_distancing2 structure).  Returns
    (values, build_s, query_s)."""
    lib = _load()
    xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
    tri = np.ascontiguousarray(triangles, dtype=np.float32).reshape(-1, 9)
    m = (xyz.shape[0] + stride - 1) // stride
    out = np.empty(m, np.float32)
    tim = np.zeros(2, np.float64)
    st = lib.pp_cpu_ion2surf(xyz.shape[0], xyz.ctypes.data, tri.shape[0], tri.ctypes.data,
                             C.c_float(R), nthreads, stride, out.ctypes.data, tim.ctypes.data)
    if st != 0:
        raise ValueError("cpu baseline rejected the arguments")
    return out, float(tim[0]), float(tim[1])


def npcorr(xyz, types, ion2surf, tasks, rmax, rincr, knn_order, nthreads=0):
    """NPCORR counts [ntasks][edge][edge][edge] the reference's way (z-slabs,
    KD-trees, nth_element); returns (counts, index_s, compute_s)."""
    lib = _load()
    xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
    types = np.ascontiguousarray(types, dtype=np.uint8)
    s2 = np.ascontiguousarray(ion2surf, dtype=np.float32)
    T = len(tasks)
    trg = np.zeros((max(T, 1), 32), np.uint8)
    env = np.zeros((max(T, 1), 32), np.uint8)
    for i, (t, e) in enumerate(tasks):
        trg[i], env[i] = _mask(t), _mask(e)
    half = int(np.ceil(np.float32(rmax) / np.float32(rincr)) + np.float32(1.0))
    edge = 1 + 2 * half
    out = np.zeros((T, edge, edge, edge), np.uint32)
    tim = np.zeros(2, np.float64)
    st = lib.pp_cpu_npcorr(xyz.shape[0], xyz.ctypes.data, types.ctypes.data, s2.ctypes.data, T,
                           trg.ctypes.data, env.ctypes.data, rmax, rincr, knn_order, nthreads,
                           out.ctypes.data, tim.ctypes.data)
    if st != 0:
        raise ValueError("cpu baseline rejected the arguments")
    return out, float(tim[0]), float(tim[1])

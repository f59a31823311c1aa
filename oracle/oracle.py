"""ctypes wrapper of the CPU oracle (``oracle/pp_oracle.c``).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product path never imports
this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
SRC = os.path.join(HERE, "pp_oracle.c")

FAM_RDF, FAM_1NN, FAM_RIPK, FAM_MKNN, FAM_CNTNB = 1, 2, 4, 8, 32


def build() -> str:
    """gcc -O2 -fopenmp -ffp-contract=off (no FMA contraction, like the GPU)."""
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
           "-fno-fast-math", SRC, "-o", tmp, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        P = C.c_void_p
        lib.pp_oracle_spatstat.argtypes = [
            C.c_int64, P, P, P, C.c_int32, P, P, C.c_float, C.c_float, C.c_float, C.c_uint32,
            C.c_int32, P, C.c_int32, C.c_int32, P, P, P, P, P, C.c_int64, P,
        ]
        lib.pp_oracle_spatstat.restype = C.c_int
        lib.pp_oracle_category.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float]
        lib.pp_oracle_category.restype = C.c_int64
        lib.pp_oracle_categories_d2.argtypes = [C.c_float, C.c_float, C.c_float, P, C.c_int64, P]
        lib.pp_oracle_hist_geometry.argtypes = [C.c_float, C.c_float, C.c_float,
                                                C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        lib.pp_oracle_hist_geometry.restype = C.c_int
        lib.pp_oracle_binedge_report.argtypes = [
            C.c_int64, P, P, P, P, P, C.c_float, C.c_float, C.c_float, C.c_int64, C.c_int32, P,
        ]
        lib.pp_oracle_binedge_report.restype = C.c_int
        lib.pp_oracle_category_sweep.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float,
                                                 C.c_float, P, P, C.c_int64]
        lib.pp_oracle_category_sweep.restype = C.c_int64
        lib.pp_oracle_closest_tri.argtypes = [P, P]
        lib.pp_oracle_closest_tri.restype = C.c_float
        lib.pp_oracle_ion2surf.argtypes = [P, C.c_int64, P, C.c_int64, C.c_float, P]
        lib.pp_oracle_pos_decode.argtypes = [P, C.c_int64, P, P, P, C.c_int32, P, P, P]
        lib.pp_oracle_epos_decode.argtypes = [P, C.c_int64, P, P, P, C.c_int32, P, P, P]
        lib.pp_oracle_npcorr.argtypes = [C.c_int64, P, P, P, C.c_int32, P, P, C.c_float,
                                         C.c_float, C.c_int64, P]
        lib.pp_oracle_npcorr.restype = C.c_int
        _lib = lib
    return _lib


def hist_geometry(rmin, rincr, rmax):
    nb = C.c_int64()
    tmp = C.c_double()
    st = _load().pp_oracle_hist_geometry(rmin, rincr, rmax, C.byref(nb), C.byref(tmp))
    return nb.value, tmp.value, st == 0


def category(rmin, rincr, rmax, d) -> int:
    return int(_load().pp_oracle_category(rmin, rincr, rmax, d))


def categories_d2(rmin, rincr, rmax, d2: np.ndarray) -> np.ndarray:
    d2 = np.ascontiguousarray(d2, dtype=np.float32)
    out = np.empty(d2.shape[0], dtype=np.int32)
    _load().pp_oracle_categories_d2(rmin, rincr, rmax, d2.ctypes.data, d2.shape[0], out.ctypes.data)
    return out


def category_sweep(rmin, rincr, rmax, d_lo, d_hi, cap=1 << 20):
    """Every float d in [d_lo, d_hi] through the literal histogram::add:
    (first float of each new category, that category), in order; raises if
    the category ever decreases."""
    first = np.zeros(cap, np.uint32)
    cat = np.zeros(cap, np.int32)
    n = _load().pp_oracle_category_sweep(rmin, rincr, rmax, d_lo, d_hi, first.ctypes.data,
                                         cat.ctypes.data, cap)
    if n < 0 or n > cap:
        raise ValueError("category sweep: not monotone, or too many changes")
    return first[:n].view(np.float32), cat[:n]


def _mask(types):
    m = np.zeros(32, dtype=np.uint8)
    for t in types:
        m[int(t) >> 3] |= np.uint8(1 << (int(t) & 7))
    return m


def spatstat(xyz, types, ion2surf, tasks, rmin, rincr, rmax, families, kvalues=(),
             brute=False, nthreads=0, cntnb_cap=65536):
    """Returns (dict of int64 histograms, stats dict).  Same layout as the C ABI."""
    lib = _load()
    xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
    types = np.ascontiguousarray(types, dtype=np.uint8)
    s2 = np.ascontiguousarray(ion2surf, dtype=np.float32)
    n = xyz.shape[0]
    T = len(tasks)
    trg = np.zeros((max(T, 1), 32), np.uint8)
    env = np.zeros((max(T, 1), 32), np.uint8)
    for i, (t, e) in enumerate(tasks):
        trg[i] = _mask(t)
        env[i] = _mask(e)
    kd = np.array(sorted((int(k) - 1 for k in kvalues), reverse=True), dtype=np.int64)
    nb, _, ok = hist_geometry(rmin, rincr, rmax)
    if not ok:
        raise ValueError("invalid histogram geometry")
    Cn = nb + 2
    out = {
        "rdf": np.zeros((T, Cn), np.int64),
        "nn1": np.zeros((T, Cn), np.int64),
        "ripk": np.zeros((T, Cn), np.int64),
        "mknn": np.zeros((T, len(kd), Cn), np.int64),
        "cntnb": np.zeros((T, cntnb_cap), np.int64),
    }
    stats = np.zeros(4, np.int64)
    st = lib.pp_oracle_spatstat(
        n, xyz.ctypes.data, types.ctypes.data, s2.ctypes.data, T, trg.ctypes.data,
        env.ctypes.data, rmin, rincr, rmax, families, len(kd), kd.ctypes.data,
        1 if brute else 0, nthreads, out["rdf"].ctypes.data, out["nn1"].ctypes.data,
        out["ripk"].ctypes.data, out["mknn"].ctypes.data, out["cntnb"].ctypes.data, cntnb_cap,
        stats.ctypes.data)
    if st == 1:
        raise ValueError("oracle rejected the arguments")
    if st == 2:
        raise OverflowError("CNTNB count above cntnb_cap")
    keep = {}
    for name, bit in (("rdf", FAM_RDF), ("nn1", FAM_1NN), ("ripk", FAM_RIPK),
                      ("mknn", FAM_MKNN), ("cntnb", FAM_CNTNB)):
        if families & bit:
            keep[name] = out[name]
    return keep, {"considered": int(stats[0]), "discarded": int(stats[1]),
                  "inball": int(stats[2]), "maxcnt": int(stats[3])}


def binedge_report(xyz, types, ion2surf, trg, env, rmin, rincr, rmax, center_stride=1,
                   nthreads=0):
    lib = _load()
    xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
    types = np.ascontiguousarray(types, dtype=np.uint8)
    s2 = np.ascontiguousarray(ion2surf, dtype=np.float32)
    out = np.zeros(4, np.int64)
    tm, em = _mask(trg), _mask(env)
    st = lib.pp_oracle_binedge_report(xyz.shape[0], xyz.ctypes.data, types.ctypes.data,
                                      s2.ctypes.data, tm.ctypes.data, em.ctypes.data, rmin, rincr,
                                      rmax, center_stride, nthreads, out.ctypes.data)
    if st != 0:
        raise ValueError("bad arguments")
    return {"pairs": int(out[0]), "bin_differs_from_fp64": int(out[1]),
            "within_one_ulp_of_edge": int(out[2]), "ball_membership_flips": int(out[3])}


def closest_tri(tri, p) -> float:
    """FP32 squared distance from p to one triangle (9 floats), the reference's
    region logic and operation order."""
    lib = _load()
    t = np.ascontiguousarray(tri, dtype=np.float32).reshape(9)
    q = np.ascontiguousarray(p, dtype=np.float32).reshape(3)
    return float(lib.pp_oracle_closest_tri(t.ctypes.data, q.ctypes.data))


def ion2surf(xyz, triangles, R) -> np.ndarray:
    """Squared ion-to-hull distance capped at fl32(R*R), by brute force over
    the triangles in index order."""
    lib = _load()
    xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
    tri = np.ascontiguousarray(triangles, dtype=np.float32).reshape(-1, 9)
    out = np.empty(xyz.shape[0], np.float32)
    lib.pp_oracle_ion2surf(xyz.ctypes.data, xyz.shape[0], tri.ctypes.data, tri.shape[0],
                           C.c_float(R), out.ctypes.data)
    return out


def pos_decode(pos: bytes, ranges=None, epos: bool = False):
    """POS (or EPOS) bytes -> (xyz, mq, types) by the reference's byte
    reversal and first-matching closed range."""
    lib = _load()
    buf = np.frombuffer(pos, dtype=np.uint8)
    n = buf.shape[0] // (44 if epos else 16)
    if ranges is None:
        lo = hi = np.zeros(0, np.float32)
        rt = np.zeros(0, np.uint8)
    else:
        lo, hi, rt = (np.ascontiguousarray(ranges[0], np.float32),
                      np.ascontiguousarray(ranges[1], np.float32),
                      np.ascontiguousarray(ranges[2], np.uint8))
    xyz = np.empty((n, 3), np.float32)
    mq = np.empty(n, np.float32)
    ty = np.empty(n, np.uint8)
    fn = lib.pp_oracle_epos_decode if epos else lib.pp_oracle_pos_decode
    fn(buf.ctypes.data, n, lo.ctypes.data, hi.ctypes.data, rt.ctypes.data, lo.shape[0],
       xyz.ctypes.data, mq.ctypes.data, ty.ctypes.data)
    return xyz, mq, ty


def npcorr(xyz, types, ion2surf, tasks, rmax, rincr, knn_order):
    """NPCORR counts [ntasks][edge][edge][edge] (z, y, x) by brute force."""
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
    st = lib.pp_oracle_npcorr(xyz.shape[0], xyz.ctypes.data, types.ctypes.data, s2.ctypes.data,
                              T, trg.ctypes.data, env.ctypes.data, rmax, rincr, knn_order,
                              out.ctypes.data)
    if st != 0:
        raise MemoryError("oracle npcorr")
    return out

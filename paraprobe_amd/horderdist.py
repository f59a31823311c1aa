"""ctypes binding of the C++ host mirror (include/pp_horderdist.h): the
reference's task-code parsing, settings keys and CSV writers, and the
end-to-end XML -> GPU -> CSV driver."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import capi

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libpp_horderdist.so")
EXPORTED = ["pph_parse_tasks", "pph_parse_mknn_code", "pph_read_settings",
            "pph_write_descrstat2", "pph_write_descrstat3", "pph_run_spatstat",
            "pph_label_randomization", "pph_shuffle_types", "pph_mt19937",
            "pph_read_rangefile", "pph_run_pos"]
RND_SEED = 4294967295     # Settings::RndDescrStatsPRNGSeed, (unsigned int) -1
RND_DISCARD = 700000      # Settings::RndDescrStatsPRNGDiscard
_lib = None


def load():
    global _lib
    if _lib is None:
        capi.load_library()
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing; run the build first")
        lib = C.CDLL(LIB_PATH)
        P = C.c_void_p
        lib.pph_parse_tasks.argtypes = [C.c_char_p, C.c_int32, P, P, C.c_int32, P, P, P, P,
                                        C.c_int32]
        lib.pph_parse_tasks.restype = C.c_int32
        lib.pph_parse_mknn_code.argtypes = [C.c_char_p, P, C.c_int32]
        lib.pph_parse_mknn_code.restype = C.c_int32
        lib.pph_read_settings.argtypes = [C.c_char_p, P, P, P, C.c_int32, P, C.c_int32]
        lib.pph_read_settings.restype = C.c_int32
        lib.pph_write_descrstat2.argtypes = [C.c_char_p, C.c_int32, C.c_char_p, C.c_float,
                                             C.c_float, C.c_float, P, C.c_uint64, C.c_float]
        lib.pph_write_descrstat2.restype = C.c_int32
        lib.pph_write_descrstat3.argtypes = [C.c_char_p, P, C.c_int64]
        lib.pph_write_descrstat3.restype = C.c_int32
        lib.pph_run_spatstat.argtypes = [P, P, C.c_char_p, C.c_uint32, C.c_int32, P, P,
                                         C.c_uint64, C.c_float, C.c_char_p]
        lib.pph_run_spatstat.restype = C.c_int32
        lib.pph_label_randomization.argtypes = [C.c_char_p]
        lib.pph_label_randomization.restype = C.c_int32
        lib.pph_shuffle_types.argtypes = [C.c_uint32, C.c_uint32, P, C.c_uint64, P]
        lib.pph_shuffle_types.restype = C.c_int32
        lib.pph_mt19937.argtypes = [C.c_uint32, C.c_uint64, P, C.c_int32]
        lib.pph_mt19937.restype = C.c_int32
        lib.pph_read_rangefile.argtypes = [C.c_char_p, P, P, P, C.c_int32,
                                           C.POINTER(C.c_int32), P, C.c_int32]
        lib.pph_read_rangefile.restype = C.c_int32
        lib.pph_run_pos.argtypes = [P, C.c_char_p, C.c_char_p, C.c_char_p, P, C.c_uint64,
                                    C.c_uint32, C.c_uint64, C.c_float, C.c_char_p]
        lib.pph_run_pos.restype = C.c_int32
        _lib = lib
    return _lib


def _registry(types: dict):
    names = (C.c_char_p * max(len(types), 1))(*[k.encode() for k in types])
    ids = np.array(list(types.values()) or [0], dtype=np.uint32)
    return names, ids


def parse_tasks(code: str, types: dict, max_tasks: int = 64):
    """-> list of (trg names, env names, trg mask, env mask)."""
    lib = load()
    names, ids = _registry(types)
    trg = np.zeros((max_tasks, 32), np.uint8)
    env = np.zeros((max_tasks, 32), np.uint8)
    stride = 256
    tn = C.create_string_buffer(max_tasks * stride)
    en = C.create_string_buffer(max_tasks * stride)
    n = lib.pph_parse_tasks(code.encode(), len(types), names, ids.ctypes.data, max_tasks,
                            trg.ctypes.data, env.ctypes.data, tn, en, stride)
    out = []
    for t in range(min(n, max_tasks)):
        a = tn.raw[t * stride:(t + 1) * stride].split(b"\0", 1)[0].decode()
        b = en.raw[t * stride:(t + 1) * stride].split(b"\0", 1)[0].decode()
        out.append((a, b, trg[t].copy(), env[t].copy()))
    return out


def parse_mknn_code(code: str):
    lib = load()
    buf = np.zeros(4096, np.int64)
    n = lib.pph_parse_mknn_code(code.encode(), buf.ctypes.data, buf.shape[0])
    return None if n < 0 else buf[:n].tolist()


def read_settings(xml: str):
    lib = load()
    fam = C.c_uint32()
    radii = np.zeros(3, np.float32)
    tc = C.create_string_buffer(4096)
    mc = C.create_string_buffer(4096)
    st = lib.pph_read_settings(xml.encode(), C.byref(fam), radii.ctypes.data, tc, 4096, mc, 4096)
    return st, fam.value, radii.tolist(), tc.value.decode(), mc.value.decode()


def write_descrstat2(path, kind, metric, rmin, rincr, rmax, hist, nions_inside=0,
                     volume_inside=0.0):
    lib = load()
    h = np.ascontiguousarray(hist, dtype=np.int64)
    return lib.pph_write_descrstat2(path.encode(), kind, metric.encode(), rmin, rincr, rmax,
                                    h.ctypes.data, nions_inside, volume_inside)


def write_descrstat3(path, cnts):
    lib = load()
    c = np.ascontiguousarray(cnts, dtype=np.int64)
    return lib.pph_write_descrstat3(path.encode(), c.ctypes.data, c.shape[0])


def run_spatstat(ctx, cloud, xml: str, simid: int, types: dict, nions_inside: int,
                 volume_inside: float, outdir: str) -> int:
    lib = load()
    names, ids = _registry(types)
    return lib.pph_run_spatstat(ctx.ctx, cloud.handle, xml.encode(), simid, len(types), names,
                                ids.ctypes.data, nions_inside, volume_inside, outdir.encode())


def label_randomization(xml: str) -> int:
    return load().pph_label_randomization(xml.encode())


def shuffle_types(types, seed: int = RND_SEED, discard: int = RND_DISCARD) -> np.ndarray:
    """The rndlabeler's std::shuffle of a type array (std::mt19937, seed, discard)."""
    t = np.ascontiguousarray(types, dtype=np.uint8)
    out = np.empty_like(t)
    if load().pph_shuffle_types(seed, discard, t.ctypes.data, t.shape[0], out.ctypes.data):
        raise ValueError("bad arguments")
    return out


def mt19937(seed: int, discard: int, count: int) -> np.ndarray:
    out = np.empty(count, np.uint32)
    if load().pph_mt19937(seed, discard, out.ctypes.data, count):
        raise ValueError("bad arguments")
    return out


def read_rangefile(path: str, cap: int = 4096):
    """RRNG file -> (type names by id, (lo, hi, type) ranges in ranging order)."""
    lib = load()
    lo = np.empty(cap, np.float32)
    hi = np.empty(cap, np.float32)
    ty = np.empty(cap, np.uint8)
    nr = C.c_int32(0)
    stride = 64
    names = C.create_string_buffer(256 * stride)
    nt = lib.pph_read_rangefile(path.encode(), lo.ctypes.data, hi.ctypes.data, ty.ctypes.data,
                                cap, C.byref(nr), names, stride)
    if nt < 0:
        raise ValueError(f"cannot read range file {path}")
    raw = names.raw
    nm = [raw[i * stride:(i + 1) * stride].split(b"\0", 1)[0].decode() for i in range(nt)]
    k = min(nr.value, cap)
    return nm, (lo[:k].copy(), hi[:k].copy(), ty[:k].copy())


def run_pos(ctx, xml: str, pos_path: str, rrng_path, triangles, simid: int,
            nions_inside: int, volume_inside: float, outdir: str) -> int:
    tri = np.ascontiguousarray(triangles if triangles is not None else
                               np.zeros((0, 9), np.float32), dtype=np.float32).reshape(-1, 9)
    return load().pph_run_pos(ctx.ctx, xml.encode(), pos_path.encode(),
                              rrng_path.encode() if rrng_path else None, tri.ctypes.data,
                              tri.shape[0], simid, nions_inside, volume_inside, outdir.encode())

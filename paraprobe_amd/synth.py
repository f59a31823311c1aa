"""Seeded synthetic tips for the measured workloads (ctypes over libpp_synth.so).

Type labels: 1 = A, 2 = B, 3 = C (0 is the reference's UNKNOWNTYPE,
src/PARAPROBE_Numerics.h:78).
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libpp_synth.so")
TYPE_NAMES = {1: "A", 2: "B", 3: "C"}

_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing; run the build first")
        lib = C.CDLL(LIB_PATH)
        lib.pp_synth_tip.argtypes = [
            C.c_uint64, C.c_uint64, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_int32,
            C.c_void_p, C.c_uint64, C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
            C.c_void_p,
        ]
        lib.pp_synth_tip.restype = C.c_int
        lib.pp_synth_lattice.argtypes = [C.c_int32, C.c_double, C.c_int32, C.c_int32, C.c_int32,
                                         C.c_void_p, C.c_int64]
        lib.pp_synth_lattice.restype = C.c_int64
        lib.pp_synth_hull.argtypes = [C.c_int32, C.c_double, C.c_double, C.c_double, C.c_int32,
                                      C.c_int32, C.c_void_p, C.c_int64]
        lib.pp_synth_hull.restype = C.c_int64
        _lib = lib
    return _lib


@dataclass
class Tip:
    xyz: np.ndarray       # (n, 3) float32
    types: np.ndarray     # (n,) uint8
    ion2surf_sqr: np.ndarray  # (n,) float32
    desc: str

    @property
    def n(self) -> int:
        return self.xyz.shape[0]


def tip(n: int, seed: int, shape: str, r0: float, r1: float, h: float, frac,
        nclusters: int = 0, crmin: float = 1.0, crmax: float = 2.0, cfrac=None) -> Tip:
    lib = _load()
    frac = np.ascontiguousarray(frac, dtype=np.float64)
    cf = None if cfrac is None else np.ascontiguousarray(cfrac, dtype=np.float64)
    xyz = np.empty((n, 3), dtype=np.float32)
    ty = np.empty(n, dtype=np.uint8)
    s2 = np.empty(n, dtype=np.float32)
    st = lib.pp_synth_tip(n, seed, 0 if shape == "cylinder" else 1, r0, r1, h, len(frac),
                          frac.ctypes.data, nclusters, crmin, crmax,
                          None if cf is None else cf.ctypes.data, xyz.ctypes.data, ty.ctypes.data,
                          s2.ctypes.data)
    if st != 0:
        raise ValueError("bad tip parameters")
    d = f"{shape} n={n} r={r0}->{r1} h={h} frac={list(frac)} clusters={nclusters} seed={seed}"
    return Tip(xyz, ty, s2, d)


def lattice(fcc: bool, a: float, nx: int, ny: int, nz: int) -> np.ndarray:
    lib = _load()
    cap = nx * ny * nz * (4 if fcc else 1)
    out = np.empty((cap, 3), dtype=np.float32)
    w = lib.pp_synth_lattice(1 if fcc else 0, a, nx, ny, nz, out.ctypes.data, cap)
    return out[:w]


def hull(shape: str, r0: float, r1: float, h: float, nseg: int, nz: int) -> np.ndarray:
    """Closed triangle hull of a cylinder / frustum (z from 0 to h, radius r0
    at z = 0, r1 at z = h): nseg x nz lateral quads as two triangles each, both
    caps as fans.  Returns (ntri, 9) float32 vertex coordinates."""
    lib = _load()
    s = 0 if shape == "cylinder" else 1
    nt = lib.pp_synth_hull(s, r0, r1, h, nseg, nz, None, 0)
    if nt < 0:
        raise ValueError("bad hull parameters")
    out = np.empty((nt, 9), dtype=np.float32)
    if lib.pp_synth_hull(s, r0, r1, h, nseg, nz, out.ctypes.data, nt) != nt:
        raise ValueError("bad hull parameters")
    return out


# ---------------------------------------------------------------------------
# the measured workloads

def cylinder_1m(seed: int = 1, n: int = 1_000_000) -> Tip:
    """1M ions uniform in a cylinder r=15 nm, h=64 nm; types A/B/C 0.90/0.08/0.02."""
    return tip(n, seed, "cylinder", 15.0, 15.0, 64.0, [0.90, 0.08, 0.02])


def cluster_fracs(base, enriched_type: int = 2, enriched: float = 0.40):
    """Type fractions inside a cluster: the enriched type at `enriched`, the
    others renormalised to the remainder."""
    base = list(base)
    i = enriched_type - 1
    rest = 1.0 - base[i]
    return [enriched if k == i else base[k] * (1.0 - enriched) / rest for k in range(len(base))]


def frustum_20m(seed: int = 2, n: int = 20_000_000, nclusters: int = 50_000) -> Tip:
    """20M ions in a conical frustum r 40 -> 20 nm over h = 250 nm (27.3 ions/nm^3
    at this N), types 0.95/0.04/0.01, Poisson clusters r ~ U(1,2) nm with B at 40%."""
    base = [0.95, 0.04, 0.01]
    return tip(n, seed, "frustum", 40.0, 20.0, 250.0, base, nclusters, 1.0, 2.0,
               cluster_fracs(base))


def frustum_100m(seed: int = 4, n: int = 100_000_000, nclusters: int = 300_000,
                 density: float = 22.0) -> Tip:
    """The large configuration: 100M ions in a frustum r 60 -> 30 nm whose
    height (689 nm) gives 22 ions/nm^3, types 0.95/0.04/0.01, 300,000 Poisson
    clusters r ~ U(1,2) nm with B at 40%.  A smaller n keeps the shape and the
    density (every length scaled by (n / 100M)^(1/3), the clusters by n)."""
    base = [0.95, 0.04, 0.01]
    h0 = 100_000_000 / density / frustum_volume(60.0, 30.0, 1.0)
    s = (n / 100_000_000) ** (1.0 / 3.0)
    nc = int(round(nclusters * n / 100_000_000))
    return tip(n, seed, "frustum", 60.0 * s, 30.0 * s, h0 * s, base, nc, 1.0, 2.0,
               cluster_fracs(base))


def frustum_volume(r0: float, r1: float, h: float) -> float:
    return math.pi * h / 3.0 * (r0 * r0 + r0 * r1 + r1 * r1)


ALL_PAIRS = [((a,), (b,)) for a in (1, 2, 3) for b in (1, 2, 3)]


def scaled_frustum(n: int, seed: int, clustered: bool, density: float = 27.3) -> Tip:
    """A frustum tip of the headline shape (r 40 -> 20 over h 250, i.e. radii
    and height in the ratio 0.16 : 0.08 : 1) scaled to hold n ions at the given
    density; clustered tips carry 50,000 clusters per 20M ions (B at 40%)."""
    base = [0.95, 0.04, 0.01]
    v = n / density
    # V = pi h / 3 (r0^2 + r0 r1 + r1^2) with r0 = 0.16 h, r1 = 0.08 h
    h = (v / (math.pi / 3.0 * (0.16 ** 2 + 0.16 * 0.08 + 0.08 ** 2))) ** (1.0 / 3.0)
    nc = int(round(50_000 * n / 20_000_000)) if clustered else 0
    return tip(n, seed, "frustum", 0.16 * h, 0.08 * h, h, base, nc, 1.0, 2.0,
               cluster_fracs(base) if clustered else None)

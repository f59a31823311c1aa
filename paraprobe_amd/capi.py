"""ctypes binding of the C ABI in ``include/pp_spatstat.h``.

This is the Python-side caller of the gfx950 library (the C++ host mirror in
``paraprobe_amd/host`` is the other).  It never falls back to anything: if
``libpp_spatstat.so`` is missing or no gfx950 device is present, it raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libpp_spatstat.so")

PP_OK = 0
STATUS_NAMES = {
    0: "PP_OK",
    1: "PP_ERR_INVALID_ARG",
    2: "PP_ERR_OOM_DEVICE",
    3: "PP_ERR_HIP",
    4: "PP_ERR_ADMISSION_REJECTED",
    5: "PP_ERR_NO_DEVICE",
    6: "PP_ERR_UNSUPPORTED",
    7: "PP_ERR_OVERFLOW",
    8: "PP_ERR_NOT_READY",
}

FAM_RDF = 1 << 0
FAM_1NN = 1 << 1
FAM_RIPK = 1 << 2
FAM_MKNN = 1 << 3
FAM_NPCORR = 1 << 4
FAM_CNTNB = 1 << 5

# every symbol declared in include/pp_spatstat.h
EXPORTED = [
    "pp_spatstat_create",
    "pp_spatstat_destroy",
    "pp_last_error_string",
    "pp_hist_geometry",
    "pp_hist_supported",
    "pp_cloud_upload",
    "pp_cloud_from_device",
    "pp_cloud_free",
    "pp_cloud_footprint",
    "pp_spatstat_run",
    "pp_spatstat_run_noalloc",
    "pp_batch_run",
    "pp_set_tuning",
    "pp_set_knn_range",
    "pp_set_cell_aspect",
    "pp_set_pair_symmetry",
    "pp_set_graph_launch",
    "pp_cloud_add_cell_list",
    "pp_serve_create",
    "pp_serve_submit",
    "pp_serve_poll",
    "pp_serve_drain",
    "pp_serve_destroy",
    "pp_ion2surf",
    "pp_ion2surf_device",
    "pp_cloud_count",
    "pp_cloud_types",
    "pp_cloud_relabel",
    "pp_pos_decode",
    "pp_pos_decode_device",
    "pp_cloud_from_pos",
    "pp_npcorr_geometry",
    "pp_npcorr_run",
    "pp_epos_decode",
    "pp_cloud_from_epos",
]


class _ServeOpts(C.Structure):
    _fields_ = [("slots", C.c_int32), ("max_batch_jobs", C.c_int32),
                ("hbm_budget_bytes", C.c_uint64), ("batch_work", C.c_double)]


class PPError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class _Job(C.Structure):
    _fields_ = [
        ("ntasks", C.c_int32),
        ("trg_mask", C.POINTER(C.c_uint8)),
        ("env_mask", C.POINTER(C.c_uint8)),
        ("rmin", C.c_float),
        ("rincr", C.c_float),
        ("rmax", C.c_float),
        ("families", C.c_uint32),
        ("nk", C.c_int32),
        ("kminus1_desc", C.POINTER(C.c_int64)),
    ]


class _Out(C.Structure):
    _fields_ = [
        ("rdf", C.POINTER(C.c_int64)),
        ("nn1", C.POINTER(C.c_int64)),
        ("ripk", C.POINTER(C.c_int64)),
        ("mknn", C.POINTER(C.c_int64)),
        ("cntnb", C.POINTER(C.c_int64)),
        ("cntnb_cap", C.c_int64),
        ("nbins", C.c_int64),
        ("cntnb_len", C.c_int64),
        ("centers_considered", C.c_int64),
        ("ions_discarded", C.c_int64),
        ("pairs_inball", C.c_int64),
        ("pairs_candidate", C.c_int64),
        ("ms_index", C.c_double),
        ("ms_prep", C.c_double),
        ("ms_query", C.c_double),
        ("ms_total", C.c_double),
        ("tile_centers", C.c_int32),
        ("cells_per_rmax", C.c_int32),
        ("counters16", C.c_int32),
        ("percenter", C.c_int32),
        ("centers_rerun", C.c_int64),
        ("ms_kernel", C.c_double),
    ]


_lib = None


def load_library() -> C.CDLL:
    """Loads libpp_spatstat.so (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    lib.pp_spatstat_create.argtypes = [C.c_int32, C.POINTER(P)]
    lib.pp_spatstat_destroy.argtypes = [P]
    lib.pp_last_error_string.argtypes = [P]
    lib.pp_last_error_string.restype = C.c_char_p
    lib.pp_hist_geometry.argtypes = [C.c_float, C.c_float, C.c_float, C.POINTER(C.c_int64)]
    lib.pp_hist_supported.argtypes = [C.c_float, C.c_float, C.c_float]
    lib.pp_cloud_upload.argtypes = [P, P, P, P, C.c_uint64, C.c_float, C.POINTER(P)]
    lib.pp_cloud_from_device.argtypes = [P, P, P, P, C.c_uint64, C.c_float, C.POINTER(P)]
    lib.pp_cloud_free.argtypes = [P, P]
    lib.pp_cloud_footprint.argtypes = [P, C.POINTER(C.c_uint64)]
    lib.pp_spatstat_run.argtypes = [P, P, C.POINTER(_Job), C.POINTER(_Out)]
    lib.pp_spatstat_run_noalloc.argtypes = [P, P, C.POINTER(_Job), C.POINTER(_Out)]
    lib.pp_batch_run.argtypes = [P, C.c_int32, C.POINTER(P), C.POINTER(_Job), C.POINTER(_Out)]
    lib.pp_set_tuning.argtypes = [P, C.c_int32, C.c_int32]
    lib.pp_set_knn_range.argtypes = [P, C.c_float]
    lib.pp_set_cell_aspect.argtypes = [P, C.c_int32]
    lib.pp_set_pair_symmetry.argtypes = [P, C.c_int32]
    lib.pp_set_graph_launch.argtypes = [P, C.c_int32]
    lib.pp_cloud_add_cell_list.argtypes = [P, P, C.c_float]
    lib.pp_serve_create.argtypes = [P, C.POINTER(_ServeOpts), C.POINTER(P)]
    lib.pp_serve_submit.argtypes = [P, P, C.POINTER(_Job), C.POINTER(_Out), C.POINTER(C.c_uint64)]
    lib.pp_serve_poll.argtypes = [P, P, P, C.c_int32, C.POINTER(C.c_int32)]
    lib.pp_serve_drain.argtypes = [P]
    lib.pp_serve_destroy.argtypes = [P]
    lib.pp_ion2surf.argtypes = [P, P, C.c_uint64, P, C.c_uint64, C.c_float, P]
    lib.pp_ion2surf_device.argtypes = [P, P, C.c_uint64, P, C.c_uint64, C.c_float, P]
    lib.pp_cloud_count.argtypes = [P, C.POINTER(C.c_uint64)]
    lib.pp_cloud_types.argtypes = [P, P, P]
    lib.pp_cloud_relabel.argtypes = [P, P, P, C.POINTER(P)]
    lib.pp_pos_decode.argtypes = [P, P, C.c_uint64, P, P, P, C.c_int32, P, P, P]
    lib.pp_pos_decode_device.argtypes = [P, P, C.c_uint64, P, P, P, C.c_int32, P, P, P]
    lib.pp_npcorr_geometry.argtypes = [C.c_float, C.c_float]
    lib.pp_npcorr_geometry.restype = C.c_int64
    lib.pp_npcorr_run.argtypes = [P, P, C.c_int32, P, P, C.c_float, C.c_float, C.c_int64, P,
                                  C.POINTER(C.c_int64)]
    lib.pp_epos_decode.argtypes = [P, P, C.c_uint64, P, P, P, C.c_int32, P, P, P]
    lib.pp_cloud_from_epos.argtypes = [P, P, C.c_uint64, P, P, P, C.c_int32, P, C.c_uint64,
                                       C.c_float, C.c_float, C.POINTER(P)]
    lib.pp_cloud_from_pos.argtypes = [P, P, C.c_uint64, P, P, P, C.c_int32, P, C.c_uint64,
                                      C.c_float, C.c_float, C.POINTER(P)]
    for name in EXPORTED:
        if name != "pp_last_error_string":
            getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib


def hist_geometry(rmin: float, rincr: float, rmax: float) -> int:
    lib = load_library()
    nb = C.c_int64()
    st = lib.pp_hist_geometry(rmin, rincr, rmax, C.byref(nb))
    if st != PP_OK:
        raise PPError(st, "invalid histogram geometry")
    return nb.value


def type_mask(types) -> np.ndarray:
    """32-byte bit mask of type ids."""
    m = np.zeros(32, dtype=np.uint8)
    for t in types:
        t = int(t)
        if not 0 <= t < 256:
            raise ValueError(f"type id {t} out of range")
        m[t >> 3] |= np.uint8(1 << (t & 7))
    return m


@dataclass
class Job:
    """One spatial-statistics job: tasks as (central types, neighbour types)."""

    tasks: list  # list of (iterable of trg type ids, iterable of env type ids)
    rmin: float
    rincr: float
    rmax: float
    families: int
    kvalues: list = field(default_factory=list)  # k values (1-based), any order

    def kminus1_desc(self) -> np.ndarray:
        return np.array(sorted((int(k) - 1 for k in self.kvalues), reverse=True), dtype=np.int64)

    def masks(self):
        trg = np.zeros((max(len(self.tasks), 1), 32), dtype=np.uint8)
        env = np.zeros((max(len(self.tasks), 1), 32), dtype=np.uint8)
        for i, (t, e) in enumerate(self.tasks):
            trg[i] = type_mask(t)
            env[i] = type_mask(e)
        return trg, env


@dataclass
class Result:
    nbins: int
    rdf: np.ndarray | None
    nn1: np.ndarray | None
    ripk: np.ndarray | None
    mknn: np.ndarray | None
    cntnb: np.ndarray | None
    cntnb_len: int
    centers_considered: int
    ions_discarded: int
    pairs_inball: int
    pairs_candidate: int
    ms_index: float
    ms_prep: float
    ms_query: float
    ms_total: float
    kernel: str = ""
    percenter: int = 0      # 0 none, 1 full-range per-center histograms, 2 kNN range only
    centers_rerun: int = 0  # centers re-run over the full range (mode 2)
    ms_kernel: float = 0.0  # the query kernel launches alone

    def arrays(self) -> dict:
        return {k: getattr(self, k) for k in ("rdf", "nn1", "ripk", "mknn", "cntnb")
                if getattr(self, k) is not None}


class _Prepared:
    """Keeps the numpy buffers a ctypes job/out pair points to alive."""

    def __init__(self, job: Job, cntnb_cap: int):
        self.job = job
        self.trg, self.env = job.masks()
        self.kd = job.kminus1_desc()
        nb = hist_geometry(job.rmin, job.rincr, job.rmax)
        self.nbins = nb
        C_ = nb + 2
        T = len(job.tasks)
        fam = job.families
        self.arr = {}
        self.cj = _Job()
        self.cj.ntasks = T
        self.cj.trg_mask = self.trg.ctypes.data_as(C.POINTER(C.c_uint8))
        self.cj.env_mask = self.env.ctypes.data_as(C.POINTER(C.c_uint8))
        self.cj.rmin, self.cj.rincr, self.cj.rmax = job.rmin, job.rincr, job.rmax
        self.cj.families = fam
        self.cj.nk = len(self.kd)
        self.cj.kminus1_desc = self.kd.ctypes.data_as(C.POINTER(C.c_int64))
        self.co = _Out()
        for name, bit, shape in (
            ("rdf", FAM_RDF, (T, C_)),
            ("nn1", FAM_1NN, (T, C_)),
            ("ripk", FAM_RIPK, (T, C_)),
            ("mknn", FAM_MKNN, (T, len(self.kd), C_)),
            ("cntnb", FAM_CNTNB, (T, cntnb_cap)),
        ):
            if fam & bit:
                a = np.zeros(shape, dtype=np.int64)
                self.arr[name] = a
                setattr(self.co, name, a.ctypes.data_as(C.POINTER(C.c_int64)))
        self.co.cntnb_cap = cntnb_cap if fam & FAM_CNTNB else 0

    def result(self) -> Result:
        o = self.co
        return Result(
            nbins=o.nbins,
            rdf=self.arr.get("rdf"),
            nn1=self.arr.get("nn1"),
            ripk=self.arr.get("ripk"),
            mknn=self.arr.get("mknn"),
            cntnb=self.arr.get("cntnb"),
            cntnb_len=o.cntnb_len,
            centers_considered=o.centers_considered,
            ions_discarded=o.ions_discarded,
            pairs_inball=o.pairs_inball,
            pairs_candidate=o.pairs_candidate,
            ms_index=o.ms_index,
            ms_prep=o.ms_prep,
            ms_query=o.ms_query,
            ms_total=o.ms_total,
            kernel="k_query<%d,%s,%s> cells/Rmax=%d" % (
                o.tile_centers, "true" if o.counters16 else "false",
                ("direct", "percenter", "hybrid")[o.percenter], o.cells_per_rmax),
            percenter=o.percenter,
            centers_rerun=o.centers_rerun,
            ms_kernel=o.ms_kernel,
        )


def _ranges(ranges):
    if ranges is None:
        z = np.zeros(1, np.float32)
        return z[:0], z[:0], np.zeros(0, np.uint8)
    lo, hi, rt = ranges
    return (np.ascontiguousarray(lo, dtype=np.float32), np.ascontiguousarray(hi, dtype=np.float32),
            np.ascontiguousarray(rt, dtype=np.uint8))


class Cloud:
    def __init__(self, ctx: "SpatStat", handle: C.c_void_p, n: int):
        self._ctx = ctx
        self.handle = handle
        self.n = n

    def add_cell_list(self, width: float):
        self._ctx._check(self._ctx.lib.pp_cloud_add_cell_list(self._ctx.ctx, self.handle, width))

    def types(self) -> np.ndarray:
        """The cloud's ion types in upload order."""
        out = np.empty(self.n, np.uint8)
        self._ctx._check(self._ctx.lib.pp_cloud_types(self._ctx.ctx, self.handle,
                                                      out.ctypes.data))
        return out

    def relabel(self, types: np.ndarray) -> "Cloud":
        """A new cloud with these positions and ion2surf and the given types."""
        types = np.ascontiguousarray(types, dtype=np.uint8)
        if types.shape[0] != self.n:
            raise ValueError("one type per ion")
        h = C.c_void_p()
        self._ctx._check(self._ctx.lib.pp_cloud_relabel(self._ctx.ctx, self.handle,
                                                        types.ctypes.data, C.byref(h)))
        return Cloud(self._ctx, h, self.n)

    def footprint(self) -> int:
        b = C.c_uint64()
        self._ctx._check(self._ctx.lib.pp_cloud_footprint(self.handle, C.byref(b)))
        return b.value

    def free(self):
        if self.handle:
            self._ctx.lib.pp_cloud_free(self._ctx.ctx, self.handle)
            self.handle = None


class SpatStat:
    """A context on one gfx950 device."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        ctx = C.c_void_p()
        st = self.lib.pp_spatstat_create(device, C.byref(ctx))
        if st != PP_OK:
            raise PPError(st, f"cannot create a context on device {device}")
        self.ctx = ctx

    def _check(self, st: int):
        if st != PP_OK:
            msg = self.lib.pp_last_error_string(self.ctx)
            raise PPError(st, msg.decode() if msg else "")

    def close(self):
        if self.ctx:
            self.lib.pp_spatstat_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_tuning(self, centers_per_tile: int = 0, cells_per_rmax: int = 0):
        self._check(self.lib.pp_set_tuning(self.ctx, centers_per_tile, cells_per_rmax))

    def set_cell_aspect(self, yz_per_x: int = 0):
        """y/z cells per x cell width of the cell lists jobs build (0: built-in:
        cubic, except the density rule's Rmax/4 and Rmax/2 cells, built twice
        as long along x)."""
        self._check(self.lib.pp_set_cell_aspect(self.ctx, yz_per_x))

    def set_pair_symmetry(self, on: bool = True):
        """Symmetric pair counting of the dominant center type (on by default);
        results do not depend on it."""
        self._check(self.lib.pp_set_pair_symmetry(self.ctx, 1 if on else 0))

    def set_knn_range(self, factor: float = 4.0):
        """kNN range factor (expected neighbours per k) of MKNN / 1NN jobs; 0
        keeps full-range per-center histograms."""
        self._check(self.lib.pp_set_knn_range(self.ctx, factor))

    def set_graph_launch(self, on: bool = True):
        """Capture each launch group of a multi-job run (batch, serving
        batch) into a HIP graph (default) or issue its launches one by one;
        single-job runs always launch one by one.  Results are identical."""
        self._check(self.lib.pp_set_graph_launch(self.ctx, 1 if on else 0))

    def ion2surf(self, xyz: np.ndarray, triangles: np.ndarray, R: float) -> np.ndarray:
        """Squared distance of every ion to the triangle hull, capped at
        fl32(R * R) (solver::surface_distancing2); R is max(SpatStatRadiusMax,
        SurfaceCellsCarvingRadius) in the reference."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        tri = np.ascontiguousarray(triangles, dtype=np.float32).reshape(-1, 9)
        out = np.empty(xyz.shape[0], np.float32)
        self._check(self.lib.pp_ion2surf(self.ctx, xyz.ctypes.data, xyz.shape[0], tri.ctypes.data,
                                         tri.shape[0], C.c_float(R), out.ctypes.data))
        return out

    def ion2surf_device(self, d_xyz: int, n: int, d_triangles: int, ntri: int, R: float,
                        d_out: int):
        """ion2surf on device pointers (n x 3 and ntri x 9 floats in, n floats out)."""
        self._check(self.lib.pp_ion2surf_device(self.ctx, C.c_void_p(d_xyz), n,
                                                C.c_void_p(d_triangles), ntri, C.c_float(R),
                                                C.c_void_p(d_out)))

    def npcorr(self, cloud: Cloud, tasks, rmax: float, rincr: float, knn_order: int):
        """NPCORR counts [ntasks][edge][edge][edge] (z, y, x) uint32."""
        edge = self.lib.pp_npcorr_geometry(rmax, rincr)
        if edge < 0:
            raise PPError(6, "NPCORR histogram too large")
        T = len(tasks)
        trg = np.zeros((max(T, 1), 32), np.uint8)
        env = np.zeros((max(T, 1), 32), np.uint8)
        for i, (t, e) in enumerate(tasks):
            trg[i], env[i] = type_mask(t), type_mask(e)
        out = np.zeros((T, edge, edge, edge), np.uint32)
        nb = C.c_int64(0)
        self._check(self.lib.pp_npcorr_run(self.ctx, cloud.handle, T, trg.ctypes.data,
                                           env.ctypes.data, rmax, rincr, knn_order,
                                           out.ctypes.data, C.byref(nb)))
        return out

    def pos_decode(self, pos: bytes, ranges=None, with_mq: bool = True, epos: bool = False):
        """POS bytes (16 per ion, big-endian x, y, z, mq) or EPOS bytes (44 per
        ion) -> (xyz, mq, types); ranges = (lo, hi, type) arrays in the
        reference's order, or None."""
        buf = np.frombuffer(pos, dtype=np.uint8)
        n = buf.shape[0] // (44 if epos else 16)
        lo, hi, rt = _ranges(ranges)
        xyz = np.empty((n, 3), np.float32)
        mq = np.empty(n, np.float32) if with_mq else None
        ty = np.empty(n, np.uint8)
        fn = self.lib.pp_epos_decode if epos else self.lib.pp_pos_decode
        self._check(fn(self.ctx, buf.ctypes.data, buf.shape[0], lo.ctypes.data, hi.ctypes.data,
                       rt.ctypes.data, lo.shape[0], xyz.ctypes.data,
                       None if mq is None else mq.ctypes.data, ty.ctypes.data))
        return xyz, mq, ty

    def cloud_from_pos(self, pos: bytes, ranges, triangles, R: float,
                       cell_width: float = 0.0, epos: bool = False) -> Cloud:
        """POS (or EPOS) bytes + ranges + hull -> a cloud, decoded and
        distanced on the GPU."""
        buf = np.frombuffer(pos, dtype=np.uint8)
        lo, hi, rt = _ranges(ranges)
        tri = np.ascontiguousarray(triangles if triangles is not None else
                                   np.zeros((0, 9), np.float32), dtype=np.float32).reshape(-1, 9)
        h = C.c_void_p()
        fn = self.lib.pp_cloud_from_epos if epos else self.lib.pp_cloud_from_pos
        self._check(fn(self.ctx, buf.ctypes.data, buf.shape[0], lo.ctypes.data, hi.ctypes.data,
                       rt.ctypes.data, lo.shape[0], tri.ctypes.data, tri.shape[0], C.c_float(R),
                       C.c_float(cell_width), C.byref(h)))
        return Cloud(self, h, buf.shape[0] // (44 if epos else 16))

    def pos_decode_device(self, d_pos: int, nbytes: int, ranges, d_xyz: int, d_mq: int,
                          d_type: int):
        lo, hi, rt = _ranges(ranges)
        self._check(self.lib.pp_pos_decode_device(self.ctx, C.c_void_p(d_pos), nbytes,
                                                  lo.ctypes.data, hi.ctypes.data,
                                                  rt.ctypes.data, lo.shape[0], C.c_void_p(d_xyz),
                                                  C.c_void_p(d_mq) if d_mq else None,
                                                  C.c_void_p(d_type)))

    def upload(self, xyz: np.ndarray, types: np.ndarray, ion2surf_sqr: np.ndarray,
               cell_width: float = 0.0) -> Cloud:
        xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        types = np.ascontiguousarray(types, dtype=np.uint8)
        s2 = np.ascontiguousarray(ion2surf_sqr, dtype=np.float32)
        n = xyz.shape[0]
        if types.shape[0] != n or s2.shape[0] != n:
            raise ValueError("xyz, types and ion2surf_sqr lengths differ")
        h = C.c_void_p()
        self._check(self.lib.pp_cloud_upload(self.ctx, xyz.ctypes.data, types.ctypes.data,
                                             s2.ctypes.data, n, cell_width, C.byref(h)))
        return Cloud(self, h, n)

    def upload_device(self, d_xyz: int, d_types: int, d_s2: int, n: int,
                      cell_width: float = 0.0) -> Cloud:
        """Cloud from device pointers (e.g. torch tensors' data_ptr())."""
        h = C.c_void_p()
        self._check(self.lib.pp_cloud_from_device(self.ctx, C.c_void_p(d_xyz), C.c_void_p(d_types),
                                                  C.c_void_p(d_s2), n, cell_width, C.byref(h)))
        return Cloud(self, h, n)

    def run(self, cloud: Cloud, job: Job, cntnb_cap: int = 65536) -> Result:
        p = _Prepared(job, cntnb_cap)
        self._check(self.lib.pp_spatstat_run(self.ctx, cloud.handle, C.byref(p.cj), C.byref(p.co)))
        return p.result()

    def run_noalloc(self, cloud: Cloud, job: Job, cntnb_cap: int = 65536) -> Result:
        """pp_spatstat_run_noalloc: the same job with every device temporary
        in the library's fixed scratch (no device allocation)."""
        p = _Prepared(job, cntnb_cap)
        self._check(self.lib.pp_spatstat_run_noalloc(self.ctx, cloud.handle, C.byref(p.cj),
                                                     C.byref(p.co)))
        return p.result()

    def run_batch(self, clouds: list, jobs: list, cntnb_cap: int = 65536) -> list:
        n = len(jobs)
        preps = [_Prepared(j, cntnb_cap) for j in jobs]
        cj = (_Job * n)(*[p.cj for p in preps])
        co = (_Out * n)(*[p.co for p in preps])
        ch = (C.c_void_p * n)(*[c.handle for c in clouds])
        self._check(self.lib.pp_batch_run(self.ctx, n, ch, cj, co))
        out = []
        for i, p in enumerate(preps):
            p.co = co[i]
            out.append(p.result())
        return out


class Server:
    """Continuous-batching request loop over resident clouds (pp_serve_*)."""

    def __init__(self, ctx: SpatStat, slots: int = 4, max_batch_jobs: int = 64,
                 hbm_budget_bytes: int = 0, batch_work: float = 0.0):
        self.ctx = ctx
        self.lib = ctx.lib
        o = _ServeOpts(slots, max_batch_jobs, hbm_budget_bytes, batch_work)
        h = C.c_void_p()
        ctx._check(self.lib.pp_serve_create(ctx.ctx, C.byref(o), C.byref(h)))
        self.h = h
        self._live = {}

    def submit(self, cloud: Cloud, job: Job, cntnb_cap: int = 65536) -> int:
        p = _Prepared(job, cntnb_cap)
        t = C.c_uint64()
        st = self.lib.pp_serve_submit(self.h, cloud.handle, C.byref(p.cj), C.byref(p.co),
                                      C.byref(t))
        if st != PP_OK:
            msg = self.lib.pp_last_error_string(self.ctx.ctx)
            raise PPError(st, msg.decode() if msg else "")
        self._live[t.value] = p
        return t.value

    def poll(self, cap: int = 4096):
        """-> list of (ticket, status, Result)."""
        tk = np.zeros(cap, np.uint64)
        stv = np.zeros(cap, np.int32)
        n = C.c_int32()
        self.ctx._check(self.lib.pp_serve_poll(self.h, tk.ctypes.data, stv.ctypes.data, cap,
                                               C.byref(n)))
        out = []
        for i in range(n.value):
            p = self._live.pop(int(tk[i]))
            out.append((int(tk[i]), int(stv[i]), p.result()))
        return out

    def drain(self):
        self.ctx._check(self.lib.pp_serve_drain(self.h))

    def close(self):
        if self.h:
            self.lib.pp_serve_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

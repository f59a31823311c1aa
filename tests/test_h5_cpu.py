"""The optional HDF5 writer (libpp_h5.so) on the CPU: the /DescrSpatStats
groups and the NPCORR datasets of the reference's results file
(src/PARAPROBE_HDF5.cpp:369-408, src/PARAPROBE_SolverHdl.cpp:4506-4575), read
back through the HDF5 library.  Skipped where the image has no HDF5."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "paraprobe_amd", "libpp_h5.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB),
                                reason="libpp_h5.so not built (no HDF5 library)")

U32, U64, F32 = 1, 2, 3


def _lib():
    lib = C.CDLL(LIB)
    P = C.c_void_p
    lib.pph5_create_results.argtypes = [C.c_char_p, C.c_int32]
    lib.pph5_report_npc3d_bincenters.argtypes = [C.c_char_p, C.c_uint64, C.c_float]
    lib.pph5_report_npc3d_histvalues.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int32,
                                                 P, C.c_uint64]
    lib.pph5_read.argtypes = [C.c_char_p, C.c_char_p, P, C.c_uint64, C.POINTER(C.c_uint64),
                              C.POINTER(C.c_uint64), C.POINTER(C.c_int32)]
    return lib


def _read(lib, fn, dset, cap=1 << 26):
    buf = np.zeros(cap // 4, np.uint32)
    nr, nc, kind = C.c_uint64(), C.c_uint64(), C.c_int32()
    rc = lib.pph5_read(fn.encode(), dset.encode(), buf.ctypes.data, buf.nbytes, C.byref(nr),
                       C.byref(nc), C.byref(kind))
    if rc != 0:
        return rc, None, None
    n = nr.value * nc.value
    dt = {U32: np.uint32, U64: np.uint64, F32: np.float32}[kind.value]
    return 0, kind.value, buf.view(dt)[:n].reshape(nr.value, nc.value).copy()


def test_exports_every_declared_symbol():
    text = open(os.path.join(ROOT, "include", "pp_h5.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(pph5_\w+)\s*\(", text)
    assert len(names) == 4
    lib = C.CDLL(LIB)
    for n in names:
        assert hasattr(lib, n), n


def test_groups_bincenters_and_counts_round_trip(tmp_path):
    lib = _lib()
    fn = str(tmp_path / "PARAPROBE.SimID.7.Results.h5")
    assert lib.pph5_create_results(fn.encode(), 1) == 0
    # npc3d(2.5, 0.05): edge 103 (halfedgelen ceil(50) + 1 = 51)
    edge, width = 103, np.float32(0.05)
    assert lib.pph5_report_npc3d_bincenters(fn.encode(), edge, width) == 0
    rc, kind, bins = _read(lib, fn, "/DescrSpatStats/TwoPointStats/BinNXNYNZ")
    assert rc == 0 and kind == U64 and bins.shape == (3, 1)
    assert bins.ravel().tolist() == [103, 103, 103]
    rc, kind, cen = _read(lib, fn, "/DescrSpatStats/TwoPointStats/BinCenterXYZ")
    assert rc == 0 and kind == F32 and cen.shape == (edge ** 3, 3)
    # npc3d::get_bincenter: float(i) * width - float((edge - 1) / 2), x fastest
    i = np.arange(edge, dtype=np.float32)
    want_x = i * width - np.float32(51)
    np.testing.assert_array_equal(cen[:edge, 0], want_x)
    assert np.all(cen[:edge, 1] == want_x[0]) and np.all(cen[:edge, 2] == want_x[0])
    z, y, x = 5, 17, 99
    np.testing.assert_array_equal(cen[(z * edge + y) * edge + x],
                                  [want_x[x], want_x[y], want_x[z]])
    # the unscaled half edge: bin 51 (the origin's cell) is at 51 * 0.05 - 51
    assert cen[51, 0] == np.float32(np.float32(51) * width - np.float32(51))
    # one u32 column per task, named <trg>_<env>_<Orig|Rndm>
    rng = np.random.default_rng(3)
    a = rng.integers(0, 2 ** 32, size=edge ** 3, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, 50, size=edge ** 3, dtype=np.uint64).astype(np.uint32)
    assert lib.pph5_report_npc3d_histvalues(fn.encode(), b"AlFe", b"Fe", 0, a.ctypes.data,
                                            a.size) == 0
    assert lib.pph5_report_npc3d_histvalues(fn.encode(), b"AlFe", b"Fe", 1, b.ctypes.data,
                                            b.size) == 0
    rc, kind, ga = _read(lib, fn, "/DescrSpatStats/TwoPointStats/AlFe_Fe_Orig")
    assert rc == 0 and kind == U32 and ga.shape == (edge ** 3, 1)
    np.testing.assert_array_equal(ga.ravel(), a)
    rc, kind, gb = _read(lib, fn, "/DescrSpatStats/TwoPointStats/AlFe_Fe_Rndm")
    np.testing.assert_array_equal(gb.ravel(), b)
    # a dataset name written twice is an error (h5Hdl would fail the same way)
    assert lib.pph5_report_npc3d_histvalues(fn.encode(), b"AlFe", b"Fe", 0, a.ctypes.data,
                                            a.size) != 0
    # absent datasets, and a buffer too small
    assert _read(lib, fn, "/DescrSpatStats/TwoPointStats/Fe_Fe_Orig")[0] != 0
    assert _read(lib, fn, "/DescrSpatStats/TwoPointStats/AlFe_Fe_Orig", cap=64)[0] == -2
    # the Python reader sizes its buffer from the dataset
    from paraprobe_amd import h5

    np.testing.assert_array_equal(h5.read(fn, "/DescrSpatStats/TwoPointStats/AlFe_Fe_Rndm"),
                                  b[:, None])
    assert h5.read(fn, "/DescrSpatStats/TwoPointStats/Fe_Fe_Orig") is None


def test_results_file_without_npcorr_has_only_the_stats_group(tmp_path):
    lib = _lib()
    fn = str(tmp_path / "r.h5")
    assert lib.pph5_create_results(fn.encode(), 0) == 0
    # no TwoPointStats group: the NPCORR datasets cannot be created
    assert lib.pph5_report_npc3d_bincenters(fn.encode(), 5, np.float32(0.5)) != 0
    # truncation: creating again empties the file
    assert lib.pph5_create_results(fn.encode(), 1) == 0
    assert lib.pph5_report_npc3d_bincenters(fn.encode(), 5, np.float32(0.5)) == 0
    assert lib.pph5_create_results(fn.encode(), 1) == 0
    assert _read(lib, fn, "/DescrSpatStats/TwoPointStats/BinNXNYNZ")[0] != 0

"""C-ABI library checks that need no GPU: it loads, it exports every symbol
declared in include/*.h, and its host-side entry points behave."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT, gpu_available

import paraprobe_amd.capi as capi


def _declared_functions():
    names = []
    inc = os.path.join(ROOT, "include")
    for f in sorted(os.listdir(inc)):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(inc, f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(pp[h]?_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = capi.load_library()
    declared = [n for n in _declared_functions() if n.startswith("pp_")]
    assert declared, "no declarations found"
    for name in declared:
        assert hasattr(lib, name), name
    assert set(capi.EXPORTED) == set(declared)


def test_hist_geometry_matches_reference_quirks():
    assert capi.hist_geometry(0.0, 0.1, 2.0) == 19
    assert capi.hist_geometry(0.0, 0.01, 5.0) == 500
    assert capi.hist_geometry(0.0, 0.005, 3.0) == 600
    with pytest.raises(capi.PPError):
        capi.hist_geometry(0.0, 1e-7, 2.0)


@pytest.mark.skipif(gpu_available(), reason="checks the no-device path")
def test_create_without_device_fails_cleanly():
    lib = capi.load_library()
    h = C.c_void_p()
    st = lib.pp_spatstat_create(0, C.byref(h))
    assert st == 5  # PP_ERR_NO_DEVICE
    assert not h.value


def test_type_mask():
    m = capi.type_mask([0, 1, 9, 255])
    assert m[0] == 0b11 and m[1] == 0b10 and m[31] == 0x80


def test_host_library_exports_every_declared_pph_symbol():
    from paraprobe_amd import horderdist as H

    lib = H.load()
    declared = [n for n in _declared_functions() if n.startswith("pph_")]
    assert declared, "no pph declarations found"
    for name in declared:
        assert hasattr(lib, name), name
    assert set(H.EXPORTED) == set(declared)


def test_hist_supported_refuses_bins_beyond_the_near_edge_bound():
    lib = capi.load_library()
    # the measured geometries
    for g in ((0.0, 0.01, 5.0), (0.0, 0.005, 3.0), (0.0, 0.1, 2.0), (4.8, 0.002, 5.0)):
        assert lib.pp_hist_supported(*g) == capi.PP_OK, g
    # over a million bins of 4.9e-6 nm up to 5 nm: the edges' offset from their
    # ideal index (0.127) plus the float index coordinate's rounding bound
    # (0.367) stays below half a bin, so the near-edge path is exact
    assert lib.pp_hist_supported(0.0, 4.9e-6, 5.0) == capi.PP_OK
    # 4.8e-6 passes the table build and the 2^20 index bound, but 0.130 +
    # 0.375 reaches half a bin: refused (the margin would otherwise be capped)
    assert lib.pp_hist_supported(0.0, 4.8e-6, 5.0) == 6  # PP_ERR_UNSUPPORTED
    # what Settings::checkUserInput rejects
    assert lib.pp_hist_supported(0.0, 1e-7, 5.0) == 1     # PP_ERR_INVALID_ARG
    assert lib.pp_hist_supported(2.0, 0.1, 1.0) == 1

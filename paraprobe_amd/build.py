"""Builds the native libraries of paraprobe_amd in-tree.

* ``libpp_spatstat.so`` -- the gfx950 product library (HIP kernels + C ABI of
  ``include/pp_spatstat.h``), compiled with hipcc for gfx950 only.
* ``libpp_synth.so`` -- the seeded synthetic-tip generator (host C++).
* ``libpp_horderdist.so`` -- the C++ host mirror of the reference's
  ``horderdist`` task interface (task parsing, settings, CSV writers) that
  drives the product library through its C ABI.

All device code is built with ``-ffp-contract=off`` (the reference's FP32
distance expression must not be contracted into FMAs) and with packed FP32
instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    "--offload-arch=gfx950",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
]

LIB_SPATSTAT = os.path.join(PKG, "libpp_spatstat.so")
LIB_SYNTH = os.path.join(PKG, "libpp_synth.so")
LIB_HORDERDIST = os.path.join(PKG, "libpp_horderdist.so")
LIB_H5 = os.path.join(PKG, "libpp_h5.so")
# the HDF5 C library of this image; the HDF5 writer is built only where it exists
HDF5_PREFIX = os.environ.get("PP_HDF5_PREFIX", "/opt/conda")


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed ({r.returncode}): {cmd[0]}")
    # surface warnings that are not the known harmless host-side feature note
    msg = "\n".join(
        ln for ln in (r.stdout + r.stderr).splitlines()
        if ln.strip() and "packed-fp32-ops" not in ln and "not a recognized feature" not in ln
    )
    if msg:
        sys.stderr.write(msg + "\n")


def build_spatstat() -> str:
    objs = []
    tmp = os.path.join(PKG, "_build")
    os.makedirs(tmp, exist_ok=True)
    for src in ("pp_kernels.hip", "pp_surfdist.hip", "pp_io.hip", "pp_npcorr.hip",
                "pp_noalloc.hip", "pp_prep.hip", "pp_capi.cpp", "pp_binning.cpp"):
        obj = os.path.join(tmp, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([HIPCC, *HIP_FLAGS, *lang, "-I", INCLUDE, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    out_tmp = LIB_SPATSTAT + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", out_tmp])
    os.replace(out_tmp, LIB_SPATSTAT)
    return LIB_SPATSTAT


def build_synth() -> str:
    out_tmp = LIB_SYNTH + ".tmp"
    _run([CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp",
          os.path.join(CSRC, "pp_synth.cpp"), "-o", out_tmp])
    os.replace(out_tmp, LIB_SYNTH)
    return LIB_SYNTH


def build_horderdist() -> str:
    srcs = [os.path.join(HOST, f) for f in sorted(os.listdir(HOST)) if f.endswith(".cpp")]
    if not srcs:
        return ""
    out_tmp = LIB_HORDERDIST + ".tmp"
    _run([CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-Wall", "-I", INCLUDE,
          *srcs, "-o", out_tmp, "-L", PKG, "-lpp_spatstat", "-Wl,-rpath,$ORIGIN"])
    os.replace(out_tmp, LIB_HORDERDIST)
    return LIB_HORDERDIST


def build_h5() -> str:
    """libpp_h5.so: the optional HDF5 writer of the /DescrSpatStats datasets
    (host C++ against the HDF5 C library; nothing else links it)."""
    inc = os.path.join(HDF5_PREFIX, "include")
    lib = os.path.join(HDF5_PREFIX, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        return ""
    out_tmp = LIB_H5 + ".tmp"
    _run([CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", INCLUDE, "-I", inc,
          os.path.join(HOST, "h5", "pp_h5.cpp"), "-o", out_tmp, "-L", lib, "-lhdf5",
          "-Wl,-rpath," + lib])
    os.replace(out_tmp, LIB_H5)
    return LIB_H5


def build_all() -> list[str]:
    out = [build_synth(), build_spatstat()]
    h = build_horderdist()
    if h:
        out.append(h)
    h5 = build_h5()
    if h5:
        out.append(h5)
    shutil.rmtree(os.path.join(PKG, "_build"), ignore_errors=True)
    return out


if __name__ == "__main__":
    for p in build_all():
        print(p)

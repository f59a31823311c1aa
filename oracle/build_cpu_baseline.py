"""Builds oracle/libcpu_baseline.so: the reference's CPU algorithm restated
for timing (g++ -O3 -fopenmp -ffp-contract=off; x86-64-v3 so the library
built here runs on any current x86 host)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libcpu_baseline.so")


def build() -> str:
    tmp = LIB + ".tmp"
    cmd = ["g++", "-O3", "-std=c++17", "-march=x86-64-v3", "-fPIC", "-shared", "-fopenmp",
           "-ffp-contract=off", os.path.join(HERE, "pp_cpu_baseline.cpp"), "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("cpu baseline build failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())

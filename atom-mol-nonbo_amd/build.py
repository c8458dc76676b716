"""Build libecg_hip.so (the HIP kernels and the C-ABI) in-tree for gfx950,
and the host-only libecg_drmng.so (the minimiser of the optimisers).

Usage: python atom-mol-nonbo_amd/build.py   (or __graft_entry__.build())
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libecg_hip.so")
SOURCES = ["ecg_kernels.hip", "ecg_linalg.hip", "ecg_screen.hip", "ecg_capi.hip"]
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-pthread",
    # packed FP32 instructions stay off (the path is FP64 throughout)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, src.replace(".hip", ".o"))
        cmd = [hipcc, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln]
        if p.returncode != 0:
            failed = True
            sys.stderr.write("FAILED: %s\n%s\n" % (" ".join(cmd), "\n".join(lines)))
        elif verbose and lines:
            sys.stderr.write("\n".join(lines) + "\n")
    if failed:
        raise RuntimeError("hipcc failed")
    tmp = LIB + ".tmp"
    subprocess.run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs], check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    build_drmng()
    build_fortran()
    return LIB


DRMNG_LIB = os.path.join(HERE, "libecg_drmng.so")


def build_drmng():
    """Compile the host-only minimiser library (csrc/drmng.cpp, the C-ABI of
    include/ecg_drmng.h).  Products are not contracted into fused
    multiply-adds: the arithmetic follows the reference's operation order."""
    cxx = os.environ.get("CXX", "g++")
    tmp = DRMNG_LIB + ".tmp"
    subprocess.run([cxx, "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-shared", "-o", tmp,
                    os.path.join(CSRC, "drmng.cpp")], check=True)
    os.replace(tmp, DRMNG_LIB)
    return DRMNG_LIB


FLANG = "/opt/rocm/lib/llvm/bin/flang"
FDIR = os.path.join(HERE, "fortran")
FBUILD = os.path.join(FDIR, "_build")
DRIVER_SRC = os.path.join(os.path.dirname(HERE), "tests", "fortran", "drive_fixture.f90")
DRIVER = os.path.join(os.path.dirname(HERE), "tests", "fortran", "drive_fixture")


def build_fortran():
    """Compile the Fortran host layer (fortran/ecg_gpu.f90) and the test
    driver that links it against libecg_hip.so.  Skipped (returns None)
    when no Fortran compiler is installed."""
    fc = os.environ.get("FC", FLANG)
    if not os.path.exists(fc):
        return None
    os.makedirs(FBUILD, exist_ok=True)
    obj = os.path.join(FBUILD, "ecg_gpu.o")
    subprocess.run([fc, "-O2", "-c", os.path.join(FDIR, "ecg_gpu.f90"), "-module-dir", FBUILD, "-o", obj],
                   check=True)
    if os.path.exists(DRIVER_SRC):
        subprocess.run([fc, "-O2", "-I", FBUILD, DRIVER_SRC, obj, "-L" + HERE, "-lecg_hip",
                        "-Wl,-rpath,$ORIGIN/../../atom-mol-nonbo_amd", "-o", DRIVER], check=True)
    return obj


if __name__ == "__main__":
    print(build(verbose=True))

"""Build recipe of the native parts (no cmake): the gfx950 engine + host
library (libpansim.so), the panSim CLI, the pybind11 module pyPanSim, and the
CPU oracle used by the tests (oracle/_build/liboracle.so).

Compilers: hipcc for .hip (gfx950 code objects), g++ for plain C++.  Every
translation unit is compiled with -ffp-contract=off so that host and device
evaluate floating point identically (no fused multiply-add), and HIP code with
packed-fp32 instruction selection disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sysconfig
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "pansim_amd")
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "_lib")
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_OUT = os.path.join(ORACLE_DIR, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

HOST_SOURCES = ["json.cpp", "scenario.cpp", "synthpop.cpp", "model.cpp", "sim_capi.cpp"]
INPUT_SOURCES = ["json.cpp", "scenario.cpp", "synthpop.cpp"]

CXXFLAGS = ["-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unknown-pragmas"]
HIPFLAGS = [
    "-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    return r


def build_engine(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    eng_obj = os.path.join(objdir, "engine.o")
    _run([HIPCC, *HIPFLAGS, "-c", os.path.join(CSRC, "device", "engine.hip"), "-o", eng_obj], verbose)
    objs.append(eng_obj)
    for src in HOST_SOURCES:
        o = os.path.join(objdir, src.replace(".cpp", ".o"))
        _run([CXX, *CXXFLAGS, "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", "-c",
              os.path.join(CSRC, "host", src), "-o", o], verbose)
        objs.append(o)
    lib = os.path.join(LIBDIR, "libpansim.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs], verbose)
    os.replace(tmp, lib)
    main_obj = os.path.join(objdir, "main.o")
    _run([CXX, *CXXFLAGS, "-c", os.path.join(CSRC, "host", "main.cpp"), "-o", main_obj], verbose)
    exe = os.path.join(LIBDIR, "panSim")
    _run([CXX, "-o", exe, main_obj, f"-L{LIBDIR}", "-lpansim", "-Wl,-rpath,$ORIGIN"], verbose)
    build_pypansim(verbose)
    shutil.rmtree(objdir, ignore_errors=True)
    return lib


def build_pypansim(verbose=False):
    """The pybind11 module pyPanSim (the reference's Python binding) over
    libpansim.so, next to it in _lib/."""
    import pybind11
    out = os.path.join(LIBDIR, "pyPanSim" + sysconfig.get_config_var("EXT_SUFFIX"))
    tmp = out + ".tmp"
    _run([CXX, *CXXFLAGS, "-fvisibility=hidden", "-shared", f"-I{pybind11.get_include()}",
          f"-I{sysconfig.get_paths()['include']}", os.path.join(CSRC, "host", "pypansim.cpp"), "-o", tmp,
          f"-L{LIBDIR}", "-lpansim", "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(tmp, out)
    return out


def build_oracle(verbose=False):
    """The oracle twice: liboracle.so (sequential, the checker of the tests)
    and liboracle_omp.so (the same code with the reference's OpenMP structure:
    parallel movement, infection and progression, a per-step stable sort by
    location and per-location sums; the CPU baseline of bench.py)."""
    os.makedirs(ORACLE_OUT, exist_ok=True)
    srcs = [os.path.join(ORACLE_DIR, "oracle.cpp")] + [os.path.join(CSRC, "host", s) for s in INPUT_SOURCES]
    for name, extra in (("liboracle.so", []), ("liboracle_omp.so", ["-O3", "-fopenmp", "-DORACLE_OMP"])):
        lib = os.path.join(ORACLE_OUT, name)
        tmp = lib + ".tmp"
        _run([CXX, *CXXFLAGS, *extra, "-shared", "-o", tmp, *srcs], verbose)
        os.replace(tmp, lib)
    return os.path.join(ORACLE_OUT, "liboracle.so")


def build_all(verbose=False):
    build_oracle(verbose)
    return build_engine(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", LIBDIR, ORACLE_OUT)

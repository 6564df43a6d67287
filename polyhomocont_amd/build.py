"""Builds the native parts in-tree.

  lib/libphc_hip.so         HIP kernels + C ABI (include/phc.h), hipcc for
                            gfx950
  lib/libpolyhomocont_amd.so  Fortran host library (module polyhomocont and
                            the C-callable layer of include/polyhomocont.h),
                            flang, linked against libphc_hip.so
  lib/mod/*.mod             Fortran module files for Fortran callers
  lib/libphc_wire.so        request/endpoint frames of a serving front end
                            (include/phc_wire.h), host C++ built with g++

Every call compiles from source (no up-to-date checks), so a build is always
the current tree.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")
LIB = os.path.join(HERE, "lib")
BUILD = os.path.join(HERE, "_build")

ARCH = os.environ.get("PHC_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # The kernels reproduce the reference's floating-point operations one by
    # one; contraction into FMA would change results.
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]

# Fortran modules in dependency order.
FORTRAN_SOURCES = [
    "config.f90", "random.f90", "linalg.f90", "system.f90", "sparse.f90",
    "homotopy.f90", "tracker.f90", "endgame.f90", "device.f90", "solve.f90",
    "polyhomocont.f90", "capi.f90",
]


def _run(cmd, cwd=None, timeout=1200):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, timeout=timeout, capture_output=True,
                       text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed ({r.returncode}): {cmd[0]}")
    return r


def build_hip():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libphc_hip.so")
    tmp = out + ".tmp"
    _run(["hipcc", *HIP_FLAGS, "-shared",
          os.path.join(CSRC, "phc_runtime.hip"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_fortran():
    """Fortran host library; skipped (returns None) when the sources are not
    present yet."""
    srcs = [os.path.join(FSRC, f) for f in FORTRAN_SOURCES]
    if not all(os.path.exists(s) for s in srcs):
        return None
    mod = os.path.join(LIB, "mod")
    shutil.rmtree(BUILD, ignore_errors=True)
    os.makedirs(BUILD)
    os.makedirs(mod, exist_ok=True)
    objs = []
    for s in srcs:
        o = os.path.join(BUILD, os.path.basename(s)[:-4] + ".o")
        _run(["flang", "-O2", "-fPIC", "-fopenmp", "-module-dir", mod, "-I",
              mod, "-c", s, "-o", o])
        objs.append(o)
    out = os.path.join(LIB, "libpolyhomocont_amd.so")
    tmp = out + ".tmp"
    _run(["flang", "-shared", "-fPIC", "-fopenmp", *objs, "-o", tmp,
          "-L" + LIB, "-lphc_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    shutil.rmtree(BUILD, ignore_errors=True)
    return out


def build_wire():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libphc_wire.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-fPIC", "-std=c++17", "-Wall", "-Wextra", "-shared",
          os.path.join(CSRC, "phc_wire.cpp"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_all():
    """The product: HIP kernels, frame codec, Fortran host library."""
    build_hip()
    build_wire()
    build_fortran()


if __name__ == "__main__":
    build_all()

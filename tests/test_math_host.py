"""The device complex arithmetic (polyhomocont_amd/csrc/phc_math.h) compiled
for the CPU and compared bit for bit with the routines the reference's flang
build calls: compiler-rt __divdc3, glibc hypot, the flang runtime's
complex ** integer, and glibc's sincos (the call flang -O2 makes for the
cos/sin pair of an arc angle).

The kernel includes the same header, so a difference here is a difference
on the GPU. Needs g++ and the ROCm LLVM runtime
archives; skips when one of them is absent.
"""
import glob
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))


def _find(pattern):
    hits = sorted(glob.glob(pattern))
    return hits[-1] if hits else None


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gxx = shutil.which("g++")
    builtins = _find("/opt/rocm*/lib/llvm/lib/clang/*/lib/linux/libclang_rt.builtins-x86_64.a")
    flang_rt = _find("/opt/rocm*/lib/llvm/lib/clang/*/lib/x86_64-unknown-linux-gnu/libflang_rt.runtime.a")
    if not (gxx and builtins and flang_rt):
        pytest.skip("g++ or the LLVM runtime archives are not available")
    exe = str(tmp_path_factory.mktemp("mathchk") / "math_host_check")
    cmd = [gxx, "-O2", "-std=c++17", "-ffp-contract=off",
           "-I", os.path.join(ROOT, "polyhomocont_amd", "csrc"),
           os.path.join(HERE, "math_host_check.cpp"), "-o", exe,
           flang_rt, builtins, "-lm", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail("host build of phc_math.h failed:\n" + r.stderr[-3000:])
    out = subprocess.run([exe, "1000000"], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr
    res = {}
    for line in out.stdout.splitlines():
        name, checked, bad = line.split()
        res[name] = (int(checked), int(bad))
    return res


@pytest.mark.parametrize("name", ["cdiv", "cabs", "zpowi", "zpowi_pair",
                                  "libm_sincos"])
def test_bit_exact_against_host_routine(checker, name):
    checked, bad = checker[name]
    assert checked > 100000
    assert bad == 0, f"{name}: {bad} of {checked} differ"

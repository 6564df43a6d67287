"""Drop-in check: the reference's own test programs (test/*.f90 with its
`testing` module), compiled unmodified against this library's Fortran
modules and shared library by `make -C oracle droptests`, must pass. Each
program prints ok/FAIL lines and exits non-zero on any failure
(test/testing.f90:116-122). solve, katsura and cyclic call `solve` on
sparse systems, i.e. the GPU; the others exercise the host layer (rng,
linear algebra, builders, homotopy, CPU tracker and end game)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DT = os.path.join(ROOT, "oracle", "_ref", "droptest")


def run(name):
    exe = os.path.join(DT, name)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built (make -C oracle droptests)")
    print(f"[droptest] {name}", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    fails = [l for l in r.stdout.splitlines() if l.startswith("FAIL")]
    assert r.returncode == 0 and not fails, r.stdout[-3000:] + r.stderr[-2000:]
    return r.stdout


@pytest.mark.parametrize("name", ["random", "linalg", "system", "sparse",
                                  "homotopy", "tracker", "endgame"])
def test_reference_host_tests(name):
    out = run(name)
    assert out.count("ok") > 5


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["solve", "katsura", "cyclic"])
def test_reference_solve_tests_on_gpu(name):
    run(name)

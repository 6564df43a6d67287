"""Mutation fuzzing of the C frame decoders under AddressSanitizer and
UndefinedBehaviorSanitizer (host code; tests/wire_fuzz.cpp): no mutated
frame may make a decoder read outside the frame or hit undefined
behaviour, and every unmutated frame decodes. Skips when g++ cannot build
with the sanitizers."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def fuzzer(tmp_path_factory):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("wirefuzz") / "wire_fuzz")
    cmd = [gxx, "-O1", "-g", "-std=c++17", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "polyhomocont_amd", "csrc", "phc_wire.cpp"),
           os.path.join(HERE, "wire_fuzz.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("sanitizer build failed: " + r.stderr[-400:])
    return exe


def test_mutated_frames_never_read_outside(fuzzer):
    env = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([fuzzer, "30000"], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    ok, bad = [int(x.rstrip(",")) for x in r.stdout.split()[1::2]]
    assert ok >= 8 and bad > 1000

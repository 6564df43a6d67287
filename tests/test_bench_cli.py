"""bench.py's contract on the CPU: more than one GPU is refused with a
message, and the CPU baselines report what they measured."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"),
                        "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_cpu_baselines_on_c1():
    import bench
    import oracle
    from polyhomocont_amd import abi, workloads
    if oracle.ref() is None and oracle.port() is None:
        pytest.skip("no CPU tracker built")
    usys = workloads.config_system("c1")
    one, rec, cnt = bench.cpu_baseline(usys, abi.default_options(), 16, 0.5)
    assert one["cores"] == 1 and one["value"] > 0 and 1 <= cnt <= 16
    assert one["kind"] in ("reference", "port")
    allc = bench.cpu_all_cores("c1", 16, 1.0 / one["value"], 0.5)
    assert allc["cores"] >= 1 and allc["value"] > 0
    assert allc["kind"] == one["kind"]


def test_percentiles():
    import bench
    p = bench.percentiles([0.001 * k for k in range(1, 101)])
    assert p["p50"] == pytest.approx(50.5)
    assert p["p99"] == pytest.approx(99.01)

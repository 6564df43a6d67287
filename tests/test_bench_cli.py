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


def test_cpu_baselines_on_a_short_stream(host):
    """The CPU baselines of the serving stream report what they measured
    and compare the reference's results with those given."""
    import bench
    import oracle
    from polyhomocont_amd import abi, workloads
    if oracle.ref() is None:
        pytest.skip("the compiled reference is not built")
    stream = workloads.serving_stream(4)
    opts = abi.default_options()
    host.set_device(-1)          # the host tracker stands in for the GPU
    try:
        results = [host.solve(u, opts)[1] for u in stream]
    finally:
        host.set_device(0)
    one, par = bench.cpu_stream_baseline(stream, results, 0.5)
    assert one["cores"] == 1 and one["value"] > 0
    assert one["kind"] == "reference"
    assert par["requests_compared"] >= 1
    assert par["counts_and_status_equal"] == par["requests_compared"]
    assert par["max_rel_solution_err"] <= 1e-10
    allc = bench.cpu_stream_all_cores(4, 1, 0.5)
    assert allc["cores"] >= 1 and allc["value"] > 0


def test_percentiles():
    import bench
    p = bench.percentiles([0.001 * k for k in range(1, 101)])
    assert p["p50"] == pytest.approx(50.5)
    assert p["p99"] == pytest.approx(99.01)

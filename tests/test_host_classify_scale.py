"""Host post-processing at scale (classification, refinement and
deduplication in path order, src/solve.f90:193-213, 254-426) on the CPU.

The product's classification (threaded per path, deduplication through the
key window) of the reference's own endpoint records must give the
reference `solve`'s counts, multiplicities and per-path solution and
singular indices: for every path of C3 and C3v (5040 each, from the
fixture) and for some of the batched systems (records from the compiled
reference, fixture counts).
"""
import os

import numpy as np
import pytest

from polyhomocont_amd import abi, workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
CLASSIFIED = ("solution_index", "singular_index", "at_infinity", "singular")


def load(name):
    p = os.path.join(GOLD, name)
    if not os.path.exists(p):
        pytest.skip(f"{name} not generated")
    return np.load(p)


def check(res, keys, counts, mult, per_path, npaths):
    c = res.counts()
    assert [c[k] for k in keys] == list(counts)
    assert list(res.multiplicities[:res.summary.nsingular]) == list(mult)
    for f, want in per_path.items():
        assert np.array_equal(res.p[f][:npaths], want), f


@pytest.mark.parametrize("name", ["c3", "c3v"])
@pytest.mark.parametrize("threads", [1, 8])
def test_c3_classification_of_reference_records(host, name, threads):
    F = load("c3_full.npz")
    keys = [str(k) for k in F["count_keys"]]
    usys = W.config_system(name)
    n = 5040
    rec = abi.EndpointArrays(n, usys.n + 1)
    for f in ("status", "winding", "eg_status", "naccepted", "nrejected",
              "t", "eg_error", "z"):
        getattr(rec, f)[:n] = F[f"{name}/{f}"]
    prev = host.set_threads(threads)
    try:
        st, res = host.classify(usys, abi.default_options(), rec)
    finally:
        host.set_threads(prev)
    assert st == 0
    check(res, keys, F[f"{name}/counts"], F[f"{name}/multiplicities"],
          {f: F[f"{name}/{f}"] for f in CLASSIFIED}, n)


def test_batch_classification_of_reference_records(host, ref):
    F = load("c4_batch.npz")
    keys = [str(k) for k in F["count_keys"]]
    off, moff = F["offsets"], F["mult_offsets"]
    npaths = F["npaths"]
    # a few systems of 300..3000 paths, one with singular endpoints if any
    pick = [s for s in range(512) if 300 <= npaths[s] <= 3000][:3]
    sing = [s for s in range(512) if F["counts"][s][3] > 0 and npaths[s] <= 3000]
    pick += sing[:1]
    opts = abi.default_options()
    for s in pick:
        usys = W.random_sparse(104 + s)
        rec = ref.endpoints(usys, opts)
        lo, hi = off[s], off[s + 1]
        assert np.array_equal(rec.status[:hi - lo], F["status"][lo:hi])
        st, res = host.classify(usys, opts, rec)
        assert st == 0
        check(res, keys, F["counts"][s], F["multiplicities"][moff[s]:moff[s + 1]],
              {f: F[f][lo:hi] for f in CLASSIFIED}, hi - lo)


@pytest.mark.parametrize("threads", [1, 8])
def test_c2_classification_of_reference_records(host, threads):
    """All 2187 C2 endpoint records of the reference, classified by the
    host library, give the reference `solve`'s counts and per-path indices
    (2187 finite regular solutions, every path its own solution)."""
    F = load("c2_full.npz")
    keys = [str(k) for k in F["count_keys"]]
    usys = W.config_system("c2")
    n = 2187
    rec = abi.EndpointArrays(n, usys.n + 1)
    for f in ("status", "winding", "eg_status", "naccepted", "nrejected",
              "t", "eg_error", "z"):
        getattr(rec, f)[:n] = F[f]
    prev = host.set_threads(threads)
    try:
        st, res = host.classify(usys, abi.default_options(), rec)
    finally:
        host.set_threads(prev)
    assert st == 0
    check(res, keys, F["counts"], F["multiplicities"],
          {f: F[f] for f in CLASSIFIED}, n)
    assert res.summary.nsolutions == 2187

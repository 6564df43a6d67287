"""Full-scale parity on the GPU against the reference's own records.

The fixtures (tests/golden/make_golden.py) hold, for every path of C3 and
its squared-factor variant C3v (5040 paths each), the compiled reference's
endpoint records and the reference `solve`'s classification; for all 512
systems of the batched workload (743,290 paths), the discrete fields of
every path, a SHA-256 digest of the endpoint doubles per system and the
reference `solve`'s counts and per-path classification.

The GPU records must be the reference's bit for bit; the host
classification of the GPU records must give the reference's counts
nsolutions, nreal, nsingular, ninfinite, nfailed, nduplicates, its
multiplicities and every path's solution and singular index
(src/solve.f90:193-213).
"""
import os

import numpy as np
import pytest

from polyhomocont_amd import abi, workloads as W

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
DISCRETE = ("status", "winding", "eg_status", "naccepted", "nrejected")
CLASSIFIED = ("solution_index", "singular_index", "at_infinity", "singular")


def load(name):
    p = os.path.join(GOLD, name)
    if not os.path.exists(p):
        pytest.skip(f"{name} not generated")
    return np.load(p)


def counts_vector(res, keys):
    c = res.counts()
    return np.array([c[k] for k in keys], np.int64)


@pytest.mark.parametrize("name", ["c3", "c3v"])
def test_c3_every_path_and_solve_counts(ctx, host, name):
    F = load("c3_full.npz")
    keys = [str(k) for k in F["count_keys"]]
    usys = W.config_system(name)
    opts = abi.default_options()
    g, b, a, ns = host.homotopy(usys, opts.seed)
    assert ns == 5040
    start = host.start_points(usys, opts.seed)
    print(f"[gpu] {name}: all {ns} paths", flush=True)
    got = ctx.track_paths(host.pack(usys), abi.HomotopyData(g, b, a), opts,
                          start)
    for f in DISCRETE:
        want = F[f"{name}/{f}"]
        bad = np.nonzero(getattr(got, f)[:ns] != want)[0]
        assert bad.size == 0, (f, bad[:10])
    for f in ("t", "eg_error"):
        assert np.array_equal(getattr(got, f)[:ns], F[f"{name}/{f}"]), f
    assert np.array_equal(got.z[:ns], F[f"{name}/z"])
    st, res = host.classify(usys, opts, got)
    assert st == 0
    assert list(counts_vector(res, keys)) == list(F[f"{name}/counts"])
    n_s = res.summary.nsingular
    assert list(res.multiplicities[:n_s]) == list(F[f"{name}/multiplicities"])
    for f in CLASSIFIED:
        assert np.array_equal(res.p[f][:ns], F[f"{name}/{f}"]), f


def test_c4_batch_every_system(ctx, host):
    F = load("c4_batch.npz")
    keys = [str(k) for k in F["count_keys"]]
    from tests.golden.make_golden import endpoint_digest, weighted_sum
    systems = W.batch_systems(512)
    opts = abi.default_options()
    items = []
    for usys in systems:
        g, b, a, ns = host.homotopy(usys, opts.seed)
        items.append((host.pack(usys), abi.HomotopyData(g, b, a), opts,
                      host.start_points(usys, opts.seed)))
    off = F["offsets"]
    assert off[-1] == sum(u.bezout for u in systems) == 743290
    print(f"[gpu] batch of 512 systems, {off[-1]} paths", flush=True)
    outs = ctx.track_batch(items)
    bad_fields, bad_digest, bad_sum, bad_class = [], [], [], []
    for s, (usys, got) in enumerate(zip(systems, outs)):
        ns = usys.bezout
        lo, hi = off[s], off[s + 1]
        assert hi - lo == ns
        for f in DISCRETE:
            if not np.array_equal(getattr(got, f)[:ns], F[f][lo:hi]):
                bad_fields.append((s, f))
        if not np.array_equal(endpoint_digest(got.z[:ns], got.t[:ns],
                                              got.eg_error[:ns]),
                              F["digest"][s]):
            bad_digest.append(s)
        wsum, asum = weighted_sum(s, got.z[:ns])
        if not abs(wsum - F["wsum"][s]) <= 1e-10 * max(F["asum"][s], 1.0):
            bad_sum.append(s)
        st, res = host.classify(usys, opts, got)
        ok = st == 0 and list(counts_vector(res, keys)) == list(F["counts"][s])
        m0, m1 = F["mult_offsets"][s], F["mult_offsets"][s + 1]
        ok = ok and list(res.multiplicities[:res.summary.nsingular]) == \
            list(F["multiplicities"][m0:m1])
        for f in CLASSIFIED:
            ok = ok and np.array_equal(res.p[f][:ns], F[f][lo:hi])
        if not ok:
            bad_class.append(s)
    print(f"[gpu] batch: {len(bad_fields)} field, {len(bad_digest)} digest, "
          f"{len(bad_sum)} sum, {len(bad_class)} classification mismatches",
          flush=True)
    assert not bad_fields, bad_fields[:10]
    assert not bad_sum, bad_sum[:10]
    assert not bad_digest, bad_digest[:10]
    assert not bad_class, bad_class[:10]

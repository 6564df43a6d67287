"""Full-scale parity on the GPU against the reference's own records.

The fixtures (tests/golden/make_golden.py) hold, for every path of C2
(2187 paths), C3 and its squared-factor variant C3v (5040 paths each), the
compiled reference's endpoint records and the reference `solve`'s
classification; for all 512 systems of the batched workload (743,290
paths), the discrete fields of every path, a SHA-256 digest of the endpoint
doubles per system and the reference `solve`'s counts and per-path
classification; for the 512 requests of the serving stream (239,777
paths), the reference `solve`'s whole result per request (discrete fields
and classification per path, counts, and a digest of the result's doubles).

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


def test_c2_every_path_and_solve_counts(ctx, host):
    """All 2187 paths of C2 on the plain tracker: every record the
    reference's bit for bit, and the reference `solve`'s classification."""
    F = load("c2_full.npz")
    keys = [str(k) for k in F["count_keys"]]
    usys = W.config_system("c2")
    opts = abi.default_options()
    g, b, a, ns = host.homotopy(usys, opts.seed)
    assert ns == 2187
    start = host.start_points(usys, opts.seed)
    print(f"[gpu] c2: all {ns} paths", flush=True)
    got = ctx.track_paths(host.pack(usys), abi.HomotopyData(g, b, a), opts,
                          start)
    for f in DISCRETE:
        bad = np.nonzero(getattr(got, f)[:ns] != F[f])[0]
        assert bad.size == 0, (f, bad[:10])
    for f in ("t", "eg_error"):
        assert np.array_equal(getattr(got, f)[:ns], F[f]), f
    assert np.array_equal(got.z[:ns], F["z"])
    st, res = host.classify(usys, opts, got)
    assert st == 0
    assert list(counts_vector(res, keys)) == list(F["counts"])
    assert list(res.multiplicities[:res.summary.nsingular]) == \
        list(F["multiplicities"])
    for f in CLASSIFIED:
        assert np.array_equal(res.p[f][:ns], F[f]), f


def test_serving_stream_every_request_on_the_engine(ctx, host):
    """The 512 requests of the serving stream (239,777 paths) through the
    continuous-batching engine, all in flight together, then classified on
    the host: every request's result is the reference `solve`'s -- counts,
    multiplicities, every path's status, winding, end game status, step
    counts and classification, and the result's doubles (digest of the
    final z, t and end game errors)."""
    from polyhomocont_amd import device
    from tests.golden.make_golden import SERVING_REQUESTS, endpoint_digest
    F = load("serving_stream.npz")
    keys = [str(k) for k in F["count_keys"]]
    stream = W.serving_stream(SERVING_REQUESTS)
    off = F["offsets"]
    assert off[-1] == sum(u.bezout for u in stream) == 239777
    opts = abi.default_options()
    srv = device.Server(ctx, max_resident_paths=16384,
                        hbm_budget_bytes=2 << 30, units_per_launch=64)
    print(f"[gpu] serving stream: {len(stream)} requests, {off[-1]} paths",
          flush=True)
    try:
        tickets = []
        for usys in stream:
            g, b, a, ns = host.homotopy(usys, opts.seed)
            tickets.append(srv.submit(host.pack(usys),
                                      abi.HomotopyData(g, b, a), opts,
                                      host.start_points(usys, opts.seed)))
        pending = {t for t, _ in tickets}
        while pending:
            if srv.progress() > 0:
                pending = {t for t in pending if not srv.done(t)}
        for t, _ in tickets:
            srv.release(t)
    finally:
        srv.close()
    bad = []
    for i, (usys, (t, rec)) in enumerate(zip(stream, tickets)):
        st, res = host.classify(usys, opts, rec)
        lo, hi = off[i], off[i + 1]
        n = hi - lo
        ok = st == 0 and res.summary.npaths == n
        ok = ok and list(counts_vector(res, keys)) == list(F["counts"][i])
        m0, m1 = F["mult_offsets"][i], F["mult_offsets"][i + 1]
        ok = ok and list(res.multiplicities[:res.summary.nsingular]) == \
            list(F["multiplicities"][m0:m1])
        for f in DISCRETE + CLASSIFIED:
            ok = ok and np.array_equal(res.p[f][:n], F[f][lo:hi])
        ok = ok and np.array_equal(
            endpoint_digest(res.p["z"][:n], res.p["t"][:n],
                            res.p["eg_error"][:n]), F["digest"][i])
        if not ok:
            bad.append(i)
    print(f"[gpu] serving stream: {len(bad)} of {len(stream)} requests "
          "differ from the reference", flush=True)
    assert not bad, bad[:10]

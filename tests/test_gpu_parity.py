"""GPU tracker vs the oracle, path by path, through the C ABI.

The device reproduces the reference's arithmetic operation by operation, so
the endpoint records are bit-identical: same status, winding number, end
game status, accepted/rejected step counts and the same endpoint doubles.
That includes the hard cases (C3: endpoints at infinity with winding
numbers up to 14, whose end game loops evaluate arcs at non-dyadic
angles), where a one-bit difference in an arc's cos or sin would change
the step sequence.
"""
import os

import numpy as np
import pytest

from polyhomocont_amd import abi, workloads as W

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIELDS = ("status", "winding", "eg_status", "naccepted", "nrejected")


def golden():
    return np.load(os.path.join(HERE, "golden", "endpoints.npz"))


def run_device(ctx, usys, opts, consts, start):
    g, b, a = consts
    return ctx.track_paths(usys.packed(), abi.HomotopyData(g, b, a), opts, start)


def assert_bit_exact(got, want, n):
    for f in FIELDS:
        assert np.array_equal(getattr(got, f)[:n], getattr(want, f)[:n]), f
    assert np.array_equal(got.z[:n], want.z[:n])
    assert np.array_equal(got.eg_error[:n], want.eg_error[:n])
    assert np.array_equal(got.t[:n], want.t[:n])


@pytest.mark.parametrize("name", ["cubic", "2x2", "infinity", "double_root",
                                  "triple_root", "mixed", "mult4", "user_cf",
                                  "univariate5", "katsura3", "katsura4",
                                  "cyclic5", "c1"])
def test_reference_systems_bit_exact_vs_golden(ctx, name):
    """Against the committed fixtures of the compiled reference."""
    G = golden()
    usys = W.config_system(name) if name == "c1" else W.ref_test_systems()[name]
    opts = abi.default_options()
    consts = (G[name + "/gamma"][0], G[name + "/b"], G[name + "/patch"])
    from polyhomocont_amd import host
    start = host.library().start_points(usys, opts.seed)
    print(f"[gpu] {name}: {len(start)} paths", flush=True)
    got = run_device(ctx, usys, opts, consts, start)
    n = len(start)
    for f in FIELDS + ("eg_error", "t"):
        assert np.array_equal(getattr(got, f)[:n], G[f"{name}/{f}"]), f
    assert np.array_equal(got.z[:n], G[name + "/z"])


def test_c2_sample_bit_exact(ctx, ref):
    usys = W.config_system("c2")
    opts = abi.default_options()
    g, b, a, ns = ref.homotopy(usys, opts.seed)
    cnt = 192
    start = ref.start_points(usys, opts.seed, 0, cnt)
    print("[gpu] c2 sample", cnt, flush=True)
    got = run_device(ctx, usys, opts, (g, b, a), start)
    want = ref.endpoints(usys, opts, 0, cnt)
    assert_bit_exact(got, want, cnt)


# c3 paths 95 and 122 took another step sequence when the device's arc
# cos/sin was correctly rounded, which differs from glibc's sincos in the
# last bit for about 1 angle in 350.
@pytest.mark.parametrize("name,first,count", [("c3v", 0, 96), ("c3", 0, 40),
                                              ("c3", 95, 1), ("c3", 122, 1)])
def test_c3_sample_bit_exact(ctx, ref, name, first, count):
    usys = W.config_system(name)
    opts = abi.default_options()
    g, b, a, ns = ref.homotopy(usys, opts.seed)
    start = ref.start_points(usys, opts.seed, first, count)
    print(f"[gpu] {name} paths {first}..{first + count}", flush=True)
    got = run_device(ctx, usys, opts, (g, b, a), start)
    want = ref.endpoints(usys, opts, first, count)
    assert_bit_exact(got, want, count)


def scaled_system(usys, s):
    return abi.UserSystemData(usys.n, [(c * s, e) for c, e in usys.equations])


# The fast unit divides through a reciprocal formed once per divisor, which
# is the compiler's own division only for numerators of magnitude in
# [2^-969, 2^768); outside it (and for NaN or infinite entries) the unit is
# redone with exact divisions. Coefficients scaled by 2^790 or 2^-990 (exact
# scalings) put the LU's and solves' numerators there.
@pytest.mark.parametrize("name", ["c1", "sparse_n8"])
@pytest.mark.parametrize("scale", [2.0 ** 790, 2.0 ** -990])
def test_extreme_magnitudes_redone_exactly(ctx, ref, name, scale):
    base = W.config_system("c1") if name == "c1" else W.random_sparse(170)
    assert name == "c1" or base.n == 8  # the 16-lane grid
    usys = scaled_system(base, scale)
    opts = abi.default_options()
    g, b, a, ns = ref.homotopy(usys, opts.seed)
    cnt = min(ns, 64)
    start = ref.start_points(usys, opts.seed, 0, cnt)
    print(f"[gpu] {name} x {scale:.3e}: {cnt} paths", flush=True)
    got = run_device(ctx, usys, opts, (g, b, a), start)
    assert ctx.exact_reruns() > 0
    want = ref.endpoints(usys, opts, 0, cnt)
    for f in FIELDS + ("eg_error", "t"):
        assert np.array_equal(getattr(got, f)[:cnt], getattr(want, f)[:cnt],
                              equal_nan=True), f
    assert np.array_equal(got.z[:cnt], want.z[:cnt], equal_nan=True)


def test_ordinary_units_rarely_redone(ctx, ref):
    usys = W.config_system("c2")
    opts = abi.default_options()
    g, b, a, ns = ref.homotopy(usys, opts.seed)
    cnt = 256
    start = ref.start_points(usys, opts.seed, 0, cnt)
    got = run_device(ctx, usys, opts, (g, b, a), start)
    units = ctx.counters()["units"]
    assert ctx.exact_reruns() <= units // 1000, (ctx.exact_reruns(), units)
    assert_bit_exact(got, ref.endpoints(usys, opts, 0, cnt), cnt)


def test_random_sparse_batch_bit_exact(ctx, ref):
    """Random sparse systems with n = 3, 4, 5, 7, 8 (2302 paths) in one
    batched call, every path against the reference."""
    from polyhomocont_amd import host
    H = host.library()
    pool = W.batch_systems(91)
    systems = [pool[i] for i in (0, 1, 16, 17, 41, 42, 68, 90)]
    opts = abi.default_options()
    items = []
    for usys in systems:
        g, b, a, ns = H.homotopy(usys, opts.seed)
        items.append((usys.packed(), abi.HomotopyData(g, b, a), opts,
                      H.start_points(usys, opts.seed)))
    print(f"[gpu] batch of {len(items)} systems, "
          f"{sum(u.bezout for u in systems)} paths", flush=True)
    outs = ctx.track_batch(items)
    for usys, got in zip(systems, outs):
        assert_bit_exact(got, ref.endpoints(usys, opts), usys.bezout)


def test_option_variants_match_reference(ctx, ref):
    S = W.ref_test_systems()
    cases = [
        ("2x2", abi.default_options(use_endgame=0)),
        ("cubic", abi.default_options(tracker_step_max=1e-3, tracker_step_init=1e-3,
                                      tracker_max_steps=10)),
        ("mixed", abi.default_options(tracker_step_min=0.02)),
        ("double_root", abi.default_options(endgame_nsamples=5, seed=3)),
        ("mult4", abi.default_options(endgame_max_winding=2, seed=9)),
        ("infinity", abi.default_options(tracker_corrector_maxit=1, seed=2)),
    ]
    for name, opts in cases:
        usys = S[name]
        g, b, a, ns = ref.homotopy(usys, opts.seed)
        start = ref.start_points(usys, opts.seed)
        got = run_device(ctx, usys, opts, (g, b, a), start)
        want = ref.endpoints(usys, opts)
        assert_bit_exact(got, want, ns)
    # the forced failure codes really occur
    got = run_device(ctx, S["cubic"], cases[1][1],
                     ref.homotopy(S["cubic"], 1)[:3], ref.start_points(S["cubic"], 1))
    assert (got.status[:3] == 2).all()


def test_batch_equals_separate_calls(ctx, host):
    """A multi-system launch gives each system exactly its own results."""
    systems = [W.random_sparse(104 + s) for s in range(6)]
    systems += [W.config_system("c1"), W.ref_test_systems()["mixed"]]
    opts = abi.default_options()
    items, single = [], []
    for usys in systems:
        g, b, a, ns = host.homotopy(usys, opts.seed)
        start = host.start_points(usys, opts.seed, 0, min(ns, 96))
        item = (host.pack(usys), abi.HomotopyData(g, b, a), opts, start)
        items.append(item)
        single.append(ctx.track_paths(*item))
    print("[gpu] batch of", len(items), flush=True)
    outs = ctx.track_batch(items)
    for o, s, it in zip(outs, single, items):
        assert_bit_exact(o, s, len(it[3]))


def test_batch_autotune_is_used_and_bit_exact(ctx, host):
    """The batch autotuner picks a launch shape per lanes class; the next
    batch of the same class shapes launches with those shapes, and every
    system's records equal its own phc_track_paths call bit for bit."""
    # lanes classes 4 (n = 3), 8 (n = 4..7) and 16 (n = 8); the first 128
    # paths of each (at most a few hundred steps per path)
    pool = W.batch_systems(93)
    systems = [pool[i] for i in (41, 42, 43, 44, 16, 17, 0, 90, 91, 92,
                                 68, 69, 70)]
    opts = abi.default_options()
    items, single = [], []
    for usys in systems:
        g, b, a, ns = host.homotopy(usys, opts.seed)
        start = host.start_points(usys, opts.seed, 0, min(ns, 128))
        item = (host.pack(usys), abi.HomotopyData(g, b, a), opts, start)
        items.append(item)
        single.append(ctx.track_paths(*item))
    try:
        choices = ctx.autotune_batch(items)
        print(f"[gpu] batch autotune of {len(items)} systems: {choices}",
              flush=True)
        assert sorted(c["G"] for c in choices) == sorted(
            {ctx.lanes_per_path(it[0].n + 1) for it in items})
        outs = ctx.track_batch(items)
        plan = {p["G"]: p for p in ctx.last_plan()}
        for c in choices:
            p = plan[c["G"]]
            assert p["mapping"] == c["mapping"], (c, p)
            assert p["threads"] == c["threads"], (c, p)
            assert p["systems"] == sum(
                ctx.lanes_per_path(it[0].n + 1) == c["G"] for it in items)
        for o, s, it in zip(outs, single, items):
            assert_bit_exact(o, s, len(it[3]))
    finally:
        ctx.clear_tuning()


@pytest.mark.parametrize("mapping", [1, 2, 3])
def test_lane_mappings_bit_exact(ctx, mapping):
    """One path per lane group, one path per wavefront (terms split across
    replicas, ordered sums in LDS) and one path per G*G-lane grid (entry per
    lane) give the reference's records."""
    G = golden()
    from polyhomocont_amd import host
    ctx.set_mapping(mapping)
    try:
        for name in ("c1", "katsura4", "cyclic5", "mixed", "triple_root",
                     "infinity"):
            usys = W.config_system(name) if name == "c1" else W.ref_test_systems()[name]
            opts = abi.default_options()
            consts = (G[name + "/gamma"][0], G[name + "/b"], G[name + "/patch"])
            start = host.library().start_points(usys, opts.seed)
            print(f"[gpu] mapping {mapping}: {name}", flush=True)
            got = run_device(ctx, usys, opts, consts, start)
            n = len(start)
            for f in FIELDS + ("eg_error", "t"):
                assert np.array_equal(getattr(got, f)[:n], G[f"{name}/{f}"]), (name, f)
            assert np.array_equal(got.z[:n], G[name + "/z"]), name
    finally:
        ctx.set_mapping(0)


@pytest.mark.parametrize("mapping", [1, 2])
def test_c2_other_mappings_match_reference(ctx, ref, mapping):
    usys = W.config_system("c2")
    opts = abi.default_options()
    g, b, a, ns = ref.homotopy(usys, opts.seed)
    start = ref.start_points(usys, opts.seed, 0, 96)
    ctx.set_mapping(mapping)
    try:
        got = run_device(ctx, usys, opts, (g, b, a), start)
    finally:
        ctx.set_mapping(0)
    assert_bit_exact(got, ref.endpoints(usys, opts, 0, 96), 96)


def test_grid_mapping_is_default_for_small_systems(ctx, host):
    """N <= 8 runs on the grid unit unless another mapping is forced."""
    from polyhomocont_amd import device
    for usys in (W.config_system("c1"), W.ref_test_systems()["cubic"]):
        opts = abi.default_options()
        g, b, a, ns = host.homotopy(usys, opts.seed)
        st = device.Staged(ctx, host.pack(usys), abi.HomotopyData(g, b, a),
                           opts, host.start_points(usys, opts.seed))
        try:
            assert st.launch_info()["mapping"] == "grid"
        finally:
            st.close()


@pytest.mark.parametrize("mapping,paths_per_wave", [(1, 8), (2, 1), (3, 1)])
def test_unit_counters(ctx, host, mapping, paths_per_wave):
    """The device counts one unit per linearization or backward error of a
    path, and offers units / (passes x paths per wavefront) slots: a
    wavefront that holds one path never loses a slot to divergence."""
    usys = W.config_system("c1")  # N = 5: 8-lane groups, 16 paths
    opts = abi.default_options()
    g, b, a, ns = host.homotopy(usys, opts.seed)
    ctx.set_mapping(mapping)
    try:
        got = ctx.track_paths(host.pack(usys), abi.HomotopyData(g, b, a), opts,
                              host.start_points(usys, opts.seed))
    finally:
        ctx.set_mapping(0)
    c = ctx.counters()
    assert c["units"] == int(got.nlin[:ns].sum()) + int(got.nbe[:ns].sum())
    assert c["unit_slots"] >= c["units"]
    if paths_per_wave == 1:
        assert c["unit_slots"] == c["units"]
    else:
        assert 0.0 < c["occupancy"] < 1.0


def _wide_systems():
    """Systems with 9..12 homogeneous unknowns (16-lane groups): a random
    sparse system with n = 8 and dense quadrics with n = 9 and n = 11."""
    sparse8 = next(s for s in (W.random_sparse(104 + k) for k in range(200))
                   if s.n == 8)
    return [("sparse n=8", sparse8), ("dense n=9 d=2", W.dense_system(9, 2, 107)),
            ("dense n=11 d=2", W.dense_system(11, 2, 108))]


@pytest.mark.parametrize("mapping", [0, 1])
def test_wide_systems_bit_exact(ctx, ref, mapping):
    """16-lane groups decode only the exponent bytes of real unknowns."""
    opts = abi.default_options()
    ctx.set_mapping(mapping)
    try:
        for name, usys in _wide_systems():
            g, b, a, ns = ref.homotopy(usys, opts.seed)
            cnt = min(ns, 24)
            first = ns // 3
            start = ref.start_points(usys, opts.seed, first, cnt)
            print(f"[gpu] wide {name}: paths {first}..{first + cnt} of {ns}",
                  flush=True)
            got = run_device(ctx, usys, opts, (g, b, a), start)
            assert_bit_exact(got, ref.endpoints(usys, opts, first, cnt), cnt)
    finally:
        ctx.set_mapping(0)


def _wide_many_factor_system(n=8):
    """n = 8 (16-lane class) with a five-variable term in every equation:
    more than four nonzero exponents per term, so the grid unit takes its
    general job layout."""
    rng = np.random.default_rng(77)
    eye = np.eye(n, dtype=np.int32)
    eqs = []
    for i in range(n):
        five = np.zeros(n, np.int32)
        five[[(i + k) % n for k in range(5)]] = 1
        e = np.vstack([five[None, :], eye, np.zeros((1, n), np.int32)])
        c = rng.normal(size=len(e)) + 1j * rng.normal(size=len(e))
        eqs.append((c, e))
    return abi.UserSystemData(n, eqs)


def test_wide_general_layout_bit_exact(ctx, ref):
    """The 16-lane grid with terms of five factors (general job layout)."""
    usys = _wide_many_factor_system()
    opts = abi.default_options(seed=5)
    g, b, a, ns = ref.homotopy(usys, opts.seed)
    first, cnt = 1000, 6
    start = ref.start_points(usys, opts.seed, first, cnt)
    print(f"[gpu] wide general layout: paths {first}..{first + cnt} of {ns}",
          flush=True)
    got = run_device(ctx, usys, opts, (g, b, a), start)
    assert_bit_exact(got, ref.endpoints(usys, opts, first, cnt), cnt)


def test_autotuned_launch_is_used_and_bit_exact(ctx, host):
    """The autotuner's choice is applied to later launches of the same
    shape, and the records stay the reference's."""
    from polyhomocont_amd import device
    G = golden()
    for name in ("c1", "katsura4"):
        usys = W.config_system(name) if name == "c1" else W.ref_test_systems()[name]
        opts = abi.default_options()
        consts = (G[name + "/gamma"][0], G[name + "/b"], G[name + "/patch"])
        start = host.start_points(usys, opts.seed)
        hom = abi.HomotopyData(*consts)
        choice = ctx.autotune(usys.packed(), hom, opts, start)
        print(f"[gpu] autotune {name}: {choice}", flush=True)
        assert choice["mapping"] in ("row", "split", "grid")
        assert choice["threads"] in (64, 128, 256)
        st = device.Staged(ctx, usys.packed(), hom, opts, start)
        try:
            info = st.launch_info()
            assert info["mapping"] == choice["mapping"]
            assert info["threads"] == choice["threads"]
            st.run()
            got = st.fetch()
        finally:
            st.close()
        n = len(start)
        for f in FIELDS + ("eg_error", "t"):
            assert np.array_equal(getattr(got, f)[:n], G[f"{name}/{f}"]), (name, f)
        assert np.array_equal(got.z[:n], G[name + "/z"]), name
    ctx.clear_tuning()

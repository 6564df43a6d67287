"""The host side of the boundary, on the CPU.

* Sparse systems the kernels cannot hold (16 variables, an exponent of
  300) are valid reference inputs: `solve` tracks them on the host and
  returns the reference's results, without needing a device.
* set_device(-1) asks for the host tracker explicitly; any device failure
  stays an error (solve never falls back to the host silently).
* The host tracker runs paths over OpenMP threads with one copy of the
  system per thread: records equal the serial run's and the reference's
  for user evaluators (the path of the reference's test/solve.f90 user
  types), and concurrent solves from two threads give the same results.
* Deduplication in path order (first match, src/solve.f90:379-400) through
  the key window equals the linear scan, on planted near-duplicates at
  the tolerance boundary.
"""
import threading

import numpy as np
import pytest

from polyhomocont_amd import abi, workloads as W

PATH_FIELDS = ("status", "winding", "eg_status", "naccepted", "nrejected",
               "solution_index", "singular_index", "at_infinity", "singular",
               "is_real")


def assert_same_solve(a, b, exact=True):
    assert a.counts() == b.counts()
    n = a.summary.npaths
    for f in PATH_FIELDS:
        assert np.array_equal(a.p[f][:n], b.p[f][:n]), f
    for f in ("t", "eg_error", "cond", "residual"):
        assert np.array_equal(a.p[f][:n], b.p[f][:n]), f
    assert np.array_equal(a.p["z"][:n], b.p["z"][:n])
    s = a.summary.nsolutions
    assert np.array_equal(a.solutions[:s], b.solutions[:s])


def wide_system(n=16, seed=31):
    """n = 16 (17 homogeneous unknowns, beyond the kernels' 16): a chain of
    linear equations closed by one quadric, Bezout number 2."""
    rng = np.random.default_rng(seed)
    eqs = []
    for i in range(n - 1):
        e = np.zeros((3, n), np.int32)
        e[0, i] = 1
        e[1, i + 1] = 1
        c = rng.normal(size=3) + 1j * rng.normal(size=3)
        eqs.append((c, e))
    e = np.zeros((3, n), np.int32)
    e[0, n - 1] = 2
    e[1, 0] = 1
    eqs.append((rng.normal(size=3) + 1j * rng.normal(size=3), e))
    return abi.UserSystemData(n, eqs)


def high_degree_system():
    """x^300 + c1 x + c0: an exponent beyond the kernels' 255."""
    e = np.array([[300], [1], [0]], np.int32)
    return abi.UserSystemData(1, [(np.array([1.0, 0.3 - 0.2j, -1.0 + 0.5j]), e)])


@pytest.mark.parametrize("make", [wide_system, high_degree_system],
                         ids=["n16", "exponent300"])
def test_beyond_kernel_limits_solved_on_host(host, ref, make):
    usys = make()
    opts = abi.default_options()
    st, mine = host.solve(usys, opts)
    assert st == 0                       # not err_invalid_input / err_device
    st2, want = ref.solve(usys, opts)
    assert st2 == 0
    assert_same_solve(mine, want)
    assert mine.summary.nsolutions >= 1


def test_host_requested_equals_reference(host, ref):
    usys = W.config_system("c1")
    opts = abi.default_options()
    host.set_device(-1)
    try:
        st, mine = host.solve(usys, opts)
    finally:
        host.set_device(0)
    assert st == 0
    st2, want = ref.solve(usys, opts)
    assert_same_solve(mine, want)


def _gpu_present():
    import ctypes as C
    from polyhomocont_amd import device
    h = C.c_void_p()
    return device.lib().phc_ctx_shared(0, C.byref(h)) == 0


def test_missing_device_is_an_error(host):
    """Without a GPU a sparse system inside the kernels' limits is an
    err_device, not a silent host run."""
    if _gpu_present():
        pytest.skip("a GPU is present")
    st, _ = host.solve(W.config_system("c1"), abi.default_options())
    assert st == -1


USER_CASES = ["user_cf", "katsura3", "mixed", "infinity", "double_root"]


@pytest.mark.parametrize("name", USER_CASES)
def test_user_evaluator_threads_match_serial_and_reference(host, ref, name):
    usys = W.ref_test_systems()[name]
    opts = abi.default_options()
    prev = host.set_threads(1)
    try:
        st1, serial = host.solve_user(usys, opts)
        host.set_threads(8)
        st8, threaded = host.solve_user(usys, opts)
    finally:
        host.set_threads(prev)
    assert st1 == st8 == 0
    assert_same_solve(threaded, serial)
    st, want = ref.solve_user(usys, opts)
    assert st == 0
    assert_same_solve(threaded, want)


def test_concurrent_solves_from_two_threads(host):
    """solve is reentrant: two threads solving at once get the results of
    solving one after the other."""
    systems = [W.ref_test_systems()["katsura3"], W.config_system("c1")]
    opts = abi.default_options()
    alone = [host.solve_user(u, opts)[1] for u in systems]
    got = [None, None]

    def work(i):
        for _ in range(3):
            st, r = host.solve_user(systems[i], opts)
            assert st == 0
            got[i] = r

    th = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for a, b in zip(got, alone):
        assert_same_solve(a, b)


def _linear_scan(xs, tol):
    """first_match in path order (src/solve.f90:379-400): index of each
    point among the stored ones, stored = points without an earlier
    match."""
    stored, out = [], []
    for x in xs:
        scale = max(1.0, float(np.abs(x).max()))
        idx = 0
        for j, p in enumerate(stored, start=1):
            if np.abs(x - p).max() <= tol * scale:
                idx = j
                break
        if idx == 0:
            stored.append(x)
            idx = len(stored)
        out.append(idx)
    return out


def test_dedup_window_equals_linear_scan(host):
    """Singular endpoints (winding 2: not refined) with planted
    near-duplicates at 0.5 and 1.5 times the singular tolerance, repeated
    points and a spread of magnitudes: the solve's singular_index equals a
    linear first-match scan over the same points."""
    usys = W.ref_test_systems()["katsura3"]          # n = 4, N = 5
    opts = abi.default_options()
    tol = opts.singular_dedup_tol
    rng = np.random.default_rng(7)
    N = usys.n + 1
    base = rng.normal(size=(40, N)) + 1j * rng.normal(size=(40, N))
    base *= 10.0 ** rng.integers(-2, 4, size=(40, 1))
    base[:, -1] = 1.0 + 0.1j                          # finite points
    pts = [base[k % 40] for k in range(40)]
    for k in range(200):
        b = base[rng.integers(0, 40)].copy()
        x = b[:-1] / b[-1]
        scale = max(1.0, np.abs(x).max())
        d = rng.choice([0.0, 0.5, 0.999, 1.5, 3.0]) * tol * scale
        j = rng.integers(0, N - 1)
        b[j] += d * b[-1] * np.exp(2j * np.pi * rng.random())
        pts.append(b)
    z = np.array(pts)
    npaths = len(z)
    rec = abi.EndpointArrays(npaths, N)
    rec.status[:] = 0
    rec.winding[:] = 2
    rec.eg_status[:] = 0
    rec.eg_error[:] = 0.0
    rec.t[:] = 0.0
    rec.z[:] = z
    st, res = host.classify(usys, opts, rec)
    assert st == 0
    assert res.summary.nsingular > 40 and res.summary.nsingular < npaths
    x = res.p["x"][:npaths]
    assert np.all(res.p["singular"][:npaths] == 1)
    want = _linear_scan(list(x), tol)
    assert list(res.p["singular_index"][:npaths]) == want
    mult = np.bincount(want)[1:]
    assert list(res.multiplicities[:res.summary.nsingular]) == list(mult)


def test_host_tracker_on_serving_requests_matches_fixture(host):
    """The host tracker (set_device(-1)) on small requests of the serving
    stream gives the reference `solve`'s results held by the fixture:
    counts, every path's discrete fields and classification, and the digest
    of the result's doubles."""
    import os
    from tests.golden.make_golden import (COUNT_KEYS, SERVING_REQUESTS,
                                          endpoint_digest)
    p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                     "serving_stream.npz")
    if not os.path.exists(p):
        pytest.skip("serving_stream.npz not generated")
    F = np.load(p)
    stream = W.serving_stream(SERVING_REQUESTS)
    off = F["offsets"]
    small = [i for i in range(len(stream)) if stream[i].bezout <= 64][:6]
    assert small
    opts = abi.default_options()
    host.set_device(-1)
    try:
        for i in small:
            st, res = host.solve(stream[i], opts)
            assert st == 0
            lo, hi = off[i], off[i + 1]
            n = hi - lo
            c = res.counts()
            assert [c[k] for k in COUNT_KEYS] == list(F["counts"][i]), i
            for f in ("status", "winding", "eg_status", "naccepted",
                      "nrejected", "solution_index", "singular_index",
                      "at_infinity", "singular"):
                assert np.array_equal(res.p[f][:n], F[f][lo:hi]), (i, f)
            assert np.array_equal(
                endpoint_digest(res.p["z"][:n], res.p["t"][:n],
                                res.p["eg_error"][:n]), F["digest"][i]), i
    finally:
        host.set_device(0)

"""Host-side constants must be the reference's bit for bit: the generator
(src/random.f90), gamma/b/patch (src/homotopy.f90:104-149) and the start
points (src/homotopy.f90:151-177). Mirrors the reference's test/random.f90
and test/homotopy.f90 checks, plus exact equality with the compiled
reference and with the Python restatement used by the workload generator."""
import numpy as np
import pytest

from polyhomocont_amd import abi, workloads as W


def test_python_generator_matches_reference(ref):
    for seed in (1, 7, 42, 2026, -5):
        g = W.Lecuyer(seed)
        got = np.array([g.uniform() for _ in range(500)])
        assert np.array_equal(got, ref.rng(seed, 0, 500))
        g = W.Lecuyer(seed)
        got = np.array([g.normal_complex() for _ in range(200)])
        assert np.array_equal(got, ref.rng(seed, 2, 200))
        g = W.Lecuyer(seed)
        got = np.array([g.unit_complex() for _ in range(200)])
        assert np.array_equal(got, ref.rng(seed, 1, 200))


def test_generator_statistics():
    # test/random.f90:30-91: range, moments, unit modulus, complex moments
    g = W.Lecuyer(42)
    u = np.array([g.uniform() for _ in range(200000)])
    assert u.min() > 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 5e-3 and abs(u.var() - 1 / 12) < 5e-3
    z = np.array([g.unit_complex() for _ in range(10000)])
    assert np.abs(np.abs(z) - 1).max() < 1e-15
    w = np.array([g.normal_complex() for _ in range(100000)])
    assert abs(np.mean(np.abs(w) ** 2) - 1.0) < 2e-2
    assert abs(w.mean()) < 2e-2


def test_fortran_host_constants_match_reference(ref, host):
    for name, usys in list(W.ref_test_systems().items())[:9] + [
            ("c1", W.config_system("c1")), ("c3v", W.config_system("c3v"))]:
        for seed in (1, 5):
            g1, b1, a1, n1 = ref.homotopy(usys, seed)
            g2, b2, a2, n2 = host.homotopy(usys, seed)
            assert n1 == n2 == usys.bezout
            assert g1 == g2 and np.array_equal(b1, b2) and np.array_equal(a1, a2)
            cnt = min(n1, 64)
            assert np.array_equal(ref.start_points(usys, seed, 0, cnt),
                                  host.start_points(usys, seed, 0, cnt))


def test_start_points_solve_start_system(ref):
    # test/homotopy.f90:62-93: H(z, 1) = 0 at every start point, distinct
    usys = W.ref_test_systems()["user_cf"]
    g, b, a, ns = ref.homotopy(usys, 5)
    assert ns == 6 and abs(abs(g) - 1) < 1e-15
    Z = ref.start_points(usys, 5)
    for z in Z:
        h, hz, ht = ref.eval(usys, 5, z, 1.0)
        assert np.abs(h).max() / max(1.0, np.abs(z).max()) ** 3 < 1e-14
    d = min(np.abs(Z[i] - Z[j]).max() for i in range(6) for j in range(i + 1, 6))
    assert d > 1e-3


def test_packing_matches_builder(host):
    # zero terms dropped, w exponent appended (src/sparse.f90:111-138)
    usys = abi.UserSystemData(2, [
        ([1.0, 1.0, -5.0], [[2, 0], [0, 2], [0, 0]]),
        ([1.0, -2.0, 0.0, 1.0], [[1, 2], [0, 1], [5, 5], [1, 0]])])
    p = host.pack(usys)
    assert list(p.first) == [0, 3, 6] and list(p.deg) == [2, 3]
    assert p.hexps.tolist() == [[2, 0, 0], [0, 2, 0], [0, 0, 2],
                                [1, 2, 0], [0, 1, 2], [1, 0, 2]]
    q = usys.packed()
    assert np.array_equal(p.hexps, q.hexps) and np.array_equal(p.coefs, q.coefs)


def test_invalid_systems_rejected(host, ref):
    # test/solve.f90:364-377 and test/sparse.f90:159-202
    bad = abi.UserSystemData(2, [([1.0, -1.0], [[2, 0], [0, 0]])])
    st1, _ = host.solve(bad, abi.default_options())
    st2, _ = ref.solve(bad, abi.default_options())
    assert st1 == st2 == 1
    neg = abi.UserSystemData(2, [([1.0], [[1, -1]]), ([1.0], [[0, 1]])])
    assert host.solve(neg, abi.default_options())[0] == 1
    const = abi.UserSystemData(2, [([3.0, 0.0], [[0, 0], [1, 1]]),
                                   ([1.0], [[0, 1]])])
    assert host.solve(const, abi.default_options())[0] == 1

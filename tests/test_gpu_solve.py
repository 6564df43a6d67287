"""End to end: the Fortran `solve` (GPU tracking + host classification)
against the reference `solve` and against the reference tests' known
answers (test/solve.f90, test/katsura.f90, test/cyclic.f90)."""
import json
import os

import numpy as np
import pytest

from polyhomocont_amd import abi, workloads as W

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def set_distance(a, b):
    """max over points of a of the distance to the nearest point of b."""
    if len(a) != len(b):
        return np.inf
    d = 0.0
    for x in a:
        d = max(d, min(np.abs(x - y).max() for y in b))
    return d


@pytest.mark.parametrize("name", ["cubic", "2x2", "infinity", "double_root",
                                  "triple_root", "mixed", "mult4", "user_cf",
                                  "univariate5", "katsura3", "katsura4",
                                  "cyclic5", "cyclic6"])
def test_solve_matches_reference(host, ref, name):
    usys = W.ref_test_systems()[name]
    opts = abi.default_options()
    print(f"[gpu] solve {name}", flush=True)
    st, mine = host.solve(usys, opts)
    assert st == 0
    st2, want = ref.solve(usys, opts)
    assert mine.counts() == want.counts()
    n = want.summary.nsolutions
    assert np.abs(mine.solutions[:n] - want.solutions[:n]).max(initial=0) <= 1e-10


def test_reference_known_answers(host):
    S = W.ref_test_systems()
    o = abi.default_options()
    c = lambda nm, **kw: host.solve(S[nm], abi.default_options(**kw))[1].counts()
    assert c("univariate5")["nsolutions"] == 5 and c("univariate5")["nreal"] == 3
    for seed in range(1, 11):                      # test/solve.f90:158-164
        r = c("2x2", seed=seed)
        assert r["nsolutions"] == 4 and r["nreal"] == 4
    r = c("infinity")
    assert (r["nsolutions"], r["ninfinite"], r["nfailed"]) == (2, 2, 0)
    r = c("double_root")
    assert (r["nsolutions"], r["nsingular"], r["multiplicities"]) == (1, 1, [2])
    r = c("mixed")
    assert (r["nsolutions"], r["nsingular"], r["ninfinite"], r["multiplicities"]) == (1, 1, 3, [2])
    r = c("mult4")
    assert (r["nsolutions"], r["nsingular"], r["multiplicities"]) == (0, 1, [4])
    assert c("2x2", use_endgame=0)["nsolutions"] == 4
    for nm, want in (("cyclic5", 70), ("cyclic6", 156)):      # test/cyclic.f90
        r = c(nm)
        assert r["nsolutions"] == want and r["nsingular"] == 0 and r["nfailed"] == 0
        assert r["nduplicates"] == 0 and r["ninfinite"] == r["npaths"] - want


def test_katsura_against_reference_vectors(host):
    gold = json.load(open(os.path.join(HERE, "golden", "katsura.json")))
    for n, key, nreal in ((3, "sols3", 6), (4, "sols4", 12)):
        v = np.array(gold[key])
        sols = (v[:, 0] + 1j * v[:, 1]).reshape(-1, n + 1)  # columns of (n+1)
        st, res = host.solve(W.katsura(n), abi.default_options())
        s = res.summary
        assert st == 0 and s.nsolutions == 2 ** n and s.nreal == nreal
        assert s.nfailed == 0 and s.nduplicates == 0
        assert set_distance(res.solutions[:s.nsolutions], sols) <= 1e-12
        assert res.p["residual"][:s.npaths].max() <= 1e-14


def test_c1_full_and_c2_counts(host):
    """Generic dense systems: Bezout-many finite regular solutions."""
    for name, npaths in (("c1", 16), ("c2", 2187)):
        print(f"[gpu] solve {name}", flush=True)
        st, res = host.solve(W.config_system(name), abi.default_options())
        r = res.counts()
        assert st == 0 and r["npaths"] == npaths
        assert r["nsolutions"] == npaths and r["nduplicates"] == 0
        assert r["nfailed"] == r["ninfinite"] == r["nsingular"] == 0
        assert res.p["residual"][:npaths].max() < 1e-12


def test_concurrent_solves_share_the_device(host):
    """solve is reentrant on the GPU path: two threads solving at once on
    the library's shared device context get the results of solving one
    after the other (src/solve.f90 keeps no module state; CLAUDE.md of the
    reference asks the same of a parallel caller)."""
    import threading
    systems = [W.ref_test_systems()["katsura4"], W.config_system("c1"),
               W.ref_test_systems()["cyclic5"]]
    opts = abi.default_options()
    alone = []
    for u in systems:
        st, r = host.solve(u, opts)
        assert st == 0
        alone.append(r)
    got = [[] for _ in systems]
    errors = []

    def work(i):
        try:
            for _ in range(3):
                st, r = host.solve(systems[i], opts)
                assert st == 0
                got[i].append(r)
        except Exception as e:  # reported by the main thread
            errors.append(e)

    print("[gpu] concurrent solves from 3 threads", flush=True)
    th = [threading.Thread(target=work, args=(i,)) for i in range(len(systems))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for rs, a in zip(got, alone):
        for r in rs:
            assert r.counts() == a.counts()
            n = a.summary.npaths
            for f in ("status", "winding", "solution_index", "singular_index"):
                assert np.array_equal(r.p[f][:n], a.p[f][:n]), f
            assert np.array_equal(r.p["z"][:n], a.p["z"][:n])

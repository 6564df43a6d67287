"""Seeded synthetic polynomial systems for tests and benchmarks.

Coefficients are drawn with the reference's own generator (L'Ecuyer's
combined multiplicative LCG, Commun. ACM 31 (1988) 742; seeding and draws as
in the reference src/random.f90:43-109), restated here in Python so the
workloads are reproducible without any compiled code.

Workloads (sizes as named in the benchmark configuration):
  dense_system(n, d, seed)     every equation has all monomials of degree
                               <= d, complex-Gaussian coefficients
  cyclic_like(seed, variant)   n = 7: equation i < 7 sums i cyclically
                               consecutive products, equation 7 is
                               c x1...x7 + c0; the variant replaces
                               equation 3 by (x1 - x2)^2 (a0 + a.x)
  random_sparse(seed)          n ~ U{3..8}, d_i ~ U{2..4}, 5..40 distinct
                               terms per equation under the degree bound,
                               one of full degree d_i
"""
import itertools
import math

import numpy as np

from .abi import UserSystemData

_M1 = 2147483563
_M2 = 2147483399
_A1 = 40014
_A2 = 40692


class Lecuyer:
    """The reference rng (src/random.f90:29-109): integer recurrence, then
    the same double-precision formulas for the derived draws."""

    def __init__(self, seed=None):
        self.s1, self.s2 = 1234567, 7654321
        if seed is not None:
            self.seed(seed)

    def seed(self, seed):
        s = abs(int(seed))
        self.s1 = s % (_M1 - 1) + 1
        self.s2 = (s * 7919 + 104729) % (_M2 - 1) + 1
        for _ in range(16):
            self.uniform()

    def uniform(self):
        self.s1 = (_A1 * self.s1) % _M1
        self.s2 = (_A2 * self.s2) % _M2
        z = self.s1 - self.s2
        if z < 1:
            z += _M1 - 1
        return float(z) / float(_M1)

    def unit_complex(self):
        phi = 2.0 * math.pi * self.uniform()
        return complex(math.cos(phi), math.sin(phi))

    def normal_complex(self):
        r = math.sqrt(-math.log(self.uniform()))
        phi = 2.0 * math.pi * self.uniform()
        return complex(r * math.cos(phi), r * math.sin(phi))

    def integer(self, lo, hi):
        """Uniform integer in [lo, hi]."""
        return lo + min(int((hi - lo + 1) * self.uniform()), hi - lo)


def monomials(n, d):
    """All exponent vectors of n variables with total degree <= d, in graded
    order (degree 0 first), lexicographic within a degree."""
    out = []
    for deg in range(d + 1):
        for combo in itertools.combinations_with_replacement(range(n), deg):
            e = [0] * n
            for j in combo:
                e[j] += 1
            out.append(e)
    # lexicographically decreasing within each degree for a stable order
    return out


def dense_system(n, d, seed):
    g = Lecuyer(seed)
    mons = np.array(monomials(n, d), np.int32)
    eqs = []
    for _ in range(n):
        c = np.array([g.normal_complex() for _ in range(len(mons))])
        eqs.append((c, mons.copy()))
    return UserSystemData(n, eqs)


def cyclic_like(seed, variant=False, n=7):
    g = Lecuyer(seed)
    eqs = []
    for i in range(1, n):
        exps = np.zeros((n, n), np.int32)
        for j in range(n):
            for l in range(i):
                exps[j, (j + l) % n] = 1
        c = np.array([g.normal_complex() for _ in range(n)])
        eqs.append((c, exps))
    exps = np.zeros((2, n), np.int32)
    exps[0, :] = 1
    eqs.append((np.array([g.normal_complex(), g.normal_complex()]), exps))
    if variant:
        # (x1 - x2)^2 (a0 + sum_j a_j x_j), expanded; degree 3 like the
        # equation it replaces, so the Bezout number stays 7!.
        a = [g.normal_complex() for _ in range(n + 1)]
        sq = {(2, 0): 1.0, (1, 1): -2.0, (0, 2): 1.0}  # powers of (x1, x2)
        terms = {}
        for (p1, p2), cs in sq.items():
            for j in range(n + 1):  # j = 0: constant a0
                e = [0] * n
                e[0] += p1
                e[1] += p2
                if j > 0:
                    e[j - 1] += 1
                key = tuple(e)
                terms[key] = terms.get(key, 0.0) + cs * a[j]
        keys = sorted(terms)
        eqs[2] = (np.array([terms[k] for k in keys]), np.array(keys, np.int32))
    return UserSystemData(n, eqs)


def random_sparse(seed):
    g = Lecuyer(seed)
    n = g.integer(3, 8)
    eqs = []
    for _ in range(n):
        d = g.integer(2, 4)
        allm = monomials(n, d)
        top = [m for m in allm if sum(m) == d]
        cap = math.comb(n + d, d)
        m = min(g.integer(5, 40), cap)
        chosen = [top[g.integer(0, len(top) - 1)]]
        seen = {tuple(chosen[0])}
        while len(chosen) < m:
            e = allm[g.integer(0, len(allm) - 1)]
            if tuple(e) not in seen:
                seen.add(tuple(e))
                chosen.append(e)
        c = np.array([g.normal_complex() for _ in range(m)])
        eqs.append((c, np.array(chosen, np.int32)))
    return UserSystemData(n, eqs)


def config_system(name):
    """Named workloads: c1, c2, c3, c3v (squared-factor variant)."""
    if name == "c1":
        return dense_system(4, 2, 101)
    if name == "c2":
        return dense_system(7, 3, 102)
    if name == "c3":
        return cyclic_like(103, variant=False)
    if name == "c3v":
        return cyclic_like(103, variant=True)
    raise KeyError(name)


def batch_systems(count=512, seed0=104):
    return [random_sparse(seed0 + s) for s in range(count)]


# Systems of the reference's own tests (test/*.f90), as (n, equations).
def ref_test_systems():
    two = lambda *rows: np.array(rows, np.int32)
    S = {}
    S["cubic"] = UserSystemData(1, [([1.0, -2.0, 1.0], two([3], [1], [0]))])
    S["2x2"] = UserSystemData(2, [([1.0, 1.0, -5.0], two([2, 0], [0, 2], [0, 0])),
                                   ([1.0, -2.0], two([1, 1], [0, 0]))])
    S["infinity"] = UserSystemData(2, [([1.0, -1.0], two([1, 1], [0, 0])),
                                        ([1.0, -4.0], two([2, 0], [0, 0]))])
    S["double_root"] = UserSystemData(1, [([1.0, -1.0, -1.0, 1.0],
                                           two([3], [2], [1], [0]))])
    S["triple_root"] = UserSystemData(1, [([1.0, -1.0, -3.0, 5.0, -2.0],
                                           two([4], [3], [2], [1], [0]))])
    S["mixed"] = UserSystemData(2, [([1.0, -1.0], two([1, 1], [0, 0])),
                                     ([1.0, -3.0, 2.0], two([3, 0], [1, 0], [0, 0]))])
    S["mult4"] = UserSystemData(2, [([1.0], two([2, 0])),
                                     ([1.0, 1.0], two([0, 2], [1, 1]))])
    S["user_cf"] = UserSystemData(2, [([1.0, 1.0, -5.0], two([2, 0], [0, 2], [0, 0])),
                                       ([1.0, -2.0, 1.0], two([1, 2], [0, 1], [1, 0]))])
    roots = [1.0, -2.0, 3j, -3j, 0.5]
    c = np.zeros(6, np.complex128)
    c[0] = 1.0
    for k, r in enumerate(roots, start=1):
        c[1:k + 1] = c[0:k] - r * c[1:k + 1]
        c[0] = -r * c[0]
    S["univariate5"] = UserSystemData(1, [(c, np.arange(6, dtype=np.int32).reshape(6, 1))])
    for nn in (5, 6):
        eqs = []
        for k in range(1, nn):
            e = np.zeros((nn, nn), np.int32)
            for i in range(nn):
                for j in range(k):
                    e[i, (i + j) % nn] = 1
            eqs.append((np.ones(nn), e))
        eqs.append(([1.0, -1.0], np.array([[1] * nn, [0] * nn], np.int32)))
        S[f"cyclic{nn}"] = UserSystemData(nn, eqs)
    for nn in (3, 4):
        S[f"katsura{nn}"] = katsura(nn)
    return S


def katsura(n):
    """Katsura-n in u_0..u_n (reference test/katsura.f90:146-198)."""
    N = n + 1
    eqs = []
    e = np.zeros((n + 2, N), np.int32)
    c = [1.0] + [2.0] * n + [-1.0]
    e[0, 0] = 1
    for k in range(1, n + 1):
        e[k, k] = 1
    eqs.append((np.array(c), e))
    for m in range(n):
        cs, es = [], []
        for l in range(-n, n + 1):
            if abs(m - l) > n:
                continue
            v = [0] * N
            v[abs(l)] += 1
            v[abs(m - l)] += 1
            cs.append(1.0)
            es.append(v)
        v = [0] * N
        v[m] = 1
        cs.append(-1.0)
        es.append(v)
        eqs.append((np.array(cs), np.array(es, np.int32)))
    return UserSystemData(N, eqs)


def serving_stream(count=512, seed=105, max_paths=4096, min_paths=8):
    """Requests of the serving benchmark: each drawn at random, with equal
    odds, from the random sparse generator (n 3..8, degree 2..4, 5..40
    terms) or as a dense system (n in 2..7, degree 2..3, all monomials);
    those with 8..4096 total-degree paths are kept."""
    g = Lecuyer(seed)
    out = []
    k = 0
    while len(out) < count:
        if g.uniform() < 0.5:
            usys = random_sparse(10_000 + seed * 1000 + k)
        else:
            n = g.integer(2, 7)
            d = g.integer(2, 3)
            usys = dense_system(n, d, 20_000 + seed * 1000 + k)
        k += 1
        if min_paths <= usys.bezout <= max_paths:
            out.append(usys)
    return out

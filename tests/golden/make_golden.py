"""Writes the golden fixtures under tests/golden/ (run in a container that
has /root/reference and a built oracle/_ref; the fixtures are committed,
this script is how they were made).

  python tests/golden/make_golden.py basic   katsura.json, endpoints.npz
  python tests/golden/make_golden.py c3      c3_full.npz
  python tests/golden/make_golden.py c4      c4_batch.npz
  python tests/golden/make_golden.py c2      c2_full.npz
  python tests/golden/make_golden.py serving serving_stream.npz
  python tests/golden/make_golden.py all     all of them

katsura.json   Katsura-3/-4 solution vectors held by the reference's own
               test (test/katsura.f90:22-137; computed there with
               sympy/mpmath by reference/ref_katsura.py), parsed from the
               Fortran literals as data.
endpoints.npz  Per-path endpoint records of the compiled reference
               (oracle/_ref, built by oracle/Makefile from the reference's
               own sources) for the C1 workload and for the systems of the
               reference's tests, at the default options (seed 1).
c3_full.npz    Every one of the 5040 paths of C3 and of its squared-factor
               variant C3v (n = 7): the reference's endpoint records
               (`endpoint`, src/solve.f90:217-252) and the result of the
               reference `solve` (src/solve.f90:154-215): counts,
               multiplicities and the per-path classification.
c2_full.npz    Every one of the 2187 paths of C2 (dense n = 7, degree 3):
               the reference's endpoint records and its `solve`'s counts
               and per-path classification.
serving_stream.npz
               The 512 requests of the serving stream (seed 105, 239,777
               paths): per request the reference `solve`'s counts and
               multiplicities, per path its status, winding, end game
               status, step counts and classification, and a SHA-256
               digest of the result's doubles (final z, t, end game
               error) per request.
c4_batch.npz   All 512 seeded random sparse systems of the batched workload
               (743,290 paths): per path the discrete endpoint fields and
               the reference `solve`'s classification; per system the
               `solve` counts and multiplicities, a SHA-256 digest of the
               endpoint doubles (z, t, end game error; bit-exact check) and
               a weighted sum of the endpoints (tolerance check).

The reference runs serially per path, so the c3/c4 sets take a few
minutes of all host cores (the two 5040-path `solve` calls about 25
minutes each); partial results are kept in --cache so an interrupted run
resumes.
"""
import argparse
import hashlib
import json
import multiprocessing as mp
import os
import re
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF = os.environ.get("PHC_REFERENCE", "/root/reference")

C3_CHUNK = 40


def katsura_vectors():
    text = open(os.path.join(REF, "test", "katsura.f90")).read()
    out = {}
    for name, n, m in (("sols3", 4, 8), ("sols4", 5, 16)):
        start = text.index(f"{name} = reshape([")
        end = text.index("])", start)
        body = text[start:end]
        pairs = re.findall(r"\(\s*([-+0-9.eE]+)_wp\s*,\s*([-+0-9.eE]+)_wp\s*\)", body)
        vals = [[float(a), float(b)] for a, b in pairs]
        assert len(vals) == n * m, (name, len(vals))
        out[name] = vals  # column-major n x m, (re, im)
    return out


def basic():
    import oracle
    from polyhomocont_amd import abi, workloads
    with open(os.path.join(HERE, "katsura.json"), "w") as f:
        json.dump(katsura_vectors(), f)
    R = oracle.ref()
    opts = abi.default_options()
    rec = {}
    systems = {"c1": workloads.config_system("c1")}
    systems.update(workloads.ref_test_systems())
    del systems["cyclic6"]
    for name, usys in systems.items():
        e = R.endpoints(usys, opts)
        g, b, a, ns = R.homotopy(usys, opts.seed)
        rec[name + "/status"] = e.status[:ns]
        rec[name + "/z"] = e.z[:ns]
        rec[name + "/winding"] = e.winding[:ns]
        rec[name + "/eg_status"] = e.eg_status[:ns]
        rec[name + "/eg_error"] = e.eg_error[:ns]
        rec[name + "/naccepted"] = e.naccepted[:ns]
        rec[name + "/nrejected"] = e.nrejected[:ns]
        rec[name + "/t"] = e.t[:ns]
        rec[name + "/gamma"] = np.array([g])
        rec[name + "/b"] = b
        rec[name + "/patch"] = a
    np.savez_compressed(os.path.join(HERE, "endpoints.npz"), **rec)


# ---------------------------------------------------------------- digests
def endpoint_digest(z, t, eg_error):
    """SHA-256 over the endpoint doubles of one system's paths in path
    order: z ([npaths][N] complex128), then t, then the end game errors."""
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(z, np.complex128).tobytes())
    h.update(np.ascontiguousarray(t, np.float64).tobytes())
    h.update(np.ascontiguousarray(eg_error, np.float64).tobytes())
    return np.frombuffer(h.digest(), np.uint8)


def endpoint_weights(seed, npaths, N):
    """Fixed unit-modulus weights of the weighted endpoint sum."""
    r = np.random.default_rng(1000 + seed)
    return np.exp(2j * np.pi * r.random((npaths, N)))


def weighted_sum(seed, z):
    """sum_k sum_j w_kj z_kj and sum |z_kj|: a size-independent checksum of
    one system's endpoints (compared within the stated 1e-10 tolerance)."""
    z = np.asarray(z, np.complex128)
    w = endpoint_weights(seed, *z.shape)
    return complex((w * z).sum()), float(np.abs(z).sum())


# ------------------------------------------------------------------ tasks
_STREAMS = {}


def _task(args):
    kind = args[0]
    import oracle
    from polyhomocont_amd import abi, workloads
    R = oracle.ref()
    opts = abi.default_options()
    t0 = time.time()
    if kind == "c3_endpoints":
        _, name, first, count = args
        e = R.endpoints(workloads.config_system(name), opts, first, count)
        out = {f: getattr(e, f)[:count].copy() for f in
               ("status", "winding", "eg_status", "naccepted", "nrejected",
                "t", "eg_error", "z")}
    elif kind == "c3_solve":  # also C2's solve
        _, name = args
        usys = workloads.config_system(name)
        st, res = R.solve(usys, opts)
        assert st == 0
        out = _solve_fields(res)
    elif kind == "c2_endpoints":
        _, first, count = args
        e = R.endpoints(workloads.config_system("c2"), opts, first, count)
        out = {f: getattr(e, f)[:count].copy() for f in
               ("status", "winding", "eg_status", "naccepted", "nrejected",
                "t", "eg_error", "z")}
    elif kind == "serving":
        _, count, i = args
        if count not in _STREAMS:  # one stream per worker process
            _STREAMS[count] = workloads.serving_stream(count)
        usys = _STREAMS[count][i]
        st, res = R.solve(usys, opts)
        assert st == 0
        np_ = res.summary.npaths
        out = {f: res.p[f][:np_].copy() for f in
               ("status", "winding", "eg_status", "naccepted", "nrejected")}
        out["digest"] = endpoint_digest(res.p["z"][:np_], res.p["t"][:np_],
                                        res.p["eg_error"][:np_])
        out.update(_solve_fields(res))
    else:  # c4 system
        _, s = args
        usys = workloads.random_sparse(104 + s)
        ns = usys.bezout
        e = R.endpoints(usys, opts)
        st, res = R.solve(usys, opts)
        assert st == 0
        out = {f: getattr(e, f)[:ns].copy() for f in
               ("status", "winding", "eg_status", "naccepted", "nrejected")}
        out["digest"] = endpoint_digest(e.z[:ns], e.t[:ns], e.eg_error[:ns])
        wsum, asum = weighted_sum(s, e.z[:ns])
        out["wsum"] = np.array([wsum])
        out["asum"] = np.array([asum])
        out.update(_solve_fields(res))
    out["_seconds"] = time.time() - t0
    return args, out


def _solve_fields(res):
    c = res.counts()
    np_ = c["npaths"]
    out = {"counts": np.array([c[k] for k in COUNT_KEYS], np.int64),
           "multiplicities": np.array(c["multiplicities"], np.int32)}
    for f in ("solution_index", "singular_index", "at_infinity", "singular"):
        out[f] = res.p[f][:np_].copy()
    return out


COUNT_KEYS = ("npaths", "nsolutions", "nreal", "nsingular", "ninfinite",
              "nfailed", "nduplicates")


def _key(args):
    return "_".join(str(a) for a in args)


def run_tasks(tasks, cache, workers):
    os.makedirs(cache, exist_ok=True)
    todo = [t for t in tasks
            if not os.path.exists(os.path.join(cache, _key(t) + ".npz"))]
    print(f"{len(tasks) - len(todo)} of {len(tasks)} tasks cached", flush=True)
    done = 0
    with mp.get_context("spawn").Pool(workers) as pool:
        for args, out in pool.imap_unordered(_task, todo):
            np.savez(os.path.join(cache, _key(args) + ".npz"), **out)
            done += 1
            print(f"[{time.strftime('%H:%M:%S')}] {done}/{len(todo)} "
                  f"{_key(args)} {out['_seconds']:.1f} s", flush=True)
    return {_key(t): dict(np.load(os.path.join(cache, _key(t) + ".npz")))
            for t in tasks}


C2_CHUNK = 81
SERVING_REQUESTS = 512


def c2_tasks():
    return [("c3_solve", "c2")] + [("c2_endpoints", f, C2_CHUNK)
                                   for f in range(0, 2187, C2_CHUNK)]


def serving_tasks():
    from polyhomocont_amd import workloads
    stream = workloads.serving_stream(SERVING_REQUESTS)
    order = sorted(range(len(stream)), key=lambda i: -stream[i].bezout)
    return [("serving", SERVING_REQUESTS, i) for i in order]


def write_c2(res):
    parts = [res[_key(("c2_endpoints", f, C2_CHUNK))]
             for f in range(0, 2187, C2_CHUNK)]
    out = {}
    for f in ("status", "winding", "eg_status"):
        out[f] = np.concatenate([p[f] for p in parts]).astype(np.int8)
    for f in ("naccepted", "nrejected"):
        out[f] = np.concatenate([p[f] for p in parts]).astype(np.int32)
    for f in ("t", "eg_error", "z"):
        out[f] = np.concatenate([p[f] for p in parts])
    sv = res[_key(("c3_solve", "c2"))]
    out["counts"] = sv["counts"]
    out["multiplicities"] = sv["multiplicities"]
    for f in ("solution_index", "singular_index"):
        out[f] = sv[f].astype(np.int32)
    for f in ("at_infinity", "singular"):
        out[f] = sv[f].astype(np.int8)
    out["count_keys"] = np.array(COUNT_KEYS)
    np.savez_compressed(os.path.join(HERE, "c2_full.npz"), **out)


def write_serving(res):
    R_ = SERVING_REQUESTS
    per = [res[_key(("serving", R_, i))] for i in range(R_)]
    out = {"count_keys": np.array(COUNT_KEYS)}
    out["npaths"] = np.array([len(p["status"]) for p in per], np.int64)
    out["offsets"] = np.concatenate([[0], np.cumsum(out["npaths"])])
    for f in ("status", "winding", "eg_status", "at_infinity", "singular"):
        out[f] = np.concatenate([p[f] for p in per]).astype(np.int8)
    for f in ("naccepted", "nrejected", "solution_index", "singular_index"):
        out[f] = np.concatenate([p[f] for p in per]).astype(np.int32)
    out["counts"] = np.stack([p["counts"] for p in per])
    mult = [p["multiplicities"] for p in per]
    out["mult_offsets"] = np.concatenate([[0], np.cumsum([len(m) for m in mult])])
    out["multiplicities"] = (np.concatenate(mult) if sum(len(m) for m in mult)
                             else np.zeros(0, np.int32)).astype(np.int32)
    out["digest"] = np.stack([p["digest"] for p in per])
    np.savez_compressed(os.path.join(HERE, "serving_stream.npz"), **out)


def c3_tasks():
    tasks = [("c3_solve", "c3"), ("c3_solve", "c3v")]
    for name in ("c3", "c3v"):
        for first in range(0, 5040, C3_CHUNK):
            tasks.append(("c3_endpoints", name, first, C3_CHUNK))
    return tasks


def c4_tasks():
    from polyhomocont_amd import workloads
    sizes = [(workloads.random_sparse(104 + s).bezout, s) for s in range(512)]
    return [("c4", s) for _, s in sorted(sizes, reverse=True)]


def write_c3(res):
    out = {}
    for name in ("c3", "c3v"):
        parts = [res[_key(("c3_endpoints", name, f, C3_CHUNK))]
                 for f in range(0, 5040, C3_CHUNK)]
        for f in ("status", "winding", "eg_status"):
            out[f"{name}/{f}"] = np.concatenate([p[f] for p in parts]).astype(np.int8)
        for f in ("naccepted", "nrejected"):
            out[f"{name}/{f}"] = np.concatenate([p[f] for p in parts]).astype(np.int32)
        for f in ("t", "eg_error", "z"):
            out[f"{name}/{f}"] = np.concatenate([p[f] for p in parts])
        sv = res[_key(("c3_solve", name))]
        out[f"{name}/counts"] = sv["counts"]
        out[f"{name}/multiplicities"] = sv["multiplicities"]
        for f in ("solution_index", "singular_index"):
            out[f"{name}/{f}"] = sv[f].astype(np.int32)
        for f in ("at_infinity", "singular"):
            out[f"{name}/{f}"] = sv[f].astype(np.int8)
    out["count_keys"] = np.array(COUNT_KEYS)
    np.savez_compressed(os.path.join(HERE, "c3_full.npz"), **out)


def write_c4(res):
    S = 512
    per = [res[_key(("c4", s))] for s in range(S)]
    out = {"count_keys": np.array(COUNT_KEYS)}
    out["npaths"] = np.array([len(p["status"]) for p in per], np.int64)
    out["offsets"] = np.concatenate([[0], np.cumsum(out["npaths"])])
    for f in ("status", "winding", "eg_status", "at_infinity", "singular"):
        out[f] = np.concatenate([p[f] for p in per]).astype(np.int8)
    for f in ("naccepted", "nrejected", "solution_index", "singular_index"):
        out[f] = np.concatenate([p[f] for p in per]).astype(np.int32)
    out["counts"] = np.stack([p["counts"] for p in per])
    mult = [p["multiplicities"] for p in per]
    out["mult_offsets"] = np.concatenate([[0], np.cumsum([len(m) for m in mult])])
    out["multiplicities"] = (np.concatenate(mult) if sum(len(m) for m in mult)
                             else np.zeros(0, np.int32)).astype(np.int32)
    out["digest"] = np.stack([p["digest"] for p in per])
    out["wsum"] = np.array([p["wsum"][0] for p in per])
    out["asum"] = np.array([p["asum"][0] for p in per])
    np.savez_compressed(os.path.join(HERE, "c4_batch.npz"), **out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=("basic", "c3", "c4", "c2", "serving",
                                     "all"))
    ap.add_argument("--cache", default="/tmp/phc_golden_cache")
    ap.add_argument("--workers", type=int, default=8)
    a = ap.parse_args()
    if a.what in ("basic", "all"):
        basic()
    if a.what != "basic":
        tasks = []
        if a.what in ("c3", "all"):
            tasks += c3_tasks()
        if a.what in ("c4", "all"):
            tasks += c4_tasks()
        if a.what in ("c2", "all"):
            tasks += c2_tasks()
        if a.what in ("serving", "all"):
            tasks += serving_tasks()
        res = run_tasks(tasks, a.cache, a.workers)
        if a.what in ("c3", "all"):
            write_c3(res)
        if a.what in ("c4", "all"):
            write_c4(res)
        if a.what in ("c2", "all"):
            write_c2(res)
        if a.what in ("serving", "all"):
            write_serving(res)


if __name__ == "__main__":
    main()

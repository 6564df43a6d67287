#!/usr/bin/env python3
"""Benchmark of the MI355X path tracker.

Headline workload (one step): the serving stream -- 512 solve requests,
each drawn at random with equal odds from the random sparse generator (n
3..8, degree 2..4, 5..40 terms per equation) or as a dense system (n 2..7,
degree 2..3, all monomials), kept when they have 8..4096 total-degree paths
(239,777 paths in all; seed 105), complex128 -- served once by the
continuous-batching engine with every request queued at the start. A
request is done when its classified result is on the host: homotopy
constants and start points are drawn, the paths tracked on the GPU, the
endpoints classified, refined and deduplicated on the host (host work runs
on a pool of host threads beside the GPU). `value` is paths/s over the K
timed passes.

Reported beside it:
  plain         the same stream served by the plain per-solve path: `solve`
                (Fortran API: constants, start points, one GPU tracking
                launch, classification) called once per request, in order;
  latency       p50/p99 per-request latency (arrival -> classified result)
                under Poisson arrivals at 70% of the plain path's saturation
                rate, for the engine and for the plain path, and a sweep of
                arrival rates;
  roofline      FP64 work of the engine's dominant kernel (the lanes class
                with the most kernel time): canonical flops (flops.py, per
                path linearization counts from the device) over its kernel
                time, measured with HIP events around each round on the
                class's stream;
  cpu_baseline  the compiled reference `solve` (oracle/_ref, flang -O2) on
                one host core over the first requests of the same stream;
  parity        engine results vs the plain path's (every request) and vs
                the reference's (the cpu_baseline sample);
  c2            the plain tracker on one dense n = 7, degree 3 system
                (2187 paths) with inputs resident in HBM, its roofline and
                CPU baseline;
  batch         the 512 seeded random sparse systems in one launch per lanes
                class.

--gpus N > 1 is not supported: the tracker runs on one GPU.
"""
import argparse
import json
import os
import sys
import threading
import time
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector, dense
UNITS_PER_LAUNCH = 256
LAUNCH_US = 2000
SUBMITS_PER_CALL = 32


def log(msg):
    """Progress on stderr (the JSON result line goes to stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr,
          flush=True)


def load_traffic(name):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary."""
    p = os.path.join(ROOT, "profiles", f"{name}_traffic.json")
    if not os.path.exists(p):
        return None
    try:
        return json.load(open(p)).get("bytes_per_launch")
    except Exception:
        return None


def percentiles(xs):
    import numpy as np
    xs = np.asarray(xs) * 1e3
    return {"p50": round(float(np.percentile(xs, 50)), 3),
            "p90": round(float(np.percentile(xs, 90)), 3),
            "p99": round(float(np.percentile(xs, 99)), 3),
            "max": round(float(xs.max()), 3),
            "mean": round(float(xs.mean()), 3)}


# ------------------------------------------------------------------ serving
def prepare(H, usys, opts):
    """What `solve` does before tracking (src/solve.f90:166-190)."""
    from polyhomocont_amd import abi
    g, b, a, _ = H.homotopy(usys, opts.seed)
    start = H.start_points(usys, opts.seed)
    return (H.pack(usys), abi.HomotopyData(g, b, a), opts, start,
            abi.EndpointArrays(start.shape[0], usys.n + 1))


class Engine:
    """The continuous-batching engine with a host pool for the host halves
    of `solve` (constants and start points before, classification after).
    serve() runs a stream and returns per-request results and latencies."""

    def __init__(self, ctx, H, workers=4, host_threads=4, resident=16384,
                 budget=2 << 30, units=UNITS_PER_LAUNCH, launch_us=LAUNCH_US,
                 submits=SUBMITS_PER_CALL, prep_wait_us=500):
        from polyhomocont_amd import device
        self.H = H
        self.submits = submits
        self.prep_wait_us = prep_wait_us
        self.srv = device.Server(ctx, max_resident_paths=resident,
                                 hbm_budget_bytes=budget,
                                 units_per_launch=units,
                                 launch_us=launch_us)
        self.pool = ThreadPoolExecutor(
            workers, initializer=lambda: H.set_threads(host_threads))

    def close(self):
        self.pool.shutdown()
        self.srv.close()

    def serve(self, stream, opts, arrivals=None, timeout=600.0):
        R = len(stream)
        prep, inflight, classify = {}, {}, {}
        results, lat, cls, recs = [None] * R, [None] * R, [None] * R, [None] * R
        arr = arrivals if arrivals is not None else [0.0] * R
        nxt, done = 0, 0
        t0 = time.perf_counter()
        while done < R:
            now = time.perf_counter() - t0
            if now > timeout:
                raise RuntimeError("serving run did not finish")
            while nxt < R and arr[nxt] <= now:
                prep[nxt] = self.pool.submit(prepare, self.H, stream[nxt], opts)
                nxt += 1
            # a bounded number of submissions between progress calls, so
            # that rounds that end meanwhile are restarted promptly
            ready = [i for i, f in prep.items() if f.done()][:self.submits]
            check = False
            for i in ready:
                req = prep.pop(i).result()
                t, out = self.srv.submit(*req)
                cls[i] = self.srv.ticket_class(t)
                inflight[t] = (i, out, req[0])
                check = check or out.npaths == 0
            if inflight:
                # tickets are polled only after a call that finished some;
                # while requests are still being prepared the call waits
                # only briefly for a round to end, so that they are
                # submitted as they become ready
                check = self.srv.progress(
                    self.prep_wait_us if prep else -1) > 0 or check
                for t in [t for t in inflight if check and self.srv.done(t)]:
                    i, out, packed = inflight.pop(t)
                    self.srv.release(t)
                    recs[i] = (packed, out)
                    classify[self.pool.submit(self.H.classify, stream[i], opts,
                                              out)] = i
            for f in [f for f in classify if f.done()]:
                i = classify.pop(f)
                st, res = f.result()
                if st != 0:
                    raise RuntimeError(f"classification of request {i}: {st}")
                results[i] = res
                lat[i] = time.perf_counter() - t0 - arr[i]
                done += 1
            if not inflight:
                # host work only (or nothing yet): wait a little, at most
                # until the next arrival
                wait = (arr[nxt] - (time.perf_counter() - t0)) if nxt < R else 0
                cap = 0.0005 if (prep or classify) else 0.05
                time.sleep(min(max(wait, 0.0), cap) if wait > 0 or not
                           (prep or classify) else cap)
        return results, lat, cls, recs, time.perf_counter() - t0


def serve_plain(H, stream, opts, arrivals=None):
    """The plain per-solve path: `solve` per request, in arrival order."""
    R = len(stream)
    results, lat = [None] * R, [None] * R
    arr = arrivals if arrivals is not None else [0.0] * R
    t0 = time.perf_counter()
    for i, usys in enumerate(stream):
        now = time.perf_counter() - t0
        if now < arr[i]:
            time.sleep(arr[i] - now)
        st, res = H.solve(usys, opts)
        if st != 0:
            raise RuntimeError(f"solve of request {i}: status {st}")
        results[i] = res
        lat[i] = time.perf_counter() - t0 - arr[i]
    return results, lat, time.perf_counter() - t0


def same_result(a, b):
    import numpy as np
    if a.counts() != b.counts():
        return False
    n = a.summary.npaths
    return all(np.array_equal(a.p[f][:n], b.p[f][:n])
               for f in ("z", "status", "winding", "solution_index",
                         "singular_index"))


def class_kernel(c):
    """Kernel name of an engine class (what rocprofv3 lists)."""
    return "phc::track_kernel<%d, false, %d>" % (int(c["G"]), int(c["mapping"]))


def stats_delta(after, before):
    return [{k: a[k] - (b[k] if k in ("launches", "kernel_ms", "units",
                                       "unit_slots") else 0)
             for k in a} for a, b in zip(after, before)]


def serving_section(ctx, H, args, out):
    import numpy as np
    from polyhomocont_amd import abi, flops, workloads
    opts = abi.default_options()
    stream = workloads.serving_stream(args.serving_requests)
    paths = sum(u.bezout for u in stream)
    log(f"serving stream: {len(stream)} requests, {paths} paths")
    eng = Engine(ctx, H, workers=args.host_workers,
                 resident=args.resident_slots, units=args.units_per_launch,
                 launch_us=args.launch_us,
                 submits=args.submits_per_call,
                 prep_wait_us=args.prep_wait_us)
    for w in range(args.warmup):
        eng.serve(stream, opts)
        log(f"warmup pass {w + 1}/{args.warmup}")
    import torch
    torch.cuda.synchronize()
    before = eng.srv.stats()
    t0 = time.perf_counter()
    for k in range(args.steps):
        res, lat, cls, recs, _ = eng.serve(stream, opts)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    after = eng.srv.stats()
    log(f"timed {args.steps} passes: {wall / args.steps:.3f} s/pass")
    value = paths * args.steps / wall
    # roofline of the dominant kernel over the timed passes: classes that
    # run the same kernel template are one kernel to a profiler, so their
    # launches, kernel time and flops are summed by kernel name
    delta = stats_delta(after, before)
    fl = np.zeros(len(delta))
    for i, (packed, rec) in enumerate(recs):
        fl[cls[i]] += flops.path_flops(packed, rec.nlin[:rec.npaths],
                                       rec.nbe[:rec.npaths]).sum()
    kern = {}
    for i, x in enumerate(delta):
        k = kern.setdefault(class_kernel(after[i]),
                            {"kernel_ms": 0.0, "launches": 0, "flops": 0.0})
        k["kernel_ms"] += x["kernel_ms"]
        k["launches"] += int(x["launches"])
        k["flops"] += fl[i] * args.steps
    dom_name = max(kern, key=lambda k: kern[k]["kernel_ms"])
    d = kern[dom_name]
    class_flops = d["flops"]
    achieved = class_flops / (d["kernel_ms"] / 1e3) / 1e12
    per_launch_ms = d["kernel_ms"] / max(d["launches"], 1)
    out["value"] = round(value, 3)
    out["ms_per_step"] = round(wall / args.steps * 1e3, 4)
    out["config"] = {
        "workload": f"serving stream: {len(stream)} requests drawn at random, "
                    "half random sparse (n 3..8, degree 2..4, 5..40 terms), "
                    "half dense (n 2..7, degree 2..3), 8..4096 paths each, "
                    f"{paths} paths; continuous-batching engine, host "
                    "classification included",
        "requests": len(stream), "paths": paths,
        "units_per_launch": args.units_per_launch,
        "launch_us": args.launch_us,
        "resident_path_slots": args.resident_slots,
        "host_pool_workers": args.host_workers}
    out["roofline"] = {
        "bound": "fp64-valu",
        "achieved": round(achieved, 4), "peak": PEAK_FP64_TFLOPS,
        "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 6),
        "traffic": load_traffic("serving"),
        "kernel": dom_name,
        "flops_per_launch": int(class_flops / max(d["launches"], 1)),
        "avg_launch_ms": round(per_launch_ms, 4),
        "launches": int(d["launches"]),
        "note": "FP64 vector ALU work (no MFMA); canonical flops of the "
                "paths tracked by the engine classes that run this kernel, "
                "over the summed durations of their launches (launches of "
                "different classes overlap in time)"}
    out["serving_classes"] = [
        {"kernel": class_kernel(x), "term_values": int(x["vmax"]),
         "kernel_ms": round(x["kernel_ms"], 2),
         "launches": int(x["launches"]), "units": int(x["units"]),
         "tflops": round(fl[i] * args.steps / max(x["kernel_ms"], 1e-9) / 1e9, 4),
         "slot_time_utilisation": round(
             x["units"] / max(x["slots"] * x["launches"] * args.units_per_launch,
                              1), 4),
         "lane_occupancy": round(x["units"] / x["unit_slots"], 4)
         if x["unit_slots"] else None}
        for i, x in enumerate(delta)]
    serving = {"engine_paths_per_s": round(value, 1)}
    if args.no_plain:
        eng.close()
        return stream, res, serving
    # plain per-solve path, saturation
    plain_res, _, plain_s = serve_plain(H, stream, opts)
    log(f"plain per-solve path: {plain_s:.2f} s")
    identical = sum(same_result(a, b) for a, b in zip(res, plain_res))
    serving.update({
        "plain_paths_per_s": round(paths / plain_s, 1),
        "speedup_vs_plain": round(value / (paths / plain_s), 3),
        "results_identical_to_plain": f"{identical}/{len(stream)}",
    })
    # latency at 70% of the plain saturation rate, then a sweep
    plain_rate = len(stream) / plain_s
    rng = np.random.default_rng(105)
    gaps = rng.exponential(1.0, len(stream))

    def arrivals(rate):
        return list(np.cumsum(gaps) / rate)

    lam = 0.7 * plain_rate
    _, lat_e, _, _, _ = eng.serve(stream, opts, arrivals(lam))
    _, lat_p, _ = serve_plain(H, stream, opts, arrivals(lam))
    serving["latency_at_fixed_rate"] = {
        "requests_per_s": round(lam, 3), "fraction_of_plain_saturation": 0.7,
        "engine_ms": percentiles(lat_e), "plain_ms": percentiles(lat_p)}
    log(f"latency at {lam:.1f} req/s: engine p50/p99 "
        f"{serving['latency_at_fixed_rate']['engine_ms']['p50']}/"
        f"{serving['latency_at_fixed_rate']['engine_ms']['p99']} ms, plain "
        f"{serving['latency_at_fixed_rate']['plain_ms']['p50']}/"
        f"{serving['latency_at_fixed_rate']['plain_ms']['p99']} ms")
    sweep = []
    if not args.no_sweep:
        # the plain path saturates at 1.0; the engine is swept beyond it
        for f in (0.5, 1.0, 2.0, 4.0, 5.0, 6.0):
            rate = f * plain_rate
            _, le, _, _, wall_f = eng.serve(stream, opts, arrivals(rate))
            # the stream's paths over its span, first arrival to last
            # result: the engine keeps up while this tracks the offered rate
            row = {"fraction_of_plain_saturation": f,
                   "requests_per_s": round(rate, 3),
                   "offered_paths_per_s": round(f * paths / plain_s, 1),
                   "served_paths_per_s": round(paths / wall_f, 1),
                   "engine_ms": percentiles(le)}
            sweep.append(row)
            log(f"sweep {f}: engine p50/p99 {row['engine_ms']['p50']}/"
                f"{row['engine_ms']['p99']} ms")
    serving["latency_sweep"] = sweep
    serving["latency_includes"] = (
        "arrival -> classified result on the host: constants and start "
        "points, tracking, classification, refinement, deduplication")
    eng.close()
    return stream, res, serving


def cpu_stream_baseline(stream, engine_results, budget_s):
    """The compiled reference `solve` on one host core over the first
    requests of the stream, within about budget_s seconds; the results are
    compared with the engine's."""
    import numpy as np
    import oracle
    from polyhomocont_amd import abi
    ref = oracle.ref()
    kind = "reference"
    if ref is None:
        ref = oracle.port()
        kind = "port"
    opts = abi.default_options()
    t0 = time.perf_counter()
    done, paths, same, maxerr = 0, 0, 0, 0.0
    for i, usys in enumerate(stream):
        if time.perf_counter() - t0 > budget_s and done:
            break
        st, want = ref.solve(usys, opts)
        got = engine_results[i]
        n = want.summary.npaths
        same += int(got.counts() == want.counts() and
                    np.array_equal(got.p["status"][:n], want.p["status"][:n]))
        s = want.summary.nsolutions
        if s:
            scale = np.maximum(np.abs(want.solutions[:s]).max(axis=1), 1.0)
            err = (np.abs(got.solutions[:s] - want.solutions[:s]).max(axis=1)
                   / scale).max()
            maxerr = max(maxerr, float(err))
        done += 1
        paths += usys.bezout
    dt = time.perf_counter() - t0
    return ({"value": round(paths / dt, 2), "unit": "paths/s", "cores": 1,
             "kind": kind,
             "sample": f"reference solve of requests 1..{done} of the same "
                       f"stream ({paths} paths), serial, 1 host core, "
                       f"{dt:.1f} s"},
            {"requests_compared": done, "counts_and_status_equal": same,
             "max_rel_solution_err": maxerr})


def _ref_solve_request(args):
    """One request of the stream solved by the reference (worker process)."""
    count, i = args
    import oracle
    from polyhomocont_amd import abi, workloads
    ref = oracle.ref() or oracle.port()
    usys = workloads.serving_stream(count)[i]
    st, _ = ref.solve(usys, abi.default_options())
    return usys.bezout


def cpu_stream_all_cores(count, first_requests, budget_s):
    """The reference `solve` over the first requests of the stream, one
    request per task, on every host core this process may use (at most
    16). Worker start-up is excluded: the pool is warmed before timing."""
    import multiprocessing as mp
    import oracle
    kind = "reference" if oracle.ref() is not None else "port"
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    m = max(cores, first_requests * cores)
    with mp.get_context("spawn").Pool(cores) as pool:
        pool.map(_ref_solve_request, [(count, 0)] * cores, chunksize=1)
        t0 = time.perf_counter()
        paths = sum(pool.imap_unordered(_ref_solve_request,
                                        [(count, i) for i in range(min(m, count))]))
        dt = time.perf_counter() - t0
    return {"value": round(paths / dt, 2), "unit": "paths/s", "cores": cores,
            "kind": kind,
            "sample": f"reference solve of requests 1..{min(m, count)} of the "
                      f"same stream ({paths} paths) over {cores} worker "
                      f"processes ({dt:.1f} s)"}


# ---------------------------------------------------------------------- C2
def c2_section(ctx, H, args):
    import numpy as np
    from polyhomocont_amd import abi, device, flops, workloads
    usys = workloads.config_system("c2")
    opts = abi.default_options()
    gamma, b, patch, npaths = H.homotopy(usys, opts.seed)
    start = H.start_points(usys, opts.seed)
    packed = H.pack(usys)
    hom = abi.HomotopyData(gamma, b, patch)
    tune = ctx.autotune(packed, hom, opts, start)
    st = device.Staged(ctx, packed, hom, opts, start)
    for _ in range(2):
        st.run()
        st.sync()
    steps = 10
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        st.run()
        kernel_ms.append(st.elapsed_ms())
    st.sync()
    wall = time.perf_counter() - t0
    rec = st.fetch()
    info = st.launch_info()
    c = ctx.counters()
    st.close()
    fl = int(flops.path_flops(packed, rec.nlin[:npaths], rec.nbe[:npaths]).sum())
    avg_s = float(np.mean(kernel_ms)) / 1e3
    achieved = fl / avg_s / 1e12
    traffic = load_traffic("c2")
    out = {
        "workload": "dense n=7, degree 3, 120 monomials/equation, 2187 paths, "
                    "plain tracker, inputs resident in HBM",
        "paths_per_s": round(npaths * steps / wall, 1),
        "ms_per_solve": round(wall / steps * 1e3, 3),
        "launch": info, "autotune": tune,
        "roofline": {"bound": "fp64-valu", "achieved": round(achieved, 4),
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64_TFLOPS, 6),
                     "traffic": traffic,
                     "kernel": "phc::track_kernel<%d, %s, %d>" % (
                         info["group"], "true" if info["staged"] else "false",
                         {"row": 0, "split": 1, "grid": 2}[info["mapping"]]),
                     "flops_per_launch": fl,
                     "avg_launch_ms": round(avg_s * 1e3, 4)},
        "lane_occupancy": round(c["occupancy"], 4),
        "slot_time_utilisation": round(c["slot_utilisation"], 4),
        "resident_path_slots": c["slots"],
        "longest_slot_units": c["max_slot_units"],
        "units": c["units"],
    }
    if not args.no_cpu:
        import oracle
        ref = oracle.ref() or oracle.port()
        t0 = time.perf_counter()
        ref.endpoints(usys, opts, 0, 4)
        per = max((time.perf_counter() - t0) / 4, 1e-4)
        cnt = int(min(npaths, max(8, args.cpu_seconds / 2 / per)))
        t0 = time.perf_counter()
        want = ref.endpoints(usys, opts, 0, cnt)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(cnt / dt, 2), "unit": "paths/s",
                               "cores": 1,
                               "sample": f"paths 1..{cnt}, serial ({dt:.1f} s)"}
        out["parity"] = {
            "paths_compared": cnt,
            "bit_identical": bool(all(np.array_equal(getattr(rec, f)[:cnt],
                                                     getattr(want, f)[:cnt])
                                      for f in ("z", "status", "winding",
                                                "eg_status", "naccepted",
                                                "nrejected")))}
    return out


# ---------------------------------------------------------------------- C3
def c3_section(ctx, H, args):
    """The end game stress system (n = 7, 5040 paths, most at infinity,
    winding numbers > 1): one plain per-solve launch with inputs resident in
    HBM, its roofline, and the longest path's share."""
    import numpy as np
    from polyhomocont_amd import abi, device, flops, workloads
    usys = workloads.config_system("c3")
    opts = abi.default_options()
    gamma, b, patch, npaths = H.homotopy(usys, opts.seed)
    start = H.start_points(usys, opts.seed)
    packed = H.pack(usys)
    st = device.Staged(ctx, packed, abi.HomotopyData(gamma, b, patch), opts,
                       start)
    t0 = time.perf_counter()
    st.run()
    kernel_ms = st.elapsed_ms()
    wall = time.perf_counter() - t0
    rec = st.fetch()
    info = st.launch_info()
    c = ctx.counters()
    st.close()
    units = rec.nlin[:npaths] + rec.nbe[:npaths]
    fl = int(flops.path_flops(packed, rec.nlin[:npaths], rec.nbe[:npaths]).sum())
    achieved = fl / (kernel_ms / 1e3) / 1e12
    longest = int(np.argmax(units))
    return {
        "workload": "cyclic-like n=7 (5040 paths, most at infinity), plain "
                    "tracker, inputs resident in HBM",
        "seconds": round(wall, 3), "paths_per_s": round(npaths / wall, 2),
        "launch": info,
        "roofline": {"bound": "fp64-valu", "achieved": round(achieved, 5),
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64_TFLOPS, 7),
                     "kernel": "phc::track_kernel<%d, %s, %d>" % (
                         info["group"], "true" if info["staged"] else "false",
                         {"row": 0, "split": 1, "grid": 2}[info["mapping"]]),
                     "flops_per_launch": fl,
                     "avg_launch_ms": round(kernel_ms, 3)},
        "units": int(units.sum()),
        "longest_path": {"path": longest, "units": int(units[longest]),
                         "us_per_unit": round(kernel_ms * 1e3 /
                                              max(int(units[longest]), 1), 3)},
        "slot_time_utilisation": round(c["slot_utilisation"], 5),
    }


# ------------------------------------------------------------------- batch
def batch_section(ctx, H, count):
    from polyhomocont_amd import abi, workloads
    items = []
    for usys in workloads.batch_systems(count):
        opts = abi.default_options()
        g, b, a, ns = H.homotopy(usys, opts.seed)
        items.append((H.pack(usys), abi.HomotopyData(g, b, a), opts,
                      H.start_points(usys, opts.seed)))
    paths = sum(len(it[3]) for it in items)
    log(f"batch: {count} systems, {paths} paths")
    t0 = time.perf_counter()
    tuned = ctx.autotune_batch(items)
    tune_s = time.perf_counter() - t0
    log(f"batch autotune ({tune_s:.2f} s): {tuned}")
    t0 = time.perf_counter()
    outs = ctx.track_batch(items)
    dt = time.perf_counter() - t0
    c = ctx.counters()
    plan = ctx.last_plan()
    ctx.clear_tuning()
    log(f"batch: {dt:.2f} s")
    return {"workload": f"{count} random sparse systems (n 3..8, degree 2..4, "
                        "5..40 terms per equation), one launch per lanes class",
            "systems": count, "paths": paths,
            "paths_per_s": round(paths / dt, 1), "seconds": round(dt, 3),
            "units": c["units"],
            "lane_occupancy_after_divergence": round(c["occupancy"], 4),
            "slot_time_utilisation": round(c["slot_utilisation"], 4),
            "longest_slot_units": c["max_slot_units"],
            "autotuned": tuned, "autotune_seconds": round(tune_s, 2),
            "launch_plan": plan,
            "includes": "host packing and autotuning excluded; upload, "
                        "kernels, download included"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3,
                    help="timed passes over the serving stream")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--serving-requests", type=int, default=512)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--batch-systems", type=int, default=512,
                    help="systems of the batched multi-system launch (0: skip)")
    ap.add_argument("--units-per-launch", type=int, default=UNITS_PER_LAUNCH,
                    help="serving engine: linearizations per path slot and "
                         "launch")
    ap.add_argument("--launch-us", type=int, default=LAUNCH_US,
                    help="serving engine: time budget of a launch in "
                         "microseconds (0: the unit budget alone)")
    ap.add_argument("--submits-per-call", type=int, default=SUBMITS_PER_CALL,
                    help="serving engine: requests submitted between two "
                         "progress calls")
    ap.add_argument("--prep-wait-us", type=int, default=500,
                    help="serving engine: longest wait of a progress call "
                         "while requests are being prepared (-1: wait for "
                         "a round to end)")
    ap.add_argument("--resident-slots", type=int, default=16384,
                    help="serving engine: resident path slots")
    ap.add_argument("--host-workers", type=int, default=4,
                    help="serving engine: host pool workers (each with 4 "
                         "threads) for constants, start points and "
                         "classification")
    ap.add_argument("--no-plain", action="store_true",
                    help="serving: engine passes only (no plain per-solve "
                         "path, no latency runs); for profiling the engine")
    ap.add_argument("--only", choices=("serving", "c2", "batch", "c3",
                                       "hbm-copy"),
                    help="run one section alone (c3, the 5040-path end game "
                         "stress solve, runs only when named here)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported; the path "
              "tracker runs on one GPU (paths would shard by index range "
              "across GPUs, with the records gathered to one host in path "
              "order for classification)", file=sys.stderr)
        sys.exit(2)

    # torch first: it loads its own HIP runtime, which the library's must
    # then share (a runtime loaded before torch's leaves torch without GPUs)
    import torch
    torch.cuda.init()
    if args.only == "hbm-copy":
        # calibration of the HBM byte counters: a plain device copy of
        # 256 MiB (read 256 MiB, write 256 MiB per launch), 5 launches
        n = 256 << 20
        src = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"hbm_copy_bytes_read": n, "hbm_copy_bytes_written": n,
                          "launches": 5}), flush=True)
        return
    from polyhomocont_amd import device, host
    H = host.library()
    ctx = device.Context(0)
    out = {
        "metric": "paths/sec at endpoint err <=1e-10 vs CPU ref; p50/p99 "
                  "solve latency @ fixed QPS",
        "value": None, "unit": "paths/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128 (fp64)",
        "data": "synthetic (seeded complex-Gaussian coefficients)",
    }
    want = (lambda name: args.only == name) if args.only else (
        lambda name: name != "c3")
    if want("serving"):
        stream, results, serving = serving_section(ctx, H, args, out)
        out["serving"] = serving
        if not args.no_cpu:
            log("cpu baseline (reference solve on the stream)")
            cb, par = cpu_stream_baseline(stream, results, args.cpu_seconds)
            out["cpu_baseline"] = cb
            out["cpu_baseline_all_cores"] = cpu_stream_all_cores(
                len(stream), par["requests_compared"], args.cpu_seconds)
            out["parity"] = par
            out["gpu_vs_cpu_core"] = round(out["value"] / cb["value"], 1)
    if want("c2") and not args.no_c2:
        log("c2 plain tracker")
        out["c2"] = c2_section(ctx, H, args)
    if want("batch") and args.batch_systems > 0:
        out["batch"] = batch_section(ctx, H, args.batch_systems)
    if want("c3"):
        log("c3 end game stress solve")
        out["c3"] = c3_section(ctx, H, args)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

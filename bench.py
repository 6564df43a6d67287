#!/usr/bin/env python3
"""Benchmark of the MI355X path tracker.

Workload (one step): one solve of C2 -- a single random dense system,
n = 7, every equation of total degree 3 with all 120 monomials of degree
<= 3, complex-Gaussian coefficients (seed 102), complex128 -- i.e. all
2187 total-degree paths tracked from t = 1 through the Cauchy end game by
the plain per-solve GPU tracker. Inputs (packed system, constants, start
points) are resident in HBM before the timed region; the timed region is K
back-to-back tracking runs.

Reported beside the throughput:
  roofline      FP64 work of the tracking kernel: canonical flops per
                launch (polyhomocont_amd/flops.py, per-path linearization
                counts from the device) over the kernel's average duration
                measured with HIP events on the library's stream;
  cpu_baseline  the compiled reference (oracle/_ref, flang -O2) on one host
                core over a bounded sample of the same paths;
  parity        endpoints of that sample vs the reference (max relative
                difference; classification fields compared exactly).

--gpus N > 1 is not supported: the tracker runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 (vector and matrix), dense
PEAK_HBM_GBS = 8000.0


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


def cpu_baseline(usys, opts, npaths, budget_s):
    import numpy as np
    import oracle
    ref = oracle.ref()
    kind = "reference"
    if ref is None:
        ref = oracle.port()
        kind = "port"
    # calibrate on a few paths, then time a sample that fills the budget
    t0 = time.perf_counter()
    ref.endpoints(usys, opts, 0, 4)
    per = max((time.perf_counter() - t0) / 4, 1e-4)
    count = int(min(npaths, max(8, budget_s / per)))
    t0 = time.perf_counter()
    rec = ref.endpoints(usys, opts, 0, count)
    dt = time.perf_counter() - t0
    return dict(value=count / dt, unit="paths/s", cores=1, kind=kind,
                sample=f"paths 1..{count} of the same solve, serial "
                       f"(1 host core, {dt:.1f} s)"), rec, count


def _cpu_chunk(args):
    """One slice of the all-core CPU baseline (runs in a worker process)."""
    name, first, count = args
    import oracle
    from polyhomocont_amd import abi, workloads
    ref = oracle.ref() or oracle.port()
    ref.endpoints(workloads.config_system(name), abi.default_options(),
                  first, count)
    return count


def cpu_all_cores(name, npaths, per_path_s, budget_s):
    """The same reference tracker on every host core this process may use
    (at most 16), paths handed out in chunks of 8 to worker processes.
    Worker start-up is excluded: the pool is warmed before timing."""
    import multiprocessing as mp
    import oracle
    kind = "reference" if oracle.ref() is not None else "port"
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    count = int(min(npaths, max(8 * cores, cores * budget_s / per_path_s)))
    chunks = [(name, f, min(8, count - f)) for f in range(0, count, 8)]
    with mp.get_context("spawn").Pool(cores) as pool:
        pool.map(_cpu_chunk, [(name, 0, 1)] * cores, chunksize=1)
        t0 = time.perf_counter()
        done = sum(pool.imap_unordered(_cpu_chunk, chunks))
        dt = time.perf_counter() - t0
    return dict(value=done / dt, unit="paths/s", cores=cores, kind=kind,
                sample=f"paths 1..{count} of the same solve over {cores} "
                       f"worker processes ({dt:.1f} s)")


def percentiles(xs):
    import numpy as np
    xs = np.asarray(xs) * 1e3
    return {"p50": round(float(np.percentile(xs, 50)), 3),
            "p99": round(float(np.percentile(xs, 99)), 3),
            "mean": round(float(xs.mean()), 3)}


def serving_benchmark(ctx, host_lib, count, units_per_launch=64):
    """The mixed request stream served by (a) the plain per-solve tracker,
    one request at a time, and (b) the continuous-batching engine:
    saturation throughput with every request queued at t = 0, then
    per-request latency (arrival -> endpoint records on the host) under
    Poisson arrivals at 70% of the plain tracker's saturation rate."""
    import numpy as np
    from polyhomocont_amd import abi, device, workloads

    systems = workloads.serving_stream(count)
    reqs = []
    for usys in systems:
        opts = abi.default_options()
        g, b, a, ns = host_lib.homotopy(usys, opts.seed)
        reqs.append((host_lib.pack(usys), abi.HomotopyData(g, b, a), opts,
                     host_lib.start_points(usys, opts.seed)))
    paths = sum(len(r[3]) for r in reqs)
    log(f"serving: {len(reqs)} requests, {paths} paths prepared")
    # warm both paths (code objects, device buffers)
    ctx.track_paths(*reqs[0])
    srv = device.Server(ctx, max_resident_paths=16384,
                        hbm_budget_bytes=2 << 30,
                        units_per_launch=units_per_launch)
    t, _ = srv.submit(*reqs[0])
    srv.wait(t)

    t0 = time.perf_counter()
    plain_out = [ctx.track_paths(*r) for r in reqs]
    plain_s = time.perf_counter() - t0
    log(f"serving: plain tracker saturation {plain_s:.2f} s")

    t0 = time.perf_counter()
    tickets = [srv.submit(*r)[0] for r in reqs]
    for t in tickets:
        srv.wait(t)
    serve_s = time.perf_counter() - t0
    log(f"serving: engine saturation {serve_s:.2f} s")
    same = all(np.array_equal(srv._keep[t][4].z, p.z) and
               np.array_equal(srv._keep[t][4].status, p.status)
               for t, p in zip(tickets, plain_out))
    for t in tickets:
        srv.release(t)

    lam = 0.7 * len(reqs) / plain_s
    rng = np.random.default_rng(105)
    arrivals = np.cumsum(rng.exponential(1.0 / lam, len(reqs)))

    # plain tracker: one FIFO server
    lat_plain = []
    t0 = time.perf_counter()
    for r, a in zip(reqs, arrivals):
        now = time.perf_counter() - t0
        if now < a:
            time.sleep(a - now)
        ctx.track_paths(*r)
        lat_plain.append(time.perf_counter() - t0 - a)
    log("serving: plain tracker latency run done")

    # serving engine: submit on arrival, progress continuously
    lat_serve = [None] * len(reqs)
    pending = {}
    nxt = 0
    t0 = time.perf_counter()
    while nxt < len(reqs) or pending:
        now = time.perf_counter() - t0
        if now > arrivals[-1] + 120.0:
            raise RuntimeError("serving latency run did not finish")
        while nxt < len(reqs) and arrivals[nxt] <= now:
            pending[srv.submit(*reqs[nxt])[0]] = nxt
            nxt += 1
        if not pending:
            time.sleep(max(0.0, arrivals[nxt] - (time.perf_counter() - t0)))
            continue
        srv.progress()
        for t in [t for t in pending if srv.done(t)]:
            i = pending.pop(t)
            lat_serve[i] = time.perf_counter() - t0 - arrivals[i]
            srv.release(t)
    srv.close()
    log("serving: engine latency run done")
    fe_s, fe_same, fe_bytes = frontend_saturation(
        ctx, host_lib, systems, plain_out, units_per_launch)
    log(f"serving: front end saturation {fe_s:.2f} s")
    return {
        "workload": f"{len(reqs)} requests, alternately random sparse "
                    "(n 3..8, degree 2..4, 5..40 terms) and dense (n 2..7, "
                    "degree 2..3), 8..4096 paths each",
        "paths": paths,
        "units_per_launch": units_per_launch,
        "plain_paths_per_s": round(paths / plain_s, 1),
        "serving_paths_per_s": round(paths / serve_s, 1),
        "serving_speedup": round(plain_s / serve_s, 2),
        "records_identical": bool(same),
        "qps": round(lam, 3),
        "latency_ms": {"serving": percentiles(lat_serve),
                       "plain": percentiles(lat_plain)},
        "front_end": {
            "paths_per_s": round(paths / fe_s, 1),
            "records_identical": bool(fe_same),
            "request_bytes": fe_bytes,
            "includes": "request frames over TCP on 127.0.0.1, decoding, "
                        "building each system and drawing its start points "
                        "on the front end's thread, the engine, reply "
                        "frames"},
    }


def frontend_saturation(ctx, host_lib, systems, plain_out, units_per_launch):
    """All requests sent at once as frames to the front end (own thread,
    own engine) by one client; returns (seconds until the last reply,
    records equal to the plain tracker's, request bytes)."""
    import threading
    import numpy as np
    from polyhomocont_amd import abi, device, frontend, wire

    frames = [wire.encode_request(u, abi.default_options(), i)
              for i, u in enumerate(systems)]
    srv = device.Server(ctx, max_resident_paths=16384,
                        hbm_budget_bytes=2 << 30,
                        units_per_launch=units_per_launch)
    fe = frontend.FrontEnd(srv, host_lib)
    stop = threading.Event()
    th = threading.Thread(target=fe.serve, args=(stop,), daemon=True)
    th.start()
    got = {}
    try:
        cl = frontend.Client(fe.address)
        t0 = time.perf_counter()
        for f in frames:
            wire.send_frame(cl.sock, f)
        for _ in frames:
            rid, rec, err, msg = cl.receive()
            if err:
                raise RuntimeError(f"front end rejected request {rid}: {msg}")
            got[rid] = rec
        dt = time.perf_counter() - t0
        cl.close()
    finally:
        stop.set()
        th.join(timeout=60)
        fe.close()
        srv.close()
    same = all(np.array_equal(got[i].z[:got[i].npaths], p.z[:got[i].npaths])
               and np.array_equal(got[i].status[:got[i].npaths],
                                  p.status[:got[i].npaths])
               for i, p in enumerate(plain_out))
    return dt, same, sum(len(f) for f in frames)


def batch_benchmark(ctx, host_lib, count):
    """The batched multi-system launch: the first `count` of the seeded
    random sparse systems (n 3..8, degree 2..4, 5..40 terms per equation),
    all their paths in one persistent launch per lanes class."""
    from polyhomocont_amd import abi, workloads
    items = []
    for usys in workloads.batch_systems(count):
        opts = abi.default_options()
        g, b, a, ns = host_lib.homotopy(usys, opts.seed)
        items.append((host_lib.pack(usys), abi.HomotopyData(g, b, a), opts,
                      host_lib.start_points(usys, opts.seed)))
    paths = sum(len(it[3]) for it in items)
    log(f"batch: {count} systems, {paths} paths")
    t0 = time.perf_counter()
    outs = ctx.track_batch(items)
    dt = time.perf_counter() - t0
    occ = ctx.counters()["occupancy"]
    ok = sum(int((o.status[:len(it[3])] == 0).sum()) for o, it in zip(outs, items))
    log(f"batch: {dt:.2f} s")
    return {"workload": f"{count} random sparse systems (n 3..8, degree 2..4, "
                        "5..40 terms per equation), one launch per lanes class",
            "systems": count, "paths": paths,
            "paths_per_s": round(paths / dt, 1), "seconds": round(dt, 3),
            "paths_success": ok,
            "lane_occupancy_after_divergence": round(occ, 4),
            "includes": "host packing excluded; upload, kernels, download included"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=48)
    ap.add_argument("--serving-units", type=int, default=64,
                    help="linearizations per path slot per engine launch")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--batch-systems", type=int, default=512,
                    help="systems of the batched multi-system launch (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported; the path "
              "tracker runs on one GPU (paths would shard by index range "
              "across GPUs, with the records gathered to one host in path "
              "order for classification)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from polyhomocont_amd import abi, device, flops, host, workloads

    usys = workloads.config_system(args.workload)
    opts = abi.default_options()
    H = host.library()
    gamma, b, patch, npaths = H.homotopy(usys, opts.seed)
    start = H.start_points(usys, opts.seed)
    packed = H.pack(usys)
    hom = abi.HomotopyData(gamma, b, patch)

    ctx = device.Context(0)
    tune = None
    if not args.no_autotune:
        tune = ctx.autotune(packed, hom, opts, start)
        log(f"autotune: {tune}")
    st = device.Staged(ctx, packed, hom, opts, start)
    log(f"staged {args.workload}: {npaths} paths, launch {st.launch_info()}")
    for _ in range(args.warmup):
        st.run()
        st.sync()

    kernel_ms = []
    torch.cuda.synchronize()
    st.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st.run()
        kernel_ms.append(st.elapsed_ms())
    st.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    log(f"timed {args.steps} steps: {(t1 - t0) / args.steps * 1e3:.2f} ms/step")
    wall = t1 - t0
    rec = st.fetch()
    info = st.launch_info()
    occ = ctx.counters()["occupancy"]

    value = npaths * args.steps / wall
    flops_launch = int(flops.path_flops(packed, rec.nlin[:npaths],
                                        rec.nbe[:npaths]).sum())
    avg_s = float(np.mean(kernel_ms)) / 1e3
    achieved = flops_launch / avg_s / 1e12
    traffic = load_traffic(args.workload)

    out = {
        "metric": "paths/sec at endpoint err <=1e-10 vs CPU ref; p50/p99 "
                  "solve latency @ fixed QPS",
        "value": round(value, 3),
        "unit": "paths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128 (fp64)",
        "data": "synthetic (seeded complex-Gaussian coefficients)",
        "config": {"workload": f"{args.workload}: dense n=7, degree 3, 120 "
                               "monomials/equation, 2187 paths, plain per-solve "
                               "GPU tracker" if args.workload == "c2" else args.workload,
                   "npaths": npaths, "lanes_per_path": info["lanes_per_path"],
                   "mapping": info["mapping"],
                   "autotuned": tune is not None,
                   "workgroups": info["workgroups"],
                   "threads_per_workgroup": info["threads"]},
        "roofline": {
            "bound": "mfma",
            "achieved": round(achieved, 4),
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP64_TFLOPS, 6),
            "traffic": traffic,
            "kernel": "phc::track_kernel<%d, %s, %d> (%s mapping)" % (
                info["group"], "true" if info["staged"] else "false",
                {"row": 0, "split": 1, "grid": 2}[info["mapping"]],
                info["mapping"]),
            "flops_per_launch": flops_launch,
            "avg_launch_ms": round(avg_s * 1e3, 4),
            "hbm_gbps": round(traffic / avg_s / 1e9, 3) if traffic else None,
            "note": "FP64 vector ALU work (no MFMA); peak is the MI355X "
                    "FP64 figure",
        },
        "path_divergence": {
            "paths_per_wavefront": 64 // info["lanes_per_path"],
            "lane_occupancy": round(occ, 4),
            "loss": round(1.0 - occ, 4),
            "definition": "1 - units run / (passes of each wavefront's "
                          "tracking loop that ran a unit x paths per "
                          "wavefront), device counters of the last timed launch"},
        "work": {"linearizations": int(rec.nlin[:npaths].sum()),
                 "backward_errors": int(rec.nbe[:npaths].sum()),
                 "steps_accepted": int(rec.naccepted[:npaths].sum()),
                 "steps_rejected": int(rec.nrejected[:npaths].sum())},
    }

    if not args.no_cpu:
        log("cpu baseline")
        cb, want, cnt = cpu_baseline(usys, opts, npaths, args.cpu_seconds)
        out["cpu_baseline"] = cb
        log("cpu baseline, all cores")
        out["cpu_baseline_all_cores"] = cpu_all_cores(
            args.workload, npaths, 1.0 / cb["value"], args.cpu_seconds / 2)
        same = all(np.array_equal(getattr(rec, f)[:cnt], getattr(want, f)[:cnt])
                   for f in ("status", "winding", "eg_status"))
        scale = np.maximum(np.abs(want.z[:cnt]).max(axis=1), 1e-300)
        err = float((np.abs(rec.z[:cnt] - want.z[:cnt]).max(axis=1) / scale).max())
        bit = float((rec.z[:cnt] == want.z[:cnt]).all(axis=1).mean())
        out["parity"] = {"paths_compared": cnt, "max_rel_endpoint_err": err,
                         "classification_equal": bool(same),
                         "bit_identical_fraction": bit}
        out["gpu_vs_cpu_core"] = round(value / cb["value"], 1)
    if args.batch_systems > 0:
        out["batch"] = batch_benchmark(ctx, H, args.batch_systems)
    if not args.no_serving:
        out["serving"] = serving_benchmark(ctx, H, args.serving_requests,
                                           args.serving_units)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

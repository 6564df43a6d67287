"""Serving engine: requests submitted while others run are admitted as
slots and memory free up; every request's endpoint records are the same
alone, in the mix, and from the plain per-solve tracker."""
import numpy as np
import pytest

from polyhomocont_amd import abi, device, workloads as W

pytestmark = pytest.mark.gpu
FIELDS = ("status", "winding", "eg_status", "naccepted", "nrejected", "t",
          "eg_error")


def request(host, usys, seed=1, limit=None):
    opts = abi.default_options(seed=seed)
    g, b, a, ns = host.homotopy(usys, seed)
    cnt = ns if limit is None else min(ns, limit)
    return (host.pack(usys), abi.HomotopyData(g, b, a), opts,
            host.start_points(usys, seed, 0, cnt))


def same(a, b, n):
    for f in FIELDS:
        assert np.array_equal(getattr(a, f)[:n], getattr(b, f)[:n]), f
    assert np.array_equal(a.z[:n], b.z[:n])


def mixed_requests(host):
    reqs = [request(host, W.random_sparse(104 + s)) for s in range(10)]
    reqs += [request(host, W.dense_system(n, d, 700 + n * 10 + d), limit=300)
             for n, d in ((2, 3), (3, 2), (5, 2), (4, 3))]
    reqs += [request(host, W.config_system("c1")),
             request(host, W.ref_test_systems()["mixed"], seed=4)]
    # 16-lane class (n = 8): the run layout of the grid unit
    reqs += [request(host, W.random_sparse(104 + 66), limit=200),
             request(host, W.random_sparse(104 + 68), seed=2, limit=150)]
    return reqs


# (budget, units per launch, launch time budget in us): the defaults; a
# 1 MiB budget with 8-unit launches (admission waits, many resumes); launches
# ended by a 100 us time budget (a slot stops after a time-dependent number
# of units, so paths resume at varying places)
@pytest.mark.parametrize("budget,units,launch_us",
                         [(0, 0, 0), (1 << 20, 8, 0), (0, 4096, 100)])
def test_mix_equals_alone_and_plain(ctx, host, budget, units, launch_us):
    reqs = mixed_requests(host)
    plain = [ctx.track_paths(*r) for r in reqs]
    srv = device.Server(ctx, max_resident_paths=2048, hbm_budget_bytes=budget,
                        units_per_launch=units, launch_us=launch_us)
    print(f"[gpu] server budget={budget} units={units} launch_us={launch_us}",
          flush=True)
    # half up front, the rest while the first half is running
    tickets = [srv.submit(*r) for r in reqs[:8]]
    srv.progress()
    tickets += [srv.submit(*r) for r in reqs[8:]]
    for (t, out), r, p in zip(tickets, reqs, plain):
        got = srv.wait(t)
        same(got, p, len(r[3]))
    # one request alone on a fresh server
    t, out = srv.submit(*reqs[3])
    same(srv.wait(t), plain[3], len(reqs[3][3]))
    srv.close()


def test_server_rejects_oversized_request(ctx, host):
    srv = device.Server(ctx, max_resident_paths=64, hbm_budget_bytes=1 << 16)
    big = request(host, W.config_system("c2"), limit=2000)
    with pytest.raises(device.DeviceError):
        srv.submit(*big)
    srv.close()


def test_released_tickets_leave_nothing_behind(ctx, host):
    """A long-running engine keeps no record of released tickets; a queued
    ticket can be cancelled; an unknown ticket is an error; the mix runs on
    several lanes classes (4-lane grid, 8- and 16-lane grids)."""
    reqs = mixed_requests(host)
    srv = device.Server(ctx, max_resident_paths=2048)
    print("[gpu] server release", flush=True)
    tickets = [srv.submit(*r)[0] for r in reqs]
    extra, _ = srv.submit(*reqs[0])
    srv.release(extra)                      # still queued: cancelled
    for t in tickets:
        srv.wait(t)
        srv.release(t)
    assert srv.tickets() == dict(queued=0, admitted=0, finished=0)
    with pytest.raises(device.DeviceError):
        srv.release(tickets[0])
    used = {(int(c["G"]), int(c["mapping"])) for c in srv.stats() if c["units"]}
    assert {(4, 2), (8, 2), (16, 2)} <= used, used
    srv.close()

"""The serving front end with the real engine: requests sent as frames over
TCP by two clients come back as the same endpoint records as the plain
per-solve tracker gives in-process, bit for bit, and as the reference's
for the systems it is run on; invalid systems come back as error frames."""
import threading

import numpy as np
import pytest

from polyhomocont_amd import abi, device, frontend
from polyhomocont_amd import workloads as W

pytestmark = pytest.mark.gpu
FIELDS = ("status", "winding", "eg_status", "naccepted", "nrejected", "t",
          "eg_error", "z")


def _zero_terms():
    base = W.random_sparse(111)
    eqs = []
    for c, e in base.equations:
        c = c.copy()
        c[0] = 0.0
        eqs.append((c, e))
    return abi.UserSystemData(base.n, eqs)


def test_front_end_matches_plain_tracker(ctx, host, ref):
    systems = {
        1: (W.config_system("c1"), 1),
        2: (W.random_sparse(104), 2),
        3: (_zero_terms(), 3),
        4: (W.dense_system(3, 2, 731), 4),
        5: (abi.UserSystemData(2, [
            (np.array([1.0 + 0j]), np.array([[0, 0]])),
            (np.array([1.0 + 0j, 1.0]), np.array([[1, 1], [0, 0]]))]), 1),
        6: (W.ref_test_systems()["mixed"], 4),
    }
    plain = {}
    for rid, (usys, seed) in systems.items():
        if rid == 5:
            continue
        opts = abi.default_options(seed=seed)
        g, b, a, ns = host.homotopy(usys, seed)
        plain[rid] = ctx.track_paths(host.pack(usys), abi.HomotopyData(g, b, a),
                                     opts, host.start_points(usys, seed))
    srv = device.Server(ctx, max_resident_paths=2048,
                        hbm_budget_bytes=64 << 20, units_per_launch=32)
    fe = frontend.FrontEnd(srv, host)
    stop = threading.Event()
    th = threading.Thread(target=fe.serve, args=(stop,), daemon=True)
    th.start()
    print("[gpu] front end at", fe.address, flush=True)
    try:
        clients = [frontend.Client(fe.address), frontend.Client(fe.address)]
        for rid, (usys, seed) in systems.items():
            clients[rid % 2].send(usys, abi.default_options(seed=seed), rid)
        replies = {}
        for cl in clients:
            for _ in range(sum(1 for r in systems if r % 2 == clients.index(cl))):
                rid, rec, err, msg = cl.receive()
                replies[rid] = (rec, err, msg)
        for cl in clients:
            cl.close()
    finally:
        stop.set()
        th.join(timeout=60)
        fe.close()
        srv.close()
    assert set(replies) == set(systems)
    assert replies[5][0] is None and replies[5][1] == 1, replies[5]
    for rid, p in plain.items():
        rec, err, msg = replies[rid]
        assert err == 0, msg
        n = rec.npaths
        assert n == systems[rid][0].bezout
        for f in FIELDS:
            assert np.array_equal(getattr(rec, f)[:n], getattr(p, f)[:n]), (rid, f)
        assert np.array_equal(rec.nlin[:n], p.nlin[:n])
    # the reference's records for two of the requests, path by path
    for rid in (1, 3):
        usys, seed = systems[rid]
        opts = abi.default_options(seed=seed)
        want = ref.endpoints(usys, opts, 0, usys.bezout)
        rec = replies[rid][0]
        for f in FIELDS:
            assert np.array_equal(getattr(rec, f)[:usys.bezout],
                                  getattr(want, f)[:usys.bezout]), (rid, f)

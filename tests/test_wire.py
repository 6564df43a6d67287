"""Request/endpoint frames of the serving front end (include/phc_wire.h) and
the front end's request loop, on the CPU.

The request frame must carry a system exactly as given to add_equation
(src/sparse.f90:83-141): the receiving side rebuilds it with the library's
builder, so zero terms are dropped and bad equations rejected there, as
for an in-process caller. These tests check that property through the
Fortran builder (pack and homotopy of the decoded system equal those of
the original, errors included), the byte layout's integrity checks, and
the front end's protocol with a stand-in engine (the GPU test of the same
loop uses the real engine)."""
import socket
import struct
import threading

import numpy as np
import pytest

from polyhomocont_amd import abi, wire
from polyhomocont_amd import workloads as W


@pytest.fixture(scope="module")
def host():
    from polyhomocont_amd import host as h
    return h.library()


def _systems():
    rng = np.random.default_rng(5)
    zero_terms = W.random_sparse(111)
    eqs = []
    for c, e in zero_terms.equations:
        c = c.copy()
        c[len(c) // 2] = 0.0  # dropped by the builder
        eqs.append((c, e))
    return {
        "c1": W.config_system("c1"),
        "sparse": W.random_sparse(107),
        "zero terms": abi.UserSystemData(zero_terms.n, eqs),
        "constant eq": abi.UserSystemData(2, [
            (np.array([1.0 + 0j, 2.0]), np.array([[1, 1], [0, 0]])),
            (np.array([3.0 + 0j]), np.array([[0, 0]]))]),
        "negative exp": abi.UserSystemData(2, [
            (np.array([1.0 + 0j, 2.0]), np.array([[1, 0], [0, -1]])),
            (np.array([1.0 + 0j, 1.0]), np.array([[0, 2], [0, 0]]))]),
        "too few eqs": abi.UserSystemData(3, [
            (rng.normal(size=3) + 0j, np.array([[1, 0, 0], [0, 2, 0], [0, 0, 0]]))]),
        "surplus eq": abi.UserSystemData(2, [
            (np.array([1.0 + 0j, -1.0]), np.array([[2, 0], [0, 0]])),
            (np.array([1.0 + 0j, -2.0]), np.array([[0, 2], [0, 0]])),
            (np.array([1.0 + 0j, -3.0]), np.array([[1, 1], [0, 0]]))]),
    }


def _same_system(a, b):
    assert a.n == b.n and len(a.equations) == len(b.equations)
    for (ca, ea), (cb, eb) in zip(a.equations, b.equations):
        assert np.array_equal(ca, cb) and ca.dtype == cb.dtype
        assert np.array_equal(ea, eb)


def _build(host, usys, seed):
    """What the builder makes of a system: packed arrays and homotopy, or
    the error text."""
    try:
        p = host.pack(usys)
        g, b, a, ns = host.homotopy(usys, seed)
    except ValueError as e:
        return ("error", str(e))
    return (p.first.tolist(), p.deg.tolist(), p.hexps.tobytes(),
            p.coefs.tobytes(), complex(g), b.tobytes(), a.tobytes(), ns)


@pytest.mark.parametrize("name", list(_systems()))
def test_request_round_trip_through_builder(host, name):
    usys = _systems()[name]
    opts = abi.default_options(seed=11)
    frame = wire.encode_request(usys, opts, request_id=1234567890123)
    assert len(frame) % 8 == 0
    assert wire.peek(frame) == (wire.REQUEST, len(frame))
    rid, back, bopts = wire.decode_request(frame)
    assert rid == 1234567890123
    _same_system(usys, back)
    assert bytes(bopts) == bytes(opts)
    assert _build(host, back, 11) == _build(host, usys, 11)


def test_builder_outcomes_cover_drop_and_reject(host):
    s = _systems()
    # zero terms vanish in the build: same packed form as without them
    zt = s["zero terms"]
    clean = abi.UserSystemData(zt.n, [(c[c != 0], e[c != 0]) for c, e in zt.equations])
    assert _build(host, zt, 1)[:4] == _build(host, clean, 1)[:4]
    for bad in ("constant eq", "negative exp", "too few eqs"):
        assert _build(host, s[bad], 1)[0] == "error", bad
    assert _build(host, s["surplus eq"], 1)[0] != "error"


def test_all_option_fields_travel():
    opts = abi.default_options()
    fields = []
    for part in ("tracker", "endgame"):
        sub = getattr(opts, part)
        fields += [(sub, f) for f, _ in sub._fields_ if f != "reserved"]
    fields += [(opts, f) for f, _ in abi.SolveOptions._fields_
               if f not in ("tracker", "endgame", "reserved")]
    assert len(fields) == 26
    for k, (obj, f) in enumerate(fields):
        v = getattr(obj, f)
        setattr(obj, f, type(v)(k + 2) if isinstance(v, int) else 0.5 + k)
    _, _, back = wire.decode_request(
        wire.encode_request(W.config_system("c1"), opts, 3))
    assert bytes(back) == bytes(opts)


def _corrupt(frame, at, value):
    b = bytearray(frame)
    b[at:at + len(value)] = value
    return bytes(b)


def test_request_integrity_checks():
    frame = wire.encode_request(W.random_sparse(104), abi.default_options(), 7)
    bad = {
        "magic": _corrupt(frame, 0, b"XXXX"),
        "version": _corrupt(frame, 4, struct.pack("<H", 2)),
        "length": _corrupt(frame, 8, struct.pack("<Q", len(frame) + 8)),
        "checksum": _corrupt(frame, len(frame) - 20, b"\x01"),
        "truncated": frame[:-8],
        "short": frame[:20],
        "endpoints magic": _corrupt(frame, 0, b"PHCE"),
    }
    for name, f in bad.items():
        with pytest.raises(wire.WireError):
            wire.decode_request(f)
    with pytest.raises(wire.WireError):
        wire.peek(b"XXXX" + frame[4:16])


def _resign(frame):
    """Recomputes the checksum after an edit of the body."""
    body = frame[32:]
    h = 1469598103934665603
    for x in body:
        h = ((h ^ x) * 1099511628211) & (2 ** 64 - 1)
    return frame[:24] + struct.pack("<Q", h) + body


def test_checksum_is_fnv1a_of_the_body():
    frame = wire.encode_request(W.config_system("c1"), abi.default_options(), 1)
    assert _resign(frame) == frame


def test_counts_must_match_the_body():
    frame = wire.encode_request(W.random_sparse(104), abi.default_options(), 7)
    n, neq, total = struct.unpack_from("<iiq", frame, 32)
    cases = {
        "total beyond body": _corrupt(frame, 40, struct.pack("<q", total + 1)),
        "huge total": _corrupt(frame, 40, struct.pack("<q", 2 ** 62)),
        "n zero": _corrupt(frame, 32, struct.pack("<i", 0)),
        "negative neq": _corrupt(frame, 36, struct.pack("<i", -1)),
        # a negative count that keeps the sum: first -1, second +1
        "negative count": _corrupt(_corrupt(frame, 224, struct.pack("<i", -1)),
                                   228, struct.pack("<i", struct.unpack_from(
                                       "<i", frame, 228)[0] + struct.unpack_from(
                                       "<i", frame, 224)[0] + 1)),
    }
    for name, f in cases.items():
        with pytest.raises(wire.WireError):
            wire.decode_request(_resign(f))
    # the untouched frame still decodes after re-signing
    wire.decode_request(_resign(frame))


def test_endpoints_round_trip_and_flags():
    rng = np.random.default_rng(3)
    npaths, N = 37, 6
    e = abi.EndpointArrays(npaths, N)
    for f in ("status", "winding", "eg_status", "naccepted", "nrejected",
              "nlin", "nbe"):
        getattr(e, f)[:] = rng.integers(-5, 1000, npaths)
    for f in ("t", "eg_error", "eg_residual"):
        getattr(e, f)[:] = rng.normal(size=npaths)
    e.z[:] = rng.normal(size=(npaths, N)) + 1j * rng.normal(size=(npaths, N))
    for flags in (0, wire.EG_RESIDUAL, wire.NLIN | wire.NBE, wire.ALL_ARRAYS):
        frame = wire.encode_endpoints(99, e, npaths, flags)
        assert wire.peek(frame) == (wire.ENDPOINTS, len(frame))
        rid, back, err, msg = wire.decode_endpoints(frame)
        assert (rid, err, msg) == (99, 0, "")
        assert back.npaths == npaths and back.N == N
        for f in ("status", "t", "z", "winding", "eg_status", "eg_error",
                  "naccepted", "nrejected"):
            assert np.array_equal(getattr(back, f), getattr(e, f)), f
        for f, bit in (("eg_residual", wire.EG_RESIDUAL), ("nlin", wire.NLIN),
                       ("nbe", wire.NBE)):
            want = getattr(e, f) if flags & bit else np.zeros_like(getattr(e, f))
            assert np.array_equal(getattr(back, f), want), f
    with pytest.raises(wire.WireError):
        wire.decode_endpoints(frame[:-8])
    with pytest.raises(wire.WireError):
        wire.decode_request(frame)


def test_error_reply():
    frame = wire.encode_error(5, wire.PHC_ERR_INVALID_INPUT, "pack status 1")
    assert wire.decode_endpoints(frame) == (5, None, 1, "pack status 1")
    with pytest.raises(wire.WireError):
        wire.encode_error(5, 0, "no code")


def test_frame_reader_splits_a_stream():
    frames = [wire.encode_request(W.random_sparse(104 + k),
                                  abi.default_options(), k) for k in range(3)]
    frames.append(wire.encode_error(9, 1, "x"))
    stream = b"".join(frames)
    r = wire.FrameReader()
    got = []
    for i in range(0, len(stream), 37):
        got += r.feed(stream[i:i + 37])
    assert got == frames and not r.buf
    with pytest.raises(wire.WireError):
        wire.FrameReader().feed(b"garbage that is not a frame header")


class StandInEngine:
    """The engine's interface with canned records: a request finishes after
    two progress rounds, its records are the start points it was given."""

    def __init__(self):
        self.next, self.live = 0, {}

    def submit(self, packed, hom, opts, start):
        self.next += 1
        out = abi.EndpointArrays(len(start), packed.n + 1)
        out.z[:len(start)] = start
        out.status[:len(start)] = 0
        out.naccepted[:len(start)] = opts.seed
        self.live[self.next] = 2
        return self.next, out

    def progress(self):
        for t in self.live:
            self.live[t] -= 1

    def done(self, t):
        return self.live[t] <= 0

    def release(self, t):
        del self.live[t]


def test_front_end_protocol(host):
    from polyhomocont_amd import frontend
    fe = frontend.FrontEnd(StandInEngine(), host)
    stop = threading.Event()
    th = threading.Thread(target=fe.serve, args=(stop,), daemon=True)
    th.start()
    try:
        s = _systems()
        cl = frontend.Client(fe.address, timeout=60)
        sent = {1: ("c1", 3), 2: ("constant eq", 1), 3: ("zero terms", 5),
                4: ("sparse", 2)}
        for rid, (name, seed) in sent.items():
            cl.send(s[name], abi.default_options(seed=seed), rid)
        replies = {}
        for _ in sent:
            rid, rec, err, msg = cl.receive()
            replies[rid] = (rec, err, msg)
        assert replies[2][0] is None and replies[2][1] == 1
        assert "invalid system" in replies[2][2]
        for rid in (1, 3, 4):
            name, seed = sent[rid]
            rec, err, _ = replies[rid]
            assert err == 0
            start = host.start_points(s[name], seed)
            assert rec.npaths == len(start)
            assert np.array_equal(rec.z[:len(start)], start)
            assert (rec.naccepted[:len(start)] == seed).all()
        # a stream that is not frames gets one error reply, then EOF
        raw = socket.create_connection(fe.address, timeout=60)
        raw.sendall(b"\x00" * 64)
        f = wire.recv_frame(raw)
        assert wire.decode_endpoints(f)[2] == 1
        assert wire.recv_frame(raw) is None
        raw.close()
        cl.close()
        assert fe.served == 3 and fe.rejected == 2
    finally:
        stop.set()
        th.join(timeout=30)
        fe.close()


class FailingEngine(StandInEngine):
    """Raises an unexpected error for requests with seed 13."""

    def submit(self, packed, hom, opts, start):
        if opts.seed == 13:
            raise RuntimeError("engine fault")
        return super().submit(packed, hom, opts, start)


def test_front_end_limits_keep_serving(host):
    """A request with more paths than the engine takes is refused from its
    degrees (no start points are built), an unexpected failure of one
    request becomes its error reply, a frame above the size limit ends only
    its own connection -- and the loop keeps serving other requests."""
    from polyhomocont_amd import frontend
    fe = frontend.FrontEnd(FailingEngine(), host, max_paths=100,
                           max_frame_bytes=16384)
    stop = threading.Event()
    th = threading.Thread(target=fe.serve, args=(stop,), daemon=True)
    th.start()
    try:
        s = _systems()
        big = W.dense_system(5, 3, 9)                 # 243 paths > 100
        cl = frontend.Client(fe.address, timeout=60)
        cl.send(big, abi.default_options(), 1)
        cl.send(s["c1"], abi.default_options(seed=13), 2)
        cl.send(s["c1"], abi.default_options(seed=4), 3)
        replies = {}
        for _ in range(3):
            rid, rec, err, msg = cl.receive()
            replies[rid] = (rec, err, msg)
        assert replies[1][1] == wire.PHC_ERR_CAPACITY and "243" in replies[1][2]
        assert replies[2][1] == -1 and "engine fault" in replies[2][2]
        assert replies[3][1] == 0 and replies[3][0].npaths == 16
        # an oversized frame: one error reply, then that connection closes
        raw = socket.create_connection(fe.address, timeout=60)
        raw.sendall(wire.encode_request(W.dense_system(7, 3, 5),
                                        abi.default_options(), 9))
        f = wire.recv_frame(raw)
        assert wire.decode_endpoints(f)[2] == 1
        assert wire.recv_frame(raw) is None
        raw.close()
        # the first client is still served
        cl.send(s["c1"], abi.default_options(seed=2), 4)
        rid, rec, err, _ = cl.receive()
        assert rid == 4 and err == 0
        cl.close()
    finally:
        stop.set()
        th.join(timeout=30)
        fe.close()

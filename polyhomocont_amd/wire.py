"""Python handle on libphc_wire.so (include/phc_wire.h): the request and
endpoint frames of an out-of-process serving front end, and helpers that
send and receive them over a byte stream.

A request frame carries a system exactly as given to `add_equation`
(src/sparse.f90:83-141) plus its `solve_options` (src/solve.f90:57-84).
The receiver rebuilds the system with the library's own builder, so
validation and zero-term dropping are those of an in-process call.
"""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libphc_wire.so")

REQUEST, ENDPOINTS = 1, 2
EG_RESIDUAL, NLIN, NBE = 1, 2, 4
ALL_ARRAYS = EG_RESIDUAL | NLIN | NBE
HEADER_BYTES = 32
PEEK_BYTES = 16
MAX_FRAME = 1 << 30
PHC_ERR_INVALID_INPUT = 1
PHC_ERR_CAPACITY = -2

_lib = None


class WireError(ValueError):
    """A frame that could not be encoded or decoded."""


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise WireError(f"{LIB_PATH} is missing: build first")
        L = C.CDLL(LIB_PATH)
        P, i32, i64, u64 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64
        L.phc_wire_last_error.restype = C.c_char_p
        L.phc_wire_request_bytes.argtypes = [P]
        L.phc_wire_request_bytes.restype = i64
        L.phc_wire_encode_request.argtypes = [u64, P, P, P, i64, P]
        L.phc_wire_decode_request.argtypes = [P, i64, P, P, P]
        L.phc_wire_endpoints_bytes.argtypes = [i64, i32, i32]
        L.phc_wire_endpoints_bytes.restype = i64
        L.phc_wire_encode_endpoints.argtypes = [u64, i64, i32, i32, P, P,
                                                i64, P]
        L.phc_wire_error_bytes.argtypes = [C.c_char_p]
        L.phc_wire_error_bytes.restype = i64
        L.phc_wire_encode_error.argtypes = [u64, i32, C.c_char_p, P, i64, P]
        L.phc_wire_decode_endpoints.argtypes = [P, i64, P, P, P, P, P, P, P,
                                                P]
        L.phc_wire_peek.argtypes = [P, i64, P, P]
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        msg = lib().phc_wire_last_error().decode(errors="replace")
        raise WireError(f"{what}: {msg} ({rc})")


def _aligned(nbytes):
    """Zeroed, 8-byte aligned buffer of at least nbytes."""
    return np.zeros((max(int(nbytes), 8) + 7) // 8, np.uint64)


def _aligned_copy(frame):
    frame = memoryview(frame).cast("B")
    buf = _aligned(len(frame))
    buf.view(np.uint8)[:len(frame)] = np.frombuffer(frame, np.uint8)
    return buf, len(frame)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def encode_request(usys, opts, request_id=0):
    """Request frame (bytes) of a system as given to add_equation and its
    solve options."""
    L = lib()
    n = L.phc_wire_request_bytes(C.byref(usys.c))
    if n < 0:
        raise WireError("malformed system")
    buf = _aligned(n)
    written = C.c_int64()
    _check(L.phc_wire_encode_request(request_id, C.byref(usys.c),
                                     C.byref(opts), _ptr(buf), buf.nbytes,
                                     C.byref(written)), "encode_request")
    return buf.view(np.uint8)[:written.value].tobytes()


def decode_request(frame):
    """(request_id, UserSystemData, SolveOptions) of a request frame. The
    system is returned as given, before the builder's checks."""
    L = lib()
    buf, n = _aligned_copy(frame)
    rid = C.c_uint64()
    us = abi.UserSystem()
    opts = abi.SolveOptions()
    _check(L.phc_wire_decode_request(_ptr(buf), n, C.byref(rid),
                                     C.byref(us), C.byref(opts)),
           "decode_request")
    counts = (np.ctypeslib.as_array(C.cast(us.nterms, C.POINTER(C.c_int32)),
                                    (us.neq,)).copy()
              if us.neq else np.zeros(0, np.int32))
    total = int(counts.sum())
    if total:
        cf = np.ctypeslib.as_array(C.cast(us.coefs, C.POINTER(C.c_double)),
                                   (total, 2))
        ex = np.ctypeslib.as_array(C.cast(us.exps, C.POINTER(C.c_int32)),
                                   (total, us.n))
        coefs = (cf[:, 0] + 1j * cf[:, 1]).astype(np.complex128)
        exps = ex.copy()
    else:
        coefs = np.zeros(0, np.complex128)
        exps = np.zeros((0, us.n), np.int32)
    eqs, off = [], 0
    for m in counts:
        eqs.append((coefs[off:off + m], exps[off:off + m]))
        off += m
    return rid.value, abi.UserSystemData(us.n, eqs), opts


def encode_endpoints(request_id, out, npaths, flags=ALL_ARRAYS):
    """Endpoints frame of the first npaths records of an EndpointArrays."""
    L = lib()
    n = L.phc_wire_endpoints_bytes(npaths, out.N, flags)
    if n < 0:
        raise WireError("endpoint sizes out of range")
    buf = _aligned(n)
    written = C.c_int64()
    _check(L.phc_wire_encode_endpoints(request_id, npaths, out.N, flags,
                                       C.byref(out.c), _ptr(buf), buf.nbytes,
                                       C.byref(written)), "encode_endpoints")
    return buf.view(np.uint8)[:written.value].tobytes()


def encode_error(request_id, code, message):
    """Error reply: a nonzero PHC_ERR_* code and a message."""
    L = lib()
    msg = message.encode()[:4096]
    buf = _aligned(L.phc_wire_error_bytes(msg))
    written = C.c_int64()
    _check(L.phc_wire_encode_error(request_id, code, msg, _ptr(buf),
                                   buf.nbytes, C.byref(written)),
           "encode_error")
    return buf.view(np.uint8)[:written.value].tobytes()


def decode_endpoints(frame):
    """(request_id, EndpointArrays or None, error code, message) of a
    reply. The arrays are copies; absent optional arrays stay zero."""
    L = lib()
    buf, n = _aligned_copy(frame)
    rid, npaths = C.c_uint64(), C.c_int64()
    N, flags, err, mlen = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
    msg = C.c_void_p()
    e = abi.Endpoints()
    _check(L.phc_wire_decode_endpoints(
        _ptr(buf), n, C.byref(rid), C.byref(npaths), C.byref(N),
        C.byref(flags), C.byref(err), C.byref(msg), C.byref(mlen),
        C.byref(e)), "decode_endpoints")
    if err.value:
        text = C.string_at(msg.value, mlen.value).decode(errors="replace") \
            if mlen.value else ""
        return rid.value, None, err.value, text
    np_, NN = npaths.value, N.value
    out = abi.EndpointArrays(np_, NN)

    def take(name, ctype, count):
        p = getattr(e, name)
        if p and count:
            src = np.ctypeslib.as_array(C.cast(p, C.POINTER(ctype)), (count,))
            dst = getattr(out, name).reshape(-1)
            if name == "z":
                dst = dst.view(np.float64)
            dst[:count] = src

    for name in ("status", "winding", "eg_status", "naccepted", "nrejected",
                 "nlin", "nbe"):
        take(name, C.c_int32, np_)
    for name in ("t", "eg_error", "eg_residual"):
        take(name, C.c_double, np_)
    take("z", C.c_double, np_ * NN * 2)
    return rid.value, out, 0, ""


def peek(head):
    """(kind, frame bytes) from the first 16 bytes of a frame."""
    L = lib()
    buf, n = _aligned_copy(bytes(head[:PEEK_BYTES]))
    kind, nbytes = C.c_int32(), C.c_int64()
    _check(L.phc_wire_peek(_ptr(buf), n, C.byref(kind), C.byref(nbytes)),
           "peek")
    return kind.value, nbytes.value


class FrameReader:
    """Splits a byte stream into whole frames, each at most `max_bytes`
    long (a longer frame is refused from its header, before its body is
    buffered)."""

    def __init__(self, max_bytes=1 << 30):
        self.buf = bytearray()
        self.max_bytes = int(max_bytes)

    def feed(self, data):
        """Appends bytes; returns the frames completed by them. Raises
        WireError on a stream that is not a frame sequence or announces a
        frame above the limit."""
        self.buf += data
        frames = []
        while len(self.buf) >= PEEK_BYTES:
            _, n = peek(self.buf)
            if n > self.max_bytes:
                raise WireError(f"frame of {n} bytes above the limit of "
                                f"{self.max_bytes}")
            if len(self.buf) < n:
                break
            frames.append(bytes(self.buf[:n]))
            del self.buf[:n]
        return frames


def send_frame(sock, frame):
    sock.sendall(frame)


def recv_frame(sock):
    """Next whole frame from a blocking socket, or None at end of stream."""
    head = b""
    while len(head) < PEEK_BYTES:
        chunk = sock.recv(PEEK_BYTES - len(head))
        if not chunk:
            if head:
                raise WireError("stream ended inside a frame header")
            return None
        head += chunk
    _, n = peek(head)
    parts, got = [head], len(head)
    while got < n:
        chunk = sock.recv(min(n - got, 1 << 20))
        if not chunk:
            raise WireError("stream ended inside a frame")
        parts.append(chunk)
        got += len(chunk)
    return b"".join(parts)

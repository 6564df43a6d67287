"""phc_track_paths_noalloc: the allocation-free variant of phc_track_paths.

The caller owns the start points and the endpoint arrays in device memory
(this test allocates them with hipMalloc through ctypes; the entry point
itself allocates nothing). Its records must be those of phc_track_paths bit
for bit, on consecutive calls that reuse the library's static scratch --
the second with a smaller input, so that stale contents of the first call
would show -- and a system whose tables exceed the scratch is refused
before anything runs.
"""
import ctypes as C

import numpy as np
import pytest

from polyhomocont_amd import abi, device, host, workloads as W

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2
FIELDS = ("status", "t", "z", "winding", "eg_status", "eg_error",
          "eg_residual", "naccepted", "nrejected", "nlin", "nbe")


class Hip:
    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t,
                                       C.c_int]
        self.lib.hipSetDevice.argtypes = [C.c_int]
        self.ptrs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), max(nbytes, 8)) == 0
        self.ptrs.append(p.value)
        return p.value

    def put(self, a):
        a = np.ascontiguousarray(a)
        d = self.alloc(a.nbytes)
        assert self.lib.hipMemcpy(d, a.ctypes.data, a.nbytes, H2D) == 0
        return d

    def get(self, d, like):
        assert self.lib.hipMemcpy(like.ctypes.data, d, like.nbytes, D2H) == 0
        return like

    def free(self):
        for p in self.ptrs:
            self.lib.hipFree(p)
        self.ptrs = []


@pytest.fixture(scope="module")
def hip():
    h = Hip()
    assert h.lib.hipSetDevice(0) == 0
    yield h
    h.free()


def noalloc_call(ctx, hip, packed, hom, opts, start):
    """Runs the entry point on device copies of `start`; returns the
    records copied back (abi.EndpointArrays) and the return code."""
    start = np.ascontiguousarray(start, np.complex128)
    npaths, N = start.shape
    host_out = abi.EndpointArrays(npaths, N)
    dev = {f: hip.alloc(getattr(host_out, f).nbytes) for f in FIELDS}
    d_out = abi.Endpoints(**{f: dev.get(f) for f, _ in abi.Endpoints._fields_})
    d_start = hip.put(start)
    rc = ctx.L.phc_track_paths_noalloc(ctx.h, C.byref(packed.c),
                                       C.byref(hom.c), C.byref(opts), npaths,
                                       d_start, C.byref(d_out))
    if rc == 0:
        for f in FIELDS:
            hip.get(dev[f], getattr(host_out, f))
    return rc, host_out


def request(usys, count=None):
    H = host.library()
    opts = abi.default_options()
    g, b, a, ns = H.homotopy(usys, opts.seed)
    start = H.start_points(usys, opts.seed, 0, count or ns)
    return H.pack(usys), abi.HomotopyData(g, b, a), opts, start


def assert_same(got, want, n):
    for f in FIELDS:
        assert np.array_equal(getattr(got, f)[:n], getattr(want, f)[:n]), f


def test_noalloc_matches_track_paths_twice(hip):
    ctx = device.Context(0)
    lib = device.lib()
    lib.phc_track_paths_noalloc.argtypes = [C.c_void_p] * 4 + [
        C.c_int64, C.c_void_p, C.c_void_p]
    lib.phc_noalloc_scratch_bytes.restype = C.c_int64
    assert lib.phc_noalloc_scratch_bytes() == 1 << 20
    # a large request (C2: 840 terms, grid tables staged in LDS), then a
    # smaller one of another shape, then the first again
    cases = [("c2 paths 0..95", request(W.config_system("c2"), 96)),
             ("c1", request(W.config_system("c1"))),
             ("mixed", request(W.ref_test_systems()["mixed"])),
             ("c2 paths 0..95 again", request(W.config_system("c2"), 96))]
    for name, (packed, hom, opts, start) in cases:
        print(f"[gpu] noalloc {name}: {len(start)} paths", flush=True)
        want = ctx.track_paths(packed, hom, opts, start)
        rc, got = noalloc_call(ctx, hip, packed, hom, opts, start)
        assert rc == 0, ctx.L.phc_thread_last_error()
        assert_same(got, want, len(start))
        c = ctx.counters()
        assert c["units"] == int(got.nlin[:len(start)].sum() +
                                 got.nbe[:len(start)].sum())
    ctx.close()


def test_noalloc_refuses_tables_beyond_scratch(hip):
    """n = 15, all monomials of degree <= 4 (58,140 terms): its term table
    alone exceeds the 1 MiB scratch; refused with PHC_ERR_CAPACITY."""
    ctx = device.Context(0)
    lib = device.lib()
    lib.phc_track_paths_noalloc.argtypes = [C.c_void_p] * 4 + [
        C.c_int64, C.c_void_p, C.c_void_p]
    n = 15
    mons = np.array(W.monomials(n, 4), np.int32)
    rng = np.random.default_rng(3)
    eqs = [(rng.normal(size=len(mons)) + 1j * rng.normal(size=len(mons)), mons)
           for _ in range(n)]
    usys = abi.UserSystemData(n, eqs)
    packed = usys.packed()
    hom = abi.HomotopyData(1.0 + 0j, np.ones(n, np.complex128),
                           np.ones(n + 1, np.complex128) / np.sqrt(n + 1))
    start = np.ones((4, n + 1), np.complex128)
    rc, _ = noalloc_call(ctx, hip, packed, hom, abi.default_options(), start)
    assert rc == -2
    assert b"scratch" in ctx.L.phc_thread_last_error()
    ctx.close()

"""Python handle on libphc_hip.so (the C ABI of include/phc.h).

Loads the in-tree library built by polyhomocont_amd/build.py and fails
loudly when it is missing: there is no CPU fallback for sparse systems.
"""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libphc_hip.so")
SERVER_MAX_CLASSES = 16  # PHC_SERVER_MAX_CLASSES of include/phc.h

_lib = None


class DeviceError(RuntimeError):
    """A failed library call; `code` is its PHC_ERR_* return value."""

    def __init__(self, message, code=-1):
        super().__init__(message)
        self.code = code


def lib():
    """The loaded C library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DeviceError(f"{LIB_PATH} is missing: build the HIP library "
                              "first (python -c 'import __graft_entry__ as g; "
                              "g.build()')")
        L = C.CDLL(LIB_PATH)
        P, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
        L.phc_default_options.argtypes = [P]
        L.phc_default_options.restype = None
        L.phc_ctx_create.argtypes = [C.c_int, C.POINTER(P)]
        L.phc_ctx_destroy.argtypes = [P]
        L.phc_ctx_lanes_per_path.argtypes = [P, C.c_int]
        L.phc_ctx_last_error.argtypes = [P]
        L.phc_ctx_last_error.restype = C.c_char_p
        L.phc_ctx_sync.argtypes = [P]
        L.phc_ctx_counters.argtypes = [P, P]
        L.phc_ctx_exact_reruns.argtypes = [P, P]
        L.phc_ctx_set_mapping.argtypes = [P, C.c_int, C.c_int]
        L.phc_ctx_autotune.argtypes = [P, P, P, P, i64, P, P, P]
        L.phc_ctx_clear_tuning.argtypes = [P]
        L.phc_ctx_autotune_batch.argtypes = [P, i32, P, P, P, P, P, P, P]
        L.phc_ctx_last_plan.argtypes = [P, i32, P]
        L.phc_track_paths.argtypes = [P, P, P, P, i64, P, P]
        L.phc_stage_create.argtypes = [P, P, P, P, i64, P, C.POINTER(P)]
        L.phc_stage_run.argtypes = [P, P]
        L.phc_stage_fetch.argtypes = [P, P, P]
        L.phc_stage_destroy.argtypes = [P, P]
        L.phc_stage_elapsed.argtypes = [P, P, P]
        L.phc_stage_launch_info.argtypes = [P, P, P]
        L.phc_track_batch.argtypes = [P, i32, P, P, P, P, P, P]
        L.phc_server_create.argtypes = [P, P, C.POINTER(P)]
        L.phc_server_submit.argtypes = [P, P, P, P, i64, P, P, P]
        L.phc_server_progress.argtypes = [P]
        L.phc_server_progress_for.argtypes = [P, i64]
        L.phc_server_poll.argtypes = [P, C.c_uint64]
        L.phc_server_wait.argtypes = [P, C.c_uint64]
        L.phc_server_release.argtypes = [P, C.c_uint64]
        L.phc_server_tickets.argtypes = [P, P]
        L.phc_server_stats.argtypes = [P, P]
        L.phc_server_ticket_class.argtypes = [P, C.c_uint64]
        L.phc_server_max_request_paths.argtypes = [P]
        L.phc_server_max_request_paths.restype = i64
        L.phc_server_destroy.argtypes = [P]
        L.phc_thread_last_error.argtypes = []
        L.phc_thread_last_error.restype = C.c_char_p
        L.phc_ctx_shared.argtypes = [C.c_int, C.POINTER(P)]
        L.phc_set_default_device.argtypes = [C.c_int]
        L.phc_set_default_device.restype = None
        L.phc_get_default_device.argtypes = []
        _lib = L
    return _lib


class Context:
    """A phc_ctx on one HIP device."""

    def __init__(self, device=0):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.phc_ctx_create(device, C.byref(h))
        if rc != 0:
            raise DeviceError(f"phc_ctx_create({device}) failed with {rc}: "
                              "no usable GPU")
        self.h = h

    def close(self):
        if self.h:
            self.L.phc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            # this thread's own message (another thread may share the context)
            msg = self.L.phc_thread_last_error().decode(errors="replace")
            raise DeviceError(f"{what} failed with {rc}: {msg}", rc)

    def counters(self):
        c = np.zeros(4, np.uint64)
        self.L.phc_ctx_counters(self.h, c.ctypes.data)
        occ = float(c[1]) / float(c[2]) if c[2] else None
        util = (float(c[1]) / (float(c[0]) * float(c[3]))
                if c[0] and c[3] else None)
        return dict(units=int(c[1]), unit_slots=int(c[2]), occupancy=occ,
                    slots=int(c[0]), max_slot_units=int(c[3]),
                    slot_utilisation=util)

    def exact_reruns(self):
        """Units of the last tracking call redone with the exact pivot rule
        and divisions (phc_ctx_exact_reruns)."""
        c = np.zeros(1, np.uint64)
        self._check(self.L.phc_ctx_exact_reruns(self.h, c.ctypes.data),
                    "phc_ctx_exact_reruns")
        return int(c[0])

    def set_mapping(self, mapping=0, wg_threads=0):
        """0: chosen per launch, 1: path per G-lane group, 2: path per
        wavefront (split evaluation), 3: path per G*G-lane grid."""
        self._check(self.L.phc_ctx_set_mapping(self.h, mapping, wg_threads),
                    "phc_ctx_set_mapping")

    def lanes_per_path(self, N):
        return self.L.phc_ctx_lanes_per_path(self.h, N)

    def autotune(self, packed, hom, opts, start):
        """Times every launch shape on (a prefix of) this request and keeps
        the fastest for later launches of the same shape; returns the choice
        and the winner's kernel time."""
        start = np.ascontiguousarray(start, np.complex128)
        choice = np.zeros(4, np.int32)
        ms = C.c_double(0.0)
        rc = self.L.phc_ctx_autotune(self.h, C.byref(packed.c), C.byref(hom.c),
                                     C.byref(opts), start.shape[0],
                                     start.ctypes.data, choice.ctypes.data,
                                     C.byref(ms))
        self._check(rc, "phc_ctx_autotune")
        return dict(mapping={1: "row", 2: "split", 3: "grid"}[int(choice[0])],
                    threads=int(choice[1]), staged=bool(choice[2]),
                    kernel_ms=ms.value)

    def clear_tuning(self):
        self._check(self.L.phc_ctx_clear_tuning(self.h), "phc_ctx_clear_tuning")

    def track_paths(self, packed, hom, opts, start):
        """Endpoint records for start points `start` [npaths][N]."""
        start = np.ascontiguousarray(start, np.complex128)
        npaths = start.shape[0]
        out = abi.EndpointArrays(npaths, packed.n + 1)
        rc = self.L.phc_track_paths(self.h, C.byref(packed.c), C.byref(hom.c),
                                    C.byref(opts), npaths, start.ctypes.data,
                                    C.byref(out.c))
        self._check(rc, "phc_track_paths")
        return out

    @staticmethod
    def _batch_args(items):
        n = len(items)
        systems = (abi.System * n)(*[it[0].c for it in items])
        homs = (abi.Homotopy * n)(*[it[1].c for it in items])
        opts = (abi.SolveOptions * n)(*[it[2] for it in items])
        starts = [np.ascontiguousarray(it[3], np.complex128) for it in items]
        npaths = np.array([s.shape[0] for s in starts], np.int64)
        sp = (C.c_void_p * n)(*[s.ctypes.data for s in starts])
        return n, systems, homs, opts, starts, npaths, sp

    def track_batch(self, items):
        """items: list of (packed, hom, opts, start); returns EndpointArrays."""
        n, systems, homs, opts, starts, npaths, sp = self._batch_args(items)
        outs = [abi.EndpointArrays(s.shape[0], it[0].n + 1)
                for s, it in zip(starts, items)]
        oc = (abi.Endpoints * n)(*[o.c for o in outs])
        rc = self.L.phc_track_batch(self.h, n, systems, homs, opts,
                                    npaths.ctypes.data, sp, oc)
        self._check(rc, "phc_track_batch")
        return outs

    def autotune_batch(self, items):
        """Times the launch shapes of every lanes class of this batch (on a
        sample of its paths, bounded launches) and keeps the one of highest
        unit throughput for later batches of the same class shapes; returns
        the choice per class."""
        n, systems, homs, opts, starts, npaths, sp = self._batch_args(items)
        ch = np.zeros(12, np.int32)
        nc = C.c_int32(0)
        rc = self.L.phc_ctx_autotune_batch(self.h, n, systems, homs, opts,
                                           npaths.ctypes.data, sp,
                                           ch.ctypes.data, C.byref(nc))
        self._check(rc, "phc_ctx_autotune_batch")
        names = {1: "row", 2: "split", 3: "grid"}
        return [dict(G=int(ch[4 * c]), mapping=names[int(ch[4 * c + 1])],
                     threads=int(ch[4 * c + 2]), staged=bool(ch[4 * c + 3]))
                for c in range(nc.value)]

    def last_plan(self):
        """Launch shape of every lanes class of the last tracking call."""
        info = np.zeros(8 * 8, np.int32)
        n = self.L.phc_ctx_last_plan(self.h, 8, info.ctypes.data)
        if n < 0:
            self._check(n, "phc_ctx_last_plan")
        names = {0: "row", 1: "split", 2: "grid"}
        return [dict(G=int(info[8 * c]), blocks=int(info[8 * c + 1]),
                     threads=int(info[8 * c + 2]), lds=int(info[8 * c + 3]),
                     staged=bool(info[8 * c + 4]),
                     mapping=names[int(info[8 * c + 5])],
                     lanes_per_path=int(info[8 * c + 6]),
                     systems=int(info[8 * c + 7]))
                for c in range(min(n, 8))]


class Staged:
    """Inputs staged in device memory for repeated timed runs."""

    def __init__(self, ctx, packed, hom, opts, start):
        self.ctx = ctx
        self.start = np.ascontiguousarray(start, np.complex128)
        self.packed, self.hom, self.opts = packed, hom, opts
        h = C.c_void_p()
        rc = ctx.L.phc_stage_create(ctx.h, C.byref(packed.c), C.byref(hom.c),
                                    C.byref(opts), self.start.shape[0],
                                    self.start.ctypes.data, C.byref(h))
        ctx._check(rc, "phc_stage_create")
        self.h = h

    def run(self):
        self.ctx._check(self.ctx.L.phc_stage_run(self.ctx.h, self.h),
                        "phc_stage_run")

    def sync(self):
        self.ctx._check(self.ctx.L.phc_ctx_sync(self.ctx.h), "phc_ctx_sync")

    def fetch(self):
        out = abi.EndpointArrays(self.start.shape[0], self.packed.n + 1)
        self.ctx._check(self.ctx.L.phc_stage_fetch(self.ctx.h, self.h,
                                                   C.byref(out.c)),
                        "phc_stage_fetch")
        return out

    def elapsed_ms(self):
        """Kernel time of the last run (HIP events on the library stream)."""
        ms = np.zeros(1)
        self.ctx._check(self.ctx.L.phc_stage_elapsed(self.ctx.h, self.h,
                                                     ms.ctypes.data),
                        "phc_stage_elapsed")
        return float(ms[0])

    def launch_info(self):
        info = np.zeros(8, np.int32)
        self.ctx._check(self.ctx.L.phc_stage_launch_info(self.ctx.h, self.h,
                                                         info.ctypes.data),
                        "phc_stage_launch_info")
        return dict(group=int(info[0]), lanes_per_path=int(info[6]),
                    workgroups=int(info[1]),
                    threads=int(info[2]), lds_bytes=int(info[3]),
                    staged=bool(info[4]),
                    mapping=("row", "split", "grid")[int(info[5])])

    def close(self):
        if self.h:
            self.ctx.L.phc_stage_destroy(self.ctx.h, self.h)
            self.h = None


class Server:
    """Continuous-batching serving engine (phc_server_* of include/phc.h)."""

    def __init__(self, ctx, max_resident_paths=0, hbm_budget_bytes=0,
                 units_per_launch=0, launch_us=0):
        self.ctx = ctx
        cfg = abi.ServerConfig(max_resident_paths, hbm_budget_bytes,
                               units_per_launch, launch_us)
        h = C.c_void_p()
        ctx._check(ctx.L.phc_server_create(ctx.h, C.byref(cfg), C.byref(h)),
                   "phc_server_create")
        self.h = h
        self._keep = {}

    def submit(self, packed, hom, opts, start, out=None):
        """Queues a request; returns (ticket, EndpointArrays filled on
        completion). `out`, when given, is the caller's EndpointArrays for
        the request's paths (made beforehand, e.g. on a worker thread)."""
        start = np.ascontiguousarray(start, np.complex128)
        if out is None:
            out = abi.EndpointArrays(start.shape[0], packed.n + 1)
        elif out.npaths != start.shape[0] or out.N != packed.n + 1:
            raise ValueError("out does not match the request's paths")
        t = C.c_uint64()
        self.ctx._check(self.ctx.L.phc_server_submit(
            self.h, C.byref(packed.c), C.byref(hom.c), C.byref(opts),
            start.shape[0], start.ctypes.data, C.byref(out.c), C.byref(t)),
            "phc_server_submit")
        self._keep[t.value] = (packed, hom, opts, start, out)
        return t.value, out

    def progress(self, max_wait_us=-1):
        """Retires the classes whose round ended, admits and starts the
        next rounds; waits for a round to end only when none has and none
        could start (at most max_wait_us microseconds when that is not
        negative). Returns the number of requests finished."""
        rc = self.ctx.L.phc_server_progress_for(self.h, max_wait_us)
        if rc < 0:
            self.ctx._check(rc, "phc_server_progress")
        return rc

    def done(self, ticket):
        return self.ctx.L.phc_server_poll(self.h, ticket) == 0

    def wait(self, ticket):
        self.ctx._check(self.ctx.L.phc_server_wait(self.h, ticket),
                        "phc_server_wait")
        return self._keep[ticket][4]

    def release(self, ticket):
        """Forgets a finished ticket (or cancels a queued one); the engine
        keeps no record of it afterwards. Raises while its paths are on the
        device, whose records are still to be copied into its arrays."""
        self.ctx._check(self.ctx.L.phc_server_release(self.h, ticket),
                        "phc_server_release")
        self._keep.pop(ticket, None)

    def stats(self):
        """Per lanes class: G, mapping, launches, kernel ms, units, unit
        slots, resident path slots."""
        out = np.zeros(8 * SERVER_MAX_CLASSES)
        n = self.ctx.L.phc_server_stats(self.h, out.ctypes.data)
        if n < 0:
            self.ctx._check(n, "phc_server_stats")
        keys = ("G", "mapping", "launches", "kernel_ms", "units",
                "unit_slots", "slots", "vmax")
        return [dict(zip(keys, out[8 * i:8 * i + 8])) for i in range(n)]

    def ticket_class(self, ticket):
        c = self.ctx.L.phc_server_ticket_class(self.h, ticket)
        if c < 0:
            self.ctx._check(c, "phc_server_ticket_class")
        return c

    def max_request_paths(self):
        """Most paths one request may have."""
        return int(self.ctx.L.phc_server_max_request_paths(self.h))

    def tickets(self):
        """Tickets queued, admitted and finished-but-unreleased."""
        c = np.zeros(3, np.int64)
        self.ctx._check(self.ctx.L.phc_server_tickets(self.h, c.ctypes.data),
                        "phc_server_tickets")
        return dict(queued=int(c[0]), admitted=int(c[1]), finished=int(c[2]))

    def close(self):
        if self.h:
            self.ctx.L.phc_server_destroy(self.h)
            self.h = None

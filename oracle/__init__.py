"""TEST INFRASTRUCTURE ONLY -- the checker, never the product.

Python bindings for the two CPU oracles of the path tracker:
  * ``ref``  -- oracle/_ref/libphc_ref.so: the unmodified reference Fortran
    compiled from its own sources (oracle/Makefile) plus oracle/ref_driver.f90;
  * ``port`` -- oracle/libphc_oracle.so: the C restatement oracle/phc_oracle.c.

Both export the same entry points (prefix phc_ref_ / phc_orc_) over the
structs of include/phc.h and include/polyhomocont.h. Only tests/,
__graft_entry__.smoke() and the cpu_baseline leg of bench.py import this.
"""
import ctypes as C
import os

import numpy as np

from polyhomocont_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
REF_LIB = os.path.join(HERE, "_ref", "libphc_ref.so")
PORT_LIB = os.path.join(HERE, "libphc_oracle.so")


def build(reference=None):
    """Builds the checkers with oracle/Makefile: the C restatement always;
    the compiled reference, and the reference's own test programs linked
    against the product library, when the reference's sources are present
    (in the build container; the GPU hosts use the prebuilt files)."""
    import subprocess
    ref = reference or os.environ.get("PHC_REFERENCE", "/root/reference")
    targets = ["port"]
    if os.path.isdir(os.path.join(ref, "src")):
        targets.append("ref")
        prod = os.path.join(os.path.dirname(HERE), "polyhomocont_amd", "lib",
                            "libpolyhomocont_amd.so")
        if os.path.exists(prod):
            targets.append("droptests")
    cmd = ["make", "-s", "-C", HERE, f"REF={ref}", *targets]
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=1200)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)


class _Oracle:
    def __init__(self, path, prefix):
        if not os.path.exists(path):
            raise FileNotFoundError(f"oracle library {path} is not built "
                                    "(run __graft_entry__.build())")
        self.lib = C.CDLL(path)
        self.prefix = prefix
        P, i32, i64, f64 = C.c_void_p, C.c_int32, C.c_int64, C.c_double
        sig = {
            "homotopy": ([P, i32, P, P, P, P], C.c_int),
            "start_points": ([P, i32, i64, i64, P], C.c_int),
            "endpoints": ([P, P, i64, i64, P], C.c_int),
            "solve": ([P, P, P, P], C.c_int),
            "eval": ([P, i32, P, P, P, P, P], C.c_int),
            "eval_sparse": ([P, P, P, P], C.c_int),
            "track_segment": ([P, i32, P, i32, P, P, f64, f64, f64, P, P, P,
                               P, P, P], C.c_int),
            "rng": ([i32, i32, i64, P], None),
            "lu": ([i32, P, P, P], C.c_int),
            "cond1": ([i32, P], C.c_double),
        }
        # entry points only the compiled reference's driver has
        optional = {"solve_user": ([P, P, P, P], C.c_int)}
        for name, (args, res) in sig.items():
            f = getattr(self.lib, prefix + name)
            f.argtypes = args
            f.restype = res
            setattr(self, "_" + name, f)
        for name, (args, res) in optional.items():
            f = getattr(self.lib, prefix + name, None)
            if f is not None:
                f.argtypes = args
                f.restype = res
            setattr(self, "_" + name, f)

    # -- thin wrappers ------------------------------------------------------
    def homotopy(self, usys, seed):
        n = usys.n
        g = np.zeros(1, np.complex128)
        b = np.zeros(n, np.complex128)
        a = np.zeros(n + 1, np.complex128)
        ns = np.zeros(1, np.int64)
        st = self._homotopy(C.byref(usys.c), seed, g.ctypes.data, b.ctypes.data,
                            a.ctypes.data, ns.ctypes.data)
        if st:
            raise ValueError(f"homotopy init status {st}")
        return g[0], b, a, int(ns[0])

    def start_points(self, usys, seed, first=0, count=None):
        if count is None:
            count = usys.bezout - first
        z = np.zeros((count, usys.n + 1), np.complex128)
        st = self._start_points(C.byref(usys.c), seed, first, count, z.ctypes.data)
        if st:
            raise ValueError(f"start points status {st}")
        return z

    def endpoints(self, usys, opts, first=0, count=None):
        if count is None:
            count = usys.bezout - first
        out = abi.EndpointArrays(count, usys.n + 1)
        st = self._endpoints(C.byref(usys.c), C.byref(opts), first, count,
                             C.byref(out.c))
        if st:
            raise ValueError(f"endpoints status {st}")
        return out

    def solve(self, usys, opts):
        res = abi.SolveArrays(max(usys.bezout, 1), usys.n)
        st = self._solve(C.byref(usys.c), C.byref(opts), C.byref(res.summary),
                         C.byref(res.paths))
        return st, res

    def solve_user(self, usys, opts):
        """solve on the system wrapped as a user evaluator (affine F only)."""
        if self._solve_user is None:
            raise NotImplementedError("only the compiled reference has it")
        res = abi.SolveArrays(max(usys.bezout, 1), usys.n)
        st = self._solve_user(C.byref(usys.c), C.byref(opts),
                              C.byref(res.summary), C.byref(res.paths))
        return st, res

    def eval(self, usys, seed, z, t):
        N = usys.n + 1
        z = np.ascontiguousarray(z, np.complex128)
        tt = np.array([t], np.complex128)
        h = np.zeros(N, np.complex128)
        hz = np.zeros(N * N, np.complex128)
        ht = np.zeros(N, np.complex128)
        st = self._eval(C.byref(usys.c), seed, z.ctypes.data, tt.ctypes.data,
                        h.ctypes.data, hz.ctypes.data, ht.ctypes.data)
        if st:
            raise ValueError(f"eval status {st}")
        return h, hz.reshape(N, N, order="F"), ht

    def eval_sparse(self, usys, z):
        n = usys.n
        z = np.ascontiguousarray(z, np.complex128)
        f = np.zeros(n, np.complex128)
        jac = np.zeros(n * (n + 1), np.complex128)
        self._eval_sparse(C.byref(usys.c), z.ctypes.data, f.ctypes.data,
                          jac.ctypes.data)
        return f, jac.reshape(n, n + 1, order="F")

    def track_segment(self, usys, seed, topts, z, line=None, arc=None):
        z = np.array(z, np.complex128)
        t0 = np.zeros(1, np.complex128)
        t1 = np.zeros(1, np.complex128)
        r = th0 = dth = 0.0
        if line is not None:
            kind = 1
            t0[0], t1[0] = line
        else:
            kind = 2
            r, th0, dth = arc
        st_ = np.zeros(1, np.int32)
        s = np.zeros(1)
        t = np.zeros(1, np.complex128)
        step = np.zeros(1)
        cnt = np.zeros(2, np.int32)
        rc = self._track_segment(C.byref(usys.c), seed, C.byref(topts), kind,
                                 t0.ctypes.data, t1.ctypes.data, r, th0, dth,
                                 z.ctypes.data, st_.ctypes.data, s.ctypes.data,
                                 t.ctypes.data, step.ctypes.data, cnt.ctypes.data)
        if rc:
            raise ValueError(f"track_segment status {rc}")
        return z, dict(status=int(st_[0]), s=float(s[0]), t=complex(t[0]),
                       step=float(step[0]), naccepted=int(cnt[0]),
                       nrejected=int(cnt[1]))

    def rng(self, seed, kind, count):
        if kind == 0:
            out = np.zeros(count)
        else:
            out = np.zeros(count, np.complex128)
        self._rng(seed, kind, count, out.ctypes.data)
        return out

    def lu(self, a, b):
        a = np.asfortranarray(a, np.complex128).copy(order="F")
        b = np.array(b, np.complex128)
        n = a.shape[0]
        ipiv = np.zeros(n, np.int32)
        st = self._lu(n, a.ctypes.data, ipiv.ctypes.data, b.ctypes.data)
        return st, a, ipiv, b

    def cond1(self, a):
        a = np.asfortranarray(a, np.complex128)
        return self._cond1(a.shape[0], a.ctypes.data)


_cache = {}


def ref():
    """The compiled reference (None if oracle/_ref was not built)."""
    if "ref" not in _cache:
        _cache["ref"] = _Oracle(REF_LIB, "phc_ref_") if os.path.exists(REF_LIB) else None
    return _cache["ref"]


def port():
    """The C restatement."""
    if "port" not in _cache:
        _cache["port"] = _Oracle(PORT_LIB, "phc_orc_")
    return _cache["port"]

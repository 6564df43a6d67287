"""Python handle on libpolyhomocont_amd.so, the Fortran host library
(include/polyhomocont.h): the reference API `solve` and its host pieces
(random constants, start points, packing, classification) over flat arrays.
"""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libpolyhomocont_amd.so")


class HostLibrary:
    def __init__(self, path=LIB_PATH):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is missing: build first")
        # libphc_hip.so is found through the library's rpath ($ORIGIN)
        self.lib = C.CDLL(path)
        P, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
        L = self.lib
        L.phc_f_solve.argtypes = [P, P, P, P]
        L.phc_f_homotopy.argtypes = [P, i32, P, P, P, P]
        L.phc_f_start_points.argtypes = [P, i32, i64, i64, P]
        L.phc_f_pack.argtypes = [P, P, P, P, P, P, P]
        L.phc_f_classify.argtypes = [P, P, i64, P, P, P]
        L.phc_f_set_device.argtypes = [i32]
        L.phc_f_set_device.restype = None
        L.phc_f_solve_user.argtypes = [P, P, P, P]
        L.phc_f_set_threads.argtypes = [i32]
        L.phc_f_set_threads.restype = i32

    def solve(self, usys, opts):
        """(status, SolveArrays) of `solve` on this system."""
        res = abi.SolveArrays(max(usys.bezout, 1), usys.n)
        st = self.lib.phc_f_solve(C.byref(usys.c), C.byref(opts),
                                  C.byref(res.summary), C.byref(res.paths))
        return st, res

    def solve_user(self, usys, opts):
        """(status, SolveArrays) of `solve` on the system wrapped as a user
        evaluator (host tracker, OpenMP threads)."""
        res = abi.SolveArrays(max(usys.bezout, 1), usys.n)
        st = self.lib.phc_f_solve_user(C.byref(usys.c), C.byref(opts),
                                       C.byref(res.summary), C.byref(res.paths))
        return st, res

    def set_threads(self, nthreads):
        """Host threads of this thread's later solves; returns the previous
        count."""
        return int(self.lib.phc_f_set_threads(nthreads))

    def homotopy(self, usys, seed):
        n = usys.n
        g = np.zeros(1, np.complex128)
        b = np.zeros(n, np.complex128)
        a = np.zeros(n + 1, np.complex128)
        ns = np.zeros(1, np.int64)
        st = self.lib.phc_f_homotopy(C.byref(usys.c), seed, g.ctypes.data,
                                     b.ctypes.data, a.ctypes.data,
                                     ns.ctypes.data)
        if st:
            raise ValueError(f"homotopy status {st}")
        return g[0], b, a, int(ns[0])

    def start_points(self, usys, seed, first=0, count=None):
        if count is None:
            count = usys.bezout - first
        z = np.zeros((max(count, 1), usys.n + 1), np.complex128)
        st = self.lib.phc_f_start_points(C.byref(usys.c), seed, first, count,
                                         z.ctypes.data)
        if st:
            raise ValueError(f"start points status {st}")
        return z[:count]

    def pack(self, usys):
        """The homogenized system as the device sees it (abi.PackedSystem)."""
        nt = C.c_int32()
        me = C.c_int32()
        self.lib.phc_f_pack(C.byref(usys.c), C.byref(nt), C.byref(me),
                            None, None, None, None)
        n = usys.n
        first = np.zeros(n + 1, np.int32)
        deg = np.zeros(n, np.int32)
        hexps = np.zeros((max(nt.value, 1), n + 1), np.int32)
        coefs = np.zeros(max(nt.value, 1), np.complex128)
        st = self.lib.phc_f_pack(C.byref(usys.c), C.byref(nt), C.byref(me),
                                 first.ctypes.data, deg.ctypes.data,
                                 hexps.ctypes.data, coefs.ctypes.data)
        if st:
            raise ValueError(f"pack status {st} (system not square)")
        return abi.PackedSystem(n, first, deg, hexps[:nt.value],
                                coefs[:nt.value])

    def classify(self, usys, opts, rec):
        """Host post-processing of solve on endpoint records."""
        npaths = rec.npaths
        res = abi.SolveArrays(max(npaths, 1), usys.n)
        st = self.lib.phc_f_classify(C.byref(usys.c), C.byref(opts), npaths,
                                     C.byref(rec.c), C.byref(res.summary),
                                     C.byref(res.paths))
        return st, res

    def set_device(self, device):
        self.lib.phc_f_set_device(device)


_lib = None


def library():
    global _lib
    if _lib is None:
        _lib = HostLibrary()
    return _lib

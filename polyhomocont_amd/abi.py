"""ctypes mirror of include/phc.h and include/polyhomocont.h.

The structs here are byte-for-byte the C structs of the two headers; the
numpy helpers allocate the struct-of-arrays outputs and build the flat
system descriptions. Nothing in this module computes anything.
"""
import ctypes as C

import numpy as np

i32 = C.c_int32
i64 = C.c_int64
f64 = C.c_double
P = C.c_void_p


class TrackerOptions(C.Structure):
    """tracker_options (reference src/tracker.f90:56-77)."""
    _fields_ = [("step_init", f64), ("step_max", f64), ("step_min", f64),
                ("corrector_tol", f64), ("contraction", f64),
                ("corrector_maxit", i32), ("grow_after", i32),
                ("max_steps", i32), ("reserved", i32)]


class EndgameOptions(C.Structure):
    """endgame_options (reference src/endgame.f90:73-101)."""
    _fields_ = [("t_start", f64), ("radius_factor", f64), ("radius_min", f64),
                ("loop_radius", f64), ("winding_agreement", f64), ("tol", f64),
                ("residual_tol", f64), ("closure_factor", f64),
                ("nsamples", i32), ("max_winding", i32)]


class SolveOptions(C.Structure):
    """solve_options (reference src/solve.f90:57-84)."""
    _fields_ = [("tracker", TrackerOptions), ("endgame", EndgameOptions),
                ("infinity_tol", f64), ("singular_cond", f64),
                ("real_tol", f64), ("dedup_tol", f64),
                ("singular_dedup_tol", f64), ("use_endgame", i32),
                ("seed", i32), ("refine_maxit", i32), ("reserved", i32)]


def default_options(**overrides):
    """The reference defaults; keyword overrides use dotted-free names of the
    solve level (seed=, use_endgame=) or tracker_/endgame_ prefixes."""
    o = SolveOptions()
    t = o.tracker
    t.step_init, t.step_max, t.step_min = 0.05, 0.1, 1.0e-14
    t.corrector_tol, t.contraction = 1.0e-10, 0.5
    t.corrector_maxit, t.grow_after, t.max_steps = 3, 5, 20000
    e = o.endgame
    e.t_start, e.radius_factor, e.radius_min = 0.1, 0.25, 1.0e-13
    e.loop_radius, e.winding_agreement, e.tol = 1.0e-5, 0.2, 1.0e-11
    e.residual_tol, e.closure_factor = 1.0e-8, 0.1
    e.nsamples, e.max_winding = 8, 16
    o.infinity_tol, o.singular_cond, o.real_tol = 1.0e-8, 1.0e10, 1.0e-8
    o.dedup_tol, o.singular_dedup_tol = 1.0e-8, 1.0e-6
    o.use_endgame, o.seed, o.refine_maxit = 1, 1, 5
    for k, v in overrides.items():
        if k.startswith("tracker_"):
            setattr(o.tracker, k[8:], v)
        elif k.startswith("endgame_"):
            setattr(o.endgame, k[8:], v)
        else:
            setattr(o, k, v)
    return o


class System(C.Structure):
    """phc_system: a homogenized sparse system (include/phc.h)."""
    _fields_ = [("n", i32), ("nterms", i32), ("maxexp", i32), ("reserved", i32),
                ("first", P), ("deg", P), ("hexps", P), ("coefs", P)]


class Homotopy(C.Structure):
    _fields_ = [("gamma", P), ("b", P), ("patch", P)]


class Endpoints(C.Structure):
    _fields_ = [("status", P), ("t", P), ("z", P), ("winding", P),
                ("eg_status", P), ("eg_error", P), ("eg_residual", P),
                ("naccepted", P), ("nrejected", P), ("nlin", P), ("nbe", P)]


class UserSystem(C.Structure):
    _fields_ = [("n", i32), ("neq", i32), ("nterms", P), ("coefs", P),
                ("exps", P)]


class SolveSummary(C.Structure):
    _fields_ = [("n", i32), ("npaths", i32), ("nsolutions", i32),
                ("nreal", i32), ("nsingular", i32), ("ninfinite", i32),
                ("nfailed", i32), ("nduplicates", i32),
                ("solutions", P), ("real_solutions", P),
                ("singular_solutions", P), ("multiplicities", P)]


class PathResults(C.Structure):
    _fields_ = [(name, P) for name in (
        "status", "t", "z", "x", "at_infinity", "singular", "is_real",
        "winding", "eg_status", "eg_error", "cond", "residual",
        "solution_index", "singular_index", "naccepted", "nrejected")]


class ServerConfig(C.Structure):
    _fields_ = [("max_resident_paths", i64), ("hbm_budget_bytes", i64),
                ("units_per_launch", i32), ("launch_us", i32)]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(P)


class UserSystemData:
    """A polynomial system as given to add_equation: per equation a list of
    complex coefficients and integer exponent vectors (n each)."""

    def __init__(self, n, equations):
        # equations: list of (coefs complex array [m], exps int array [m][n])
        self.n = int(n)
        self.equations = [(np.asarray(c, dtype=np.complex128),
                           np.asarray(e, dtype=np.int32).reshape(-1, n))
                          for c, e in equations]
        self._nterms = np.array([len(c) for c, _ in self.equations], np.int32)
        if len(self.equations):
            self._coefs = np.ascontiguousarray(
                np.concatenate([c for c, _ in self.equations]))
            self._exps = np.ascontiguousarray(
                np.concatenate([e for _, e in self.equations]), dtype=np.int32)
        else:
            self._coefs = np.zeros(1, np.complex128)
            self._exps = np.zeros((1, n), np.int32)
        self.c = UserSystem(self.n, len(self.equations), _ptr(self._nterms),
                            _ptr(self._coefs), _ptr(self._exps))

    @property
    def degrees(self):
        out = []
        for c, e in self.equations:
            keep = c != 0
            out.append(int(e[keep].sum(axis=1).max()) if keep.any() else 0)
        return out

    @property
    def bezout(self):
        b = 1
        for d in self.degrees:
            b *= d
        return b

    def packed(self):
        """Homogenized form (what add_equation stores, src/sparse.f90:119-138):
        zero terms dropped, w exponent d_i - |a| appended."""
        first, deg, hexps, coefs = [0], [], [], []
        for c, e in self.equations:
            keep = c != 0
            d = int(e[keep].sum(axis=1).max())
            deg.append(d)
            for ck, ek in zip(c[keep], e[keep]):
                hexps.append(list(ek) + [d - int(ek.sum())])
                coefs.append(ck)
            first.append(len(coefs))
        return PackedSystem(self.n, np.array(first, np.int32),
                            np.array(deg, np.int32),
                            np.array(hexps, np.int32).reshape(-1, self.n + 1),
                            np.array(coefs, np.complex128))


class PackedSystem:
    """Arrays of a phc_system plus the ctypes struct pointing at them."""

    def __init__(self, n, first, deg, hexps, coefs):
        self.n = n
        self.first = np.ascontiguousarray(first, np.int32)
        self.deg = np.ascontiguousarray(deg, np.int32)
        self.hexps = np.ascontiguousarray(hexps, np.int32)
        self.coefs = np.ascontiguousarray(coefs, np.complex128)
        self.nterms = len(self.coefs)
        self.maxexp = int(self.hexps.max()) if self.nterms else 0
        self.c = System(n, self.nterms, self.maxexp, 0, _ptr(self.first),
                        _ptr(self.deg), _ptr(self.hexps), _ptr(self.coefs))


class HomotopyData:
    def __init__(self, gamma, b, patch):
        self.gamma = np.array([gamma], np.complex128)
        self.b = np.ascontiguousarray(b, np.complex128)
        self.patch = np.ascontiguousarray(patch, np.complex128)
        self.c = Homotopy(_ptr(self.gamma), _ptr(self.b), _ptr(self.patch))


class EndpointArrays:
    """Caller-owned struct-of-arrays for phc_endpoints."""

    def __init__(self, npaths, N):
        self.npaths, self.N = int(npaths), int(N)
        m = max(self.npaths, 1)
        self.status = np.full(m, -99, np.int32)
        self.t = np.zeros(m)
        self.z = np.zeros((m, N), np.complex128)
        self.winding = np.zeros(m, np.int32)
        self.eg_status = np.zeros(m, np.int32)
        self.eg_error = np.zeros(m)
        self.eg_residual = np.zeros(m)
        self.naccepted = np.zeros(m, np.int32)
        self.nrejected = np.zeros(m, np.int32)
        self.nlin = np.zeros(m, np.int32)
        self.nbe = np.zeros(m, np.int32)
        self.c = Endpoints(*[_ptr(getattr(self, f)) for f, _ in Endpoints._fields_])


class SolveArrays:
    """Caller-owned outputs of a full solve (summary + classified paths)."""

    def __init__(self, npaths, n):
        m = max(int(npaths), 1)
        self.n = n
        self.solutions = np.zeros((m, n), np.complex128)
        self.real_solutions = np.zeros((m, n))
        self.singular_solutions = np.zeros((m, n), np.complex128)
        self.multiplicities = np.zeros(m, np.int32)
        self.summary = SolveSummary(0, 0, 0, 0, 0, 0, 0, 0,
                                    _ptr(self.solutions),
                                    _ptr(self.real_solutions),
                                    _ptr(self.singular_solutions),
                                    _ptr(self.multiplicities))
        self.p = {}
        for name, _ in PathResults._fields_:
            if name in ("t", "eg_error", "cond", "residual"):
                a = np.zeros(m)
            elif name == "z":
                a = np.zeros((m, n + 1), np.complex128)
            elif name == "x":
                a = np.zeros((m, n), np.complex128)
            else:
                a = np.zeros(m, np.int32)
            self.p[name] = a
        self.paths = PathResults(*[_ptr(self.p[f]) for f, _ in PathResults._fields_])

    def counts(self):
        s = self.summary
        return dict(n=s.n, npaths=s.npaths, nsolutions=s.nsolutions,
                    nreal=s.nreal, nsingular=s.nsingular,
                    ninfinite=s.ninfinite, nfailed=s.nfailed,
                    nduplicates=s.nduplicates,
                    multiplicities=list(self.multiplicities[:s.nsingular]))

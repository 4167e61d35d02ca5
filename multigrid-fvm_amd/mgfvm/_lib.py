"""ctypes binding of libmgfvm.so (the C ABI in include/mgfvm.h).

This is a thin host-side mirror for Python callers: every call goes straight
to the native library, which runs the HIP kernels.  There is no fallback: if
the library or a GPU is missing, the calls raise.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence, Tuple

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libmgfvm.so")

# declared entry points (kept in sync with include/mgfvm.h; tests check both)
EXPORTS = [
    "mgf_problem_read", "mgf_problem_free", "mgf_options_default", "mgf_error_string",
    "mgf_abi_version", "mgf_create", "mgf_destroy", "mgf_last_error", "mgf_num_levels",
    "mgf_level_dims", "mgf_stream", "mgf_relax", "mgf_restrict", "mgf_prolong",
    "mgf_vcycle", "mgf_solve", "mgf_get_field", "mgf_set_field", "mgf_get_fields",
    "mgf_time_kernel", "mgf_reset", "mgf_synchronize", "mgf_batch_create", "mgf_batch_size",
    "mgf_batch_num_levels", "mgf_batch_level_dims", "mgf_batch_solve", "mgf_batch_history",
    "mgf_batch_get_field", "mgf_pattern_bernoulli", "mgf_pattern_staggered",
    "mgf_pattern_particles", "mgf_set_inlet_profile", "mgf_solve_noalloc", "mgf_autotune",
]

PREC_FP64 = 0
PREC_MIXED = 1
KERNEL_SWEEP_PP = 0
KERNEL_SWEEP_U = 1
KERNEL_RELAX_U = 2
KERNEL_RELAX_PP = 3
KERNEL_RELAX_UVS = 4


class MgfError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mgfvm error {code}: {msg}")
        self.code = code


class Options(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("device", ctypes.c_int), ("strict_limits", ctypes.c_int),
                ("sweeps_per_launch", ctypes.c_int), ("sweep_threads", ctypes.c_int),
                ("split_assembly", ctypes.c_int), ("gs_order", ctypes.c_int),
                ("compact", ctypes.c_int), ("lanes", ctypes.c_int),
                ("tuning", ctypes.c_char_p)]


class ProblemStruct(ctypes.Structure):
    _fields_ = [
        ("nfx", ctypes.c_int), ("nfy", ctypes.c_int), ("ngrid", ctypes.c_int),
        ("ig_obs", ctypes.c_int), ("ibeg_obs", ctypes.c_int), ("ig_inlt", ctypes.c_int),
        ("ig_outlt", ctypes.c_int), ("nvcycle", ctypes.c_int), ("nswpm", ctypes.c_int),
        ("nswpp", ctypes.c_int), ("isolve", ctypes.c_int), ("omega", ctypes.c_double),
        ("convergence", ctypes.c_double), ("xl", ctypes.c_double), ("yl", ctypes.c_double),
        ("amuf", ctypes.c_double), ("prsc", ctypes.c_double), ("relxm", ctypes.c_double),
        ("relxp", ctypes.c_double), ("ibc", ctypes.c_int * 4),
        ("bc", (ctypes.c_double * 4) * 4), ("guess", ctypes.c_double * 4),
        ("sc_obs", ctypes.c_double), ("nobs", ctypes.c_int),
        ("obs", ctypes.POINTER(ctypes.c_int)), ("ninlt", ctypes.c_int),
        ("noutlt", ctypes.c_int), ("inlt_j", ctypes.POINTER(ctypes.c_int)),
        ("outlt_j", ctypes.POINTER(ctypes.c_int)),
        ("inlt_v", ctypes.POINTER(ctypes.c_double)),
        ("outlt_v", ctypes.POINTER(ctypes.c_double)),
        ("dpdxm", ctypes.c_double), ("dpdym", ctypes.c_double),
        ("inlet_u", ctypes.POINTER(ctypes.c_double)),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"native library missing: {LIB_PATH}; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, dp, cp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_char_p
        P_i, P_d = ctypes.POINTER(ip), ctypes.POINTER(dp)
        L.mgf_problem_read.argtypes = [cp, cp, ctypes.POINTER(ctypes.POINTER(ProblemStruct))]
        L.mgf_problem_free.argtypes = [ctypes.POINTER(ProblemStruct)]
        L.mgf_options_default.argtypes = [ctypes.POINTER(Options)]
        L.mgf_error_string.restype = cp
        L.mgf_create.argtypes = [ctypes.POINTER(ProblemStruct), ctypes.POINTER(Options), ctypes.POINTER(vp)]
        L.mgf_destroy.argtypes = [vp]
        L.mgf_last_error.argtypes = [vp]
        L.mgf_last_error.restype = cp
        L.mgf_num_levels.argtypes = [vp]
        L.mgf_level_dims.argtypes = [vp, ip, P_i, P_i]
        L.mgf_stream.argtypes = [vp]
        L.mgf_stream.restype = vp
        L.mgf_relax.argtypes = [vp, ip, ip, P_d]
        L.mgf_restrict.argtypes = [vp, ip]
        L.mgf_prolong.argtypes = [vp, ip]
        L.mgf_vcycle.argtypes = [vp, ip, P_d]
        L.mgf_solve.argtypes = [vp, ip, dp, P_d, P_i, P_i]
        L.mgf_get_field.argtypes = [vp, ip, cp, P_d]
        L.mgf_set_field.argtypes = [vp, ip, cp, P_d]
        L.mgf_time_kernel.argtypes = [vp, ip, ip, ip, P_d, P_d]
        L.mgf_autotune.argtypes = [vp, ip, ctypes.c_char_p, ctypes.c_size_t]
        L.mgf_reset.argtypes = [vp]
        L.mgf_synchronize.argtypes = [vp]
        L.mgf_set_inlet_profile.argtypes = [vp, ip, P_d]
        L.mgf_solve_noalloc.argtypes = [ctypes.POINTER(ProblemStruct), ctypes.POINTER(Options), vp,
                                        P_d, P_i, P_i, P_d, P_d, P_d, P_d]
        PP = ctypes.POINTER(ProblemStruct)
        L.mgf_batch_create.argtypes = [ctypes.POINTER(PP), ip, ctypes.POINTER(Options),
                                       ctypes.POINTER(vp)]
        L.mgf_batch_size.argtypes = [vp]
        L.mgf_batch_num_levels.argtypes = [vp, ip]
        L.mgf_batch_level_dims.argtypes = [vp, ip, ip, P_i, P_i]
        L.mgf_batch_solve.argtypes = [vp, P_i, P_i]
        L.mgf_batch_history.argtypes = [vp, ip, P_d, ip]
        L.mgf_batch_get_field.argtypes = [vp, ip, ip, cp, P_d]
        u64 = ctypes.c_uint64
        L.mgf_pattern_bernoulli.argtypes = [ip, dp, u64, ip, P_i, ip]
        L.mgf_pattern_staggered.argtypes = [ip, ip, ip, ip, ip, u64, ip, P_i, ip]
        L.mgf_pattern_particles.argtypes = [ip, dp, dp, dp, dp, u64, ip, P_i, ip]
        _lib = L
    return _lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def read_problem(inputdata: str, pattern: str):
    """Parse the file pair with the library's own reader; returns an owned
    pointer to be released with free_problem()."""
    L = lib()
    p = ctypes.POINTER(ProblemStruct)()
    rc = L.mgf_problem_read(inputdata.encode(), pattern.encode(), ctypes.byref(p))
    if rc != 0:
        raise MgfError(rc, L.mgf_error_string().decode())
    return p


def free_problem(p):
    lib().mgf_problem_free(p)


def _profile_ptr(inlet_u, nfy: int):
    """A pointer to nfy inlet values (it keeps the array alive), or NULL."""
    if inlet_u is None:
        return ctypes.POINTER(ctypes.c_double)()
    a = np.ascontiguousarray(inlet_u, dtype=np.float64)
    if a.size != nfy:
        raise ValueError(f"inlet profile needs {nfy} values")
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class InletProfile:
    """Attach an inlet u profile (nfy values, rows 1..nfy of the level-0 west
    face) to a problem read by the library, for as long as the `with` block
    runs; the problem never owns the buffer."""

    def __init__(self, prob, inlet_u):
        self.prob = prob
        self.buf = None
        if inlet_u is not None:
            a = np.ascontiguousarray(inlet_u, dtype=np.float64)
            if a.size != prob.contents.nfy:
                raise ValueError(f"inlet profile needs {prob.contents.nfy} values")
            self.buf = (ctypes.c_double * a.size)(*a.tolist())

    def __enter__(self):
        if self.buf is not None:
            self.prob.contents.inlet_u = ctypes.cast(self.buf, ctypes.POINTER(ctypes.c_double))
        return self.prob

    def __exit__(self, *a):
        self.prob.contents.inlet_u = ctypes.POINTER(ctypes.c_double)()


class Solver:
    """One problem resident on one GPU."""

    def __init__(self, inputdata: str, pattern: str, precision: int = PREC_FP64,
                 use_graph: bool = True, device: int = -1, strict_limits: bool = False,
                 sweeps_per_launch: int = 0, sweep_threads: int = 0,
                 split_assembly: bool = False, inlet_u=None, gs_lex: bool = False):
        L = lib()
        prob = read_problem(inputdata, pattern)
        try:
            with InletProfile(prob, inlet_u):
                self._init_from(prob, precision, use_graph, device, strict_limits,
                                sweeps_per_launch, sweep_threads, split_assembly, gs_lex)
        finally:
            free_problem(prob)

    def _init_from(self, prob, precision, use_graph, device, strict_limits, sweeps_per_launch=0,
                   sweep_threads=0, split_assembly=False, gs_lex=False):
        L = lib()
        opt = Options()
        L.mgf_options_default(ctypes.byref(opt))
        opt.precision = precision
        opt.use_graph = 1 if use_graph else 0
        opt.device = device
        opt.strict_limits = 1 if strict_limits else 0
        opt.sweeps_per_launch = sweeps_per_launch
        opt.sweep_threads = sweep_threads
        opt.split_assembly = 1 if split_assembly else 0
        opt.gs_order = 1 if gs_lex else 0
        h = ctypes.c_void_p()
        rc = L.mgf_create(prob, ctypes.byref(opt), ctypes.byref(h))
        if rc != 0:
            raise MgfError(rc, L.mgf_error_string().decode())
        self.h = h
        self.ngrid = L.mgf_num_levels(h)
        self.nvcycle = prob.contents.nvcycle
        self.convergence = prob.contents.convergence

    def _check(self, rc: int):
        if rc != 0:
            raise MgfError(rc, lib().mgf_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().mgf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def dims(self, ig: int) -> Tuple[int, int]:
        a, b = ctypes.c_int(), ctypes.c_int()
        self._check(lib().mgf_level_dims(self.h, ig, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def stream(self) -> int:
        return lib().mgf_stream(self.h)

    def get(self, ig: int, name: str) -> np.ndarray:
        nx, ny = self.dims(ig)
        out = np.empty(nx * ny, np.float64)
        self._check(lib().mgf_get_field(self.h, ig, name.encode(), _dptr(out)))
        return out.reshape(ny, nx)

    def set(self, ig: int, name: str, arr: np.ndarray):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        self._check(lib().mgf_set_field(self.h, ig, name.encode(), _dptr(a)))

    def reset(self):
        """Back to the initial state, on the device (asynchronous)."""
        self._check(lib().mgf_reset(self.h))

    def set_inlet_profile(self, inlet_u=None):
        """A new inlet u profile (nfy values; None: the InputData inlet), and
        the state reset to the initial state built with it."""
        self._check(lib().mgf_set_inlet_profile(self.h, 0,
                                                _profile_ptr(inlet_u, self.dims(0)[1] - 2)))

    def synchronize(self):
        self._check(lib().mgf_synchronize(self.h))

    def relax(self, ig: int, leg: int) -> float:
        e = ctypes.c_double()
        self._check(lib().mgf_relax(self.h, ig, leg, ctypes.byref(e)))
        return e.value

    def restrict(self, ig: int):
        self._check(lib().mgf_restrict(self.h, ig))

    def prolong(self, igc: int):
        self._check(lib().mgf_prolong(self.h, igc))

    def vcycle(self, icyc: int) -> float:
        e = ctypes.c_double()
        self._check(lib().mgf_vcycle(self.h, icyc, ctypes.byref(e)))
        return e.value

    def solve(self, nvcycle: Optional[int] = None, tol: Optional[float] = None):
        nv = self.nvcycle if nvcycle is None else nvcycle
        tl = self.convergence if tol is None else tol
        hist = np.zeros(max(nv, 1), np.float64)
        nc, cv = ctypes.c_int(), ctypes.c_int()
        rc = lib().mgf_solve(self.h, nv, tl, _dptr(hist), ctypes.byref(nc), ctypes.byref(cv))
        if rc not in (0, -8):
            self._check(rc)
        return hist[: nc.value].copy(), nc.value, bool(cv.value)

    def time_kernel(self, kernel: int, ig: int, reps: int):
        us, b = ctypes.c_double(), ctypes.c_double()
        self._check(lib().mgf_time_kernel(self.h, kernel, ig, reps, ctypes.byref(us), ctypes.byref(b)))
        return us.value, b.value


class Batch:
    """Many independent problems solved together on one GPU."""

    def __init__(self, pairs: Sequence[Tuple[str, str]], precision: int = PREC_FP64,
                 use_graph: bool = True, device: int = -1, sweeps_per_launch: int = 0,
                 sweep_threads: int = 0, split_assembly: bool = False, compact: bool = True,
                 gs_lex: bool = False, lanes: int = 0, tuning: Optional[str] = None):
        L = lib()
        probs = [read_problem(a, b) for a, b in pairs]
        tuning_b = tuning.encode() if tuning else None
        try:
            arr = (ctypes.POINTER(ProblemStruct) * len(probs))(*probs)
            opt = Options()
            L.mgf_options_default(ctypes.byref(opt))
            opt.precision = precision
            opt.use_graph = 1 if use_graph else 0
            opt.device = device
            opt.sweeps_per_launch = sweeps_per_launch
            opt.sweep_threads = sweep_threads
            opt.split_assembly = 1 if split_assembly else 0
            opt.compact = 1 if compact else 0
            opt.gs_order = 1 if gs_lex else 0
            opt.lanes = lanes
            opt.tuning = tuning_b
            h = ctypes.c_void_p()
            rc = L.mgf_batch_create(arr, len(probs), ctypes.byref(opt), ctypes.byref(h))
            if rc != 0:
                raise MgfError(rc, L.mgf_error_string().decode())
            self.h = h
            self.n = len(probs)
        finally:
            for p in probs:
                free_problem(p)

    def _check(self, rc: int):
        if rc != 0:
            raise MgfError(rc, lib().mgf_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().mgf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def autotune(self, reps: int = 20) -> str:
        """mgf_autotune: measure and apply the launch shapes; returns the
        table (the text mgf_options.tuning / tuning= reads)."""
        buf = ctypes.create_string_buffer(8192)
        n = lib().mgf_autotune(self.h, reps, buf, len(buf))
        if n < 0:
            self._check(n)
        return buf.value.decode()

    def reset(self):
        self._check(lib().mgf_reset(self.h))

    def set_inlet_profile(self, b: int, inlet_u=None):
        """Problem b's inlet u profile (None: its InputData inlet); problem b
        is reset to the initial state built with it."""
        self._check(lib().mgf_set_inlet_profile(self.h, b,
                                                _profile_ptr(inlet_u, self.dims(b, 0)[1] - 2)))

    def synchronize(self):
        self._check(lib().mgf_synchronize(self.h))

    def solve(self):
        nc = (ctypes.c_int * self.n)()
        cv = (ctypes.c_int * self.n)()
        rc = lib().mgf_batch_solve(self.h, nc, cv)
        if rc not in (0, -8):
            self._check(rc)
        return list(nc), [bool(x) for x in cv]

    def history(self, b: int) -> np.ndarray:
        cap = 1 << 16
        out = np.zeros(cap, np.float64)
        n = lib().mgf_batch_history(self.h, b, _dptr(out), cap)
        if n < 0:
            self._check(n)
        return out[:n].copy()

    def num_levels(self, b: int) -> int:
        return lib().mgf_batch_num_levels(self.h, b)

    def dims(self, b: int, ig: int) -> Tuple[int, int]:
        a, c = ctypes.c_int(), ctypes.c_int()
        self._check(lib().mgf_batch_level_dims(self.h, b, ig, ctypes.byref(a), ctypes.byref(c)))
        return a.value, c.value

    def get(self, b: int, ig: int, name: str) -> np.ndarray:
        nx, ny = self.dims(b, ig)
        out = np.empty(nx * ny, np.float64)
        self._check(lib().mgf_batch_get_field(self.h, b, ig, name.encode(), _dptr(out)))
        return out.reshape(ny, nx)

    def time_kernel(self, kernel: int, ig: int, reps: int):
        """Mean microseconds and algorithmic bytes of one smoother-sweep
        launch over every member's level at problem 0's depth of level ig."""
        us, b = ctypes.c_double(), ctypes.c_double()
        self._check(lib().mgf_time_kernel(self.h, kernel, ig, reps, ctypes.byref(us),
                                          ctypes.byref(b)))
        return us.value, b.value


def _rects(fn, *args):
    n = fn(*args, None, 0)
    if n < 0:
        raise MgfError(n, lib().mgf_error_string().decode())
    buf = (ctypes.c_int * (4 * max(n, 1)))()
    fn(*args, buf, n)
    return [tuple(buf[4 * k:4 * k + 4]) for k in range(n)]


def pattern_bernoulli(n_base: int, porosity: float, seed: int, free_edge_columns: bool = True):
    """Rectangles of mgf_pattern_bernoulli (the native generator)."""
    return _rects(lib().mgf_pattern_bernoulli, n_base, porosity, seed, int(free_edge_columns))


def pattern_staggered(n_base: int, spacing_x: int, spacing_y: int, scale: int, variant: int,
                      seed: int, free_edge_columns: bool = True):
    """Rectangles of mgf_pattern_staggered (the native generator)."""
    return _rects(lib().mgf_pattern_staggered, n_base, spacing_x, spacing_y, scale, variant, seed,
                  int(free_edge_columns))


def pattern_particles(n_grid: int, ratio: float, pct1: float, pct2: float, pct3: float,
                      seed: int, free_edge_columns: bool = True):
    """Rectangles of mgf_pattern_particles (the native generator)."""
    return _rects(lib().mgf_pattern_particles, n_grid, ratio, pct1, pct2, pct3, seed,
                  int(free_edge_columns))


def solve_noalloc(inputdata: str, pattern: str, precision: int = PREC_FP64, stream=None,
                  device: int = -1):
    """mgf_solve_noalloc on the file pair: the whole solve in the library's
    static device scratch (no device allocation).  Returns (history,
    ncycles, converged, {"u", "v", "p", "sc": level-0 arrays})."""
    L = lib()
    p = read_problem(inputdata, pattern)
    try:
        q = p.contents
        npx, npy = q.nfx + 2, q.nfy + 2
        opt = Options()
        L.mgf_options_default(ctypes.byref(opt))
        opt.precision = precision
        opt.device = device
        hist = np.zeros(max(q.nvcycle, 1))
        out = {k: np.zeros(npx * npy) for k in ("u", "v", "p", "sc")}
        nc, cv = ctypes.c_int(0), ctypes.c_int(0)
        rc = L.mgf_solve_noalloc(p, ctypes.byref(opt), stream, _dptr(hist), ctypes.byref(nc),
                                 ctypes.byref(cv), _dptr(out["u"]), _dptr(out["v"]),
                                 _dptr(out["p"]), _dptr(out["sc"]))
        if rc != 0:
            raise MgfError(rc, L.mgf_error_string().decode())
        return hist[:nc.value], nc.value, bool(cv.value), out
    finally:
        free_problem(p)

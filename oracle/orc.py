"""ctypes binding of the CPU restatement (liborc.so) and the dump-format
reader shared with the reference harnesses.

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke entry and the CPU-baseline
leg of bench.py as the checker; the product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import struct
from typing import Dict, List, Optional, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")
RUN_PATH = os.path.join(HERE, "orc_run")

# order of the arrays in a dump (oracle/orc_main.c, oracle/ref/harness_*.c)
DUMP_NAMES = ["u", "v", "p", "pp", "sc", "su", "rs", "resu", "resv", "ressc",
              "c", "ae", "aw", "an", "as", "ap", "apu", "apv", "spu", "spv",
              "spsc", "cx", "cy", "areax", "areay", "tag"]
FIELDS = ["u", "v", "p", "pp", "sc"]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run build())")
        L = ctypes.CDLL(LIB_PATH)
        vp, cp, ip, dp = ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_double
        L.orc_last_error.restype = cp
        L.orc_create_from_files.argtypes = [cp, cp, ctypes.POINTER(vp)]
        L.orc_create_from_files_inlet.argtypes = [cp, cp, ctypes.POINTER(dp), ctypes.POINTER(vp)]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_verbose.argtypes = [vp, ip]
        L.orc_set_gs_lex.argtypes = [vp, ip]
        L.orc_num_levels.argtypes = [vp]
        L.orc_level_dims.argtypes = [vp, ip, ctypes.POINTER(ip), ctypes.POINTER(ip)]
        L.orc_get.argtypes = [vp, ip, cp, ctypes.POINTER(dp)]
        L.orc_set.argtypes = [vp, ip, cp, ctypes.POINTER(dp)]
        L.orc_relax.argtypes = [vp, ip, ip]
        L.orc_relax.restype = dp
        L.orc_restrict.argtypes = [vp, ip]
        L.orc_prolong.argtypes = [vp, ip]
        L.orc_vcycle.argtypes = [vp, ip]
        L.orc_vcycle.restype = dp
        L.orc_solve.argtypes = [vp, ip, dp, ctypes.POINTER(dp), ctypes.POINTER(ip), ctypes.POINTER(ip)]
        L.orc_call.argtypes = [vp, cp, ip, ctypes.POINTER(dp)]
        _lib = L
    return _lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """CPU restatement of the reference solver on one InputData/pattern pair."""

    def __init__(self, inputdata: str, pattern: str, gs_lex: bool = False, inlet_u=None):
        """inlet_u: optional level-0 inlet u profile (nfy values, rows 1..nfy),
        set after the initial state and carried to the coarse levels as the
        boundary restriction carries it (mgfvm_oracle.c init/rest_geom)."""
        L = lib()
        h = ctypes.c_void_p()
        if inlet_u is not None:
            prof = np.ascontiguousarray(inlet_u, dtype=np.float64)
            rc = L.orc_create_from_files_inlet(inputdata.encode(), pattern.encode(), _dptr(prof),
                                               ctypes.byref(h))
        else:
            rc = L.orc_create_from_files(inputdata.encode(), pattern.encode(), ctypes.byref(h))
        if rc != 0:
            raise RuntimeError(f"oracle create failed ({rc}): {L.orc_last_error().decode()}")
        self.h = h
        if gs_lex:
            L.orc_set_gs_lex(h, 1)
        self.ngrid = L.orc_num_levels(h)

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def dims(self, ig: int) -> Tuple[int, int]:
        a, b = ctypes.c_int(), ctypes.c_int()
        lib().orc_level_dims(self.h, ig, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def get(self, ig: int, name: str) -> np.ndarray:
        nx, ny = self.dims(ig)
        out = np.empty(nx * ny, np.float64)
        if lib().orc_get(self.h, ig, name.encode(), _dptr(out)) != 0:
            raise KeyError(name)
        return out.reshape(ny, nx)

    def set(self, ig: int, name: str, arr: np.ndarray):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        if lib().orc_set(self.h, ig, name.encode(), _dptr(a)) != 0:
            raise KeyError(name)

    def relax(self, ig: int, leg: int) -> float:
        return lib().orc_relax(self.h, ig, leg)

    def restrict(self, ig: int):
        lib().orc_restrict(self.h, ig)

    def prolong(self, igc: int):
        lib().orc_prolong(self.h, igc)

    def vcycle(self, icyc: int) -> float:
        return lib().orc_vcycle(self.h, icyc)

    def solve(self, nvcycle: int, tol: float):
        hist = np.zeros(max(nvcycle, 1), np.float64)
        nc, cv = ctypes.c_int(), ctypes.c_int()
        lib().orc_solve(self.h, nvcycle, tol, _dptr(hist), ctypes.byref(nc), ctypes.byref(cv))
        return hist[: nc.value].copy(), nc.value, bool(cv.value)

    def call(self, fn: str, ig: int) -> float:
        r = ctypes.c_double()
        rc = lib().orc_call(self.h, fn.encode(), ig, ctypes.byref(r))
        if rc != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        return r.value


# ----------------------------------------------------------------------
# dump readers
# ----------------------------------------------------------------------
def _read_levels(buf: bytes, off: int, ngrid: int):
    levels = []
    for _ in range(ngrid):
        npx, npy = struct.unpack_from("<ii", buf, off)
        off += 8
        n = npx * npy
        lv = {}
        for name in DUMP_NAMES:
            lv[name] = np.frombuffer(buf, np.float64, n, off).reshape(npy, npx).copy()
            off += 8 * n
        levels.append(lv)
    return levels, off


def read_dump(path: str):
    """Whole-solve dump (orc_run / ref_a): (levels, history)."""
    with open(path, "rb") as f:
        buf = f.read()
    assert buf[:4] == b"MGFD", "bad dump magic"
    ver, ngrid, nh = struct.unpack_from("<iii", buf, 4)
    levels, off = _read_levels(buf, 16, ngrid)
    hist = np.frombuffer(buf, np.float64, nh, off).copy()
    return levels, hist


def read_kat(path: str, ngrid: int):
    """Per-routine records written by ref_b: list of (name, level, ret, levels)."""
    with open(path, "rb") as f:
        buf = f.read()
    off, recs = 0, []
    while off < len(buf):
        name = buf[off:off + 32].split(b"\0", 1)[0].decode()
        level, = struct.unpack_from("<i", buf, off + 32)
        ret, = struct.unpack_from("<d", buf, off + 36)
        levels, off = _read_levels(buf, off + 44, ngrid)
        recs.append((name, level, ret, levels))
    return recs

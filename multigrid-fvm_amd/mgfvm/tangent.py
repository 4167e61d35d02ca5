"""Forward sensitivities of the converged steady solution: ctypes binding of
libmgfvm_tangent.so (include/mgfvm_tangent.h), the FP64 solver built with a
dual-number scalar, and a torch.autograd.Function over it.

A batch of n copies of one problem, each seeded with one input direction,
gives the derivative of the objectives J1 (pressure drop between the first
and last interior columns) and J2 (outlet scalar flux) along n directions in
one solve.  `gradients()` builds that batch for the inlet u profile (every
row), amuf and prsc.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

from ._lib import PKG_DIR, InletProfile, MgfError, Options, ProblemStruct

TAN_PATH = os.path.join(PKG_DIR, "libmgfvm_tangent.so")

EXPORTS = ["mgf_tangent_batch_create", "mgf_tangent_get_field", "mgf_tangent_objectives",
           "mgf_tangent_gradient"]


class Seed(ctypes.Structure):
    _fields_ = [("d_amuf", ctypes.c_double), ("d_prsc", ctypes.c_double),
                ("d_inlet_u", ctypes.POINTER(ctypes.c_double)), ("d_u_xp", ctypes.c_double)]


_tlib = None


def tlib():
    global _tlib
    if _tlib is None:
        if not os.path.exists(TAN_PATH):
            raise RuntimeError(f"native library missing: {TAN_PATH}; run __graft_entry__.build()")
        L = ctypes.CDLL(TAN_PATH)
        vp, ip, dp, cp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_char_p
        P_i, P_d = ctypes.POINTER(ip), ctypes.POINTER(dp)
        PP = ctypes.POINTER(ProblemStruct)
        L.mgf_problem_read.argtypes = [cp, cp, ctypes.POINTER(PP)]
        L.mgf_problem_free.argtypes = [PP]
        L.mgf_options_default.argtypes = [ctypes.POINTER(Options)]
        L.mgf_error_string.restype = cp
        L.mgf_last_error.argtypes = [vp]
        L.mgf_last_error.restype = cp
        L.mgf_destroy.argtypes = [vp]
        L.mgf_batch_solve.argtypes = [vp, P_i, P_i]
        L.mgf_batch_history.argtypes = [vp, ip, P_d, ip]
        L.mgf_batch_level_dims.argtypes = [vp, ip, ip, P_i, P_i]
        L.mgf_tangent_batch_create.argtypes = [ctypes.POINTER(PP), ctypes.POINTER(Seed), ip,
                                               ctypes.POINTER(Options), ctypes.POINTER(vp)]
        L.mgf_tangent_get_field.argtypes = [vp, ip, ip, cp, P_d, P_d]
        L.mgf_tangent_objectives.argtypes = [vp, ip, P_d, P_d]
        L.mgf_tangent_gradient.argtypes = [PP, ctypes.POINTER(Options), ip, P_d, P_d, P_d, P_d,
                                           P_i]
        L.mgf_create.argtypes = [PP, ctypes.POINTER(Options), ctypes.POINTER(vp)]
        _tlib = L
    return _tlib


class TangentBatch:
    """n problems (pairs of InputData / pattern.txt paths) with one seed each
    (dicts with optional keys d_amuf, d_prsc, d_inlet_u); `inlet_u` sets the
    value of the inlet profile of every problem."""

    def __init__(self, pairs: Sequence, seeds: Sequence[dict], inlet_u=None, device: int = -1,
                 use_graph: bool = True):
        if len(pairs) != len(seeds):
            raise ValueError("one seed per problem")
        L = tlib()
        probs, keep = [], []
        try:
            for a, b in pairs:
                p = ctypes.POINTER(ProblemStruct)()
                rc = L.mgf_problem_read(a.encode(), b.encode(), ctypes.byref(p))
                if rc != 0:
                    raise MgfError(rc, L.mgf_error_string().decode())
                probs.append(p)
            profiles = [InletProfile(p, inlet_u) for p in probs]
            for pr in profiles:
                pr.__enter__()
            arr = (ctypes.POINTER(ProblemStruct) * len(probs))(*probs)
            sarr = (Seed * len(seeds))()
            for k, sd in enumerate(seeds):
                sarr[k].d_amuf = float(sd.get("d_amuf", 0.0))
                sarr[k].d_prsc = float(sd.get("d_prsc", 0.0))
                sarr[k].d_u_xp = float(sd.get("d_u_xp", 0.0))
                if sd.get("d_inlet_u") is not None:
                    buf = (ctypes.c_double * len(sd["d_inlet_u"]))(*map(float, sd["d_inlet_u"]))
                    keep.append(buf)
                    sarr[k].d_inlet_u = ctypes.cast(buf, ctypes.POINTER(ctypes.c_double))
            opt = Options()
            L.mgf_options_default(ctypes.byref(opt))
            opt.device = device
            opt.use_graph = 1 if use_graph else 0
            h = ctypes.c_void_p()
            rc = L.mgf_tangent_batch_create(arr, sarr, len(probs), ctypes.byref(opt),
                                            ctypes.byref(h))
            for pr in profiles:
                pr.__exit__()
            if rc != 0:
                raise MgfError(rc, L.mgf_error_string().decode())
            self.h = h
            self.n = len(probs)
        finally:
            for p in probs:
                L.mgf_problem_free(p)

    def _check(self, rc):
        if rc != 0:
            raise MgfError(rc, tlib().mgf_last_error(self.h).decode())

    def solve(self):
        nc = (ctypes.c_int * self.n)()
        cv = (ctypes.c_int * self.n)()
        rc = tlib().mgf_batch_solve(self.h, nc, cv)
        if rc not in (0, -8):
            self._check(rc)
        return list(nc), [bool(x) for x in cv]

    def objectives(self, b: int):
        J = (ctypes.c_double * 2)()
        dJ = (ctypes.c_double * 2)()
        self._check(tlib().mgf_tangent_objectives(self.h, b, J, dJ))
        return np.array(J[:]), np.array(dJ[:])

    def get(self, b: int, ig: int, name: str):
        npx, npy = ctypes.c_int(), ctypes.c_int()
        self._check(tlib().mgf_batch_level_dims(self.h, b, ig, ctypes.byref(npx),
                                                ctypes.byref(npy)))
        v = np.empty((npy.value, npx.value))
        d = np.empty_like(v)
        self._check(tlib().mgf_tangent_get_field(
            self.h, b, ig, name.encode(), v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            d.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return v, d

    def close(self):
        if getattr(self, "h", None):
            tlib().mgf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gradients(inputdata: str, pattern: str, inlet_u=None, rows: Optional[Sequence[int]] = None,
              device: int = -1):
    """J = (J1, J2) and its gradient with respect to the inlet profile
    (rows 1..nfy, or the given 1-based rows), amuf and prsc, from one
    batched tangent solve.  An inlet-row direction also raises the east
    boundary u by 1/nfy, so inflow and outflow stay balanced (a fixed east u
    admits a steady solution only then).  Returns (J, dJ_dinlet [2 x rows],
    dJ_damuf [2], dJ_dprsc [2], converged)."""
    from .inputs import parse
    nfy = parse(inputdata, pattern).nfy
    rows = list(range(1, nfy + 1)) if rows is None else list(rows)
    seeds = []
    for r in rows:
        e = np.zeros(nfy)
        e[r - 1] = 1.0
        seeds.append({"d_inlet_u": e, "d_u_xp": 1.0 / nfy})
    seeds.append({"d_amuf": 1.0})
    seeds.append({"d_prsc": 1.0})
    tb = TangentBatch([(inputdata, pattern)] * len(seeds), seeds, inlet_u=inlet_u, device=device)
    try:
        _, conv = tb.solve()
        outs = [tb.objectives(b) for b in range(len(seeds))]
    finally:
        tb.close()
    J = outs[0][0]
    dJ = np.array([o[1] for o in outs])            # directions x 2
    return J, dJ[:len(rows)].T, dJ[len(rows)], dJ[len(rows) + 1], all(conv)


def gradient_c(inputdata: str, pattern: str, inlet_u=None, max_batch: int = 0,
               device: int = -1):
    """The same gradient through the library's own entry point,
    mgf_tangent_gradient (every inlet row, amuf, prsc; batches of at most
    max_batch directions, 0 = one batch).  Returns (J, dJ_dinlet [2 x nfy],
    dJ_damuf [2], dJ_dprsc [2], converged)."""
    L = tlib()
    p = ctypes.POINTER(ProblemStruct)()
    rc = L.mgf_problem_read(inputdata.encode(), pattern.encode(), ctypes.byref(p))
    if rc != 0:
        raise MgfError(rc, L.mgf_error_string().decode())
    try:
        nfy = p.contents.nfy
        opt = Options()
        L.mgf_options_default(ctypes.byref(opt))
        opt.device = device
        J = (ctypes.c_double * 2)()
        gi = (ctypes.c_double * (2 * nfy))()
        ga = (ctypes.c_double * 2)()
        gp = (ctypes.c_double * 2)()
        conv = ctypes.c_int()
        with InletProfile(p, inlet_u):
            rc = L.mgf_tangent_gradient(p, ctypes.byref(opt), max_batch, J, gi, ga, gp,
                                        ctypes.byref(conv))
        if rc != 0:
            raise MgfError(rc, L.mgf_error_string().decode())
        return (np.array(J[:]), np.array(gi[:]).reshape(2, nfy), np.array(ga[:]),
                np.array(gp[:]), bool(conv.value))
    finally:
        L.mgf_problem_free(p)


def _problem_files(problem, inlet_u, amuf, prsc, directory):
    """Write the problem with the given amuf, prsc and an east boundary u
    equal to the mean inlet velocity (inflow = outflow)."""
    bc = tuple(problem.bc_xp)
    p = problem.replace(amuf=float(amuf), prsc=float(prsc),
                        bc_xp=(float(np.mean(inlet_u)), bc[1], bc[2], bc[3]))
    return p.write(directory)


def steady_objectives(inlet_u, amuf, prsc, problem, device: int = -1):
    """J = (J1, J2) of `problem` (an inputs.Problem) with the given inlet
    profile, amuf and prsc, and its gradient, from one batched tangent
    solve.  Returns (J, dJ_dinlet [2 x nfy], dJ_damuf [2], dJ_dprsc [2],
    converged)."""
    import tempfile
    with tempfile.TemporaryDirectory(prefix="mgf_tan_") as d:
        pin, ppat = _problem_files(problem, inlet_u, amuf, prsc, d)
        return gradients(pin, ppat, inlet_u=np.asarray(inlet_u, dtype=np.float64), device=device)


def torch_function():
    """The forward-mode torch.autograd.Function over steady_objectives:
    forward(inlet_u [nfy], amuf, prsc, problem) -> J [2] (float64, on
    inlet_u's device) computes the whole Jacobian in the same batched solve
    (nfy + 2 directions); backward returns its vector-Jacobian products for
    the tensor inputs that need them (None for float hyperparameters), each
    on its input's device and dtype.  The reverse-mode Function, which
    solves only the primal in forward and one adjoint per backward, is
    mgfvm.adjoint.torch_function."""
    import torch

    class SteadyObjectives(torch.autograd.Function):
        @staticmethod
        def forward(ctx, inlet_u, amuf, prsc, problem):
            J, d_in, d_mu, d_pr, conv = steady_objectives(
                inlet_u.detach().cpu().double().numpy(), float(amuf), float(prsc), problem)
            if not conv:
                raise RuntimeError("steady solve did not converge")
            ctx.save_for_backward(torch.as_tensor(d_in), torch.as_tensor(d_mu),
                                  torch.as_tensor(d_pr))
            ctx.like = (inlet_u, amuf if torch.is_tensor(amuf) else None,
                        prsc if torch.is_tensor(prsc) else None)
            return torch.as_tensor(J, dtype=torch.float64).to(inlet_u.device)

        @staticmethod
        def backward(ctx, g):
            d_in, d_mu, d_pr = ctx.saved_tensors
            g = g.detach().to("cpu", torch.float64)
            outs = [None, None, None, None]
            for k, (d, t) in enumerate(zip((d_in, d_mu, d_pr), ctx.like)):
                if t is None or not ctx.needs_input_grad[k]:
                    continue
                v = g @ d
                outs[k] = (v if k == 0 else v.reshape(())).to(t.dtype).to(t.device)
            return tuple(outs)

    return SteadyObjectives

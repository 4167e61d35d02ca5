"""ctypes binding of the reverse-mode adjoint (include/mgfvm_adjoint.h, in
libmgfvm.so) and its torch.autograd.Function.

The adjoint runs on a solver's converged state: solve first, then ask for
the gradient of one objective (0: pressure drop J1, 1: outlet scalar flux
J2) with respect to the inlet u profile (every row; a row direction also
raises the east boundary u by 1/nfy), amuf and prsc.  One reverse solve per
objective, on the device, for every problem of a batch at once.
"""
from __future__ import annotations

import ctypes
import os
import tempfile
from typing import Optional

import numpy as np

from . import _lib
from ._lib import MgfError, _dptr, lib

EXPORTS = ["mgf_adjoint_opts_default", "mgf_objectives", "mgf_adjoint", "mgf_batch_adjoint",
           "mgf_adjoint_get_state", "mgf_adjoint_apply"]

OBJ_PRESSURE_DROP = 0
OBJ_OUTLET_FLUX = 1


class AdjointOpts(ctypes.Structure):
    _fields_ = [("rtol", ctypes.c_double), ("max_iter", ctypes.c_int), ("restart", ctypes.c_int),
                ("only_converged", ctypes.c_int), ("single_colour", ctypes.c_int),
                ("precond_cycles", ctypes.c_int)]


_bound = False


def _bind():
    global _bound
    if _bound:
        return lib()
    L = lib()
    vp, ip = ctypes.c_void_p, ctypes.c_int
    P_d, P_i = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    L.mgf_adjoint_opts_default.argtypes = [ctypes.POINTER(AdjointOpts)]
    L.mgf_objectives.argtypes = [vp, ip, P_d]
    L.mgf_adjoint.argtypes = [vp, ip, P_d, P_d, P_d]
    L.mgf_batch_adjoint.argtypes = [vp, ip, ctypes.POINTER(AdjointOpts), P_d, P_d, P_d, P_d, P_i,
                                    P_d]
    L.mgf_adjoint_get_state.argtypes = [vp, ip, P_d, P_d, P_d, P_d]
    L.mgf_adjoint_apply.argtypes = [vp, ip, ip, P_d, P_d]
    _bound = True
    return L


def _check(h, rc, allow_nonfinite=False):
    if rc == 0 or (allow_nonfinite and rc == -8):
        return
    raise MgfError(rc, lib().mgf_last_error(h).decode())


def opts(rtol: Optional[float] = None, max_iter: Optional[int] = None,
         restart: Optional[int] = None, only_converged: Optional[bool] = None,
         single_colour: Optional[bool] = None,
         precond_cycles: Optional[int] = None) -> AdjointOpts:
    o = AdjointOpts()
    _bind().mgf_adjoint_opts_default(ctypes.byref(o))
    if rtol is not None:
        o.rtol = rtol
    if max_iter is not None:
        o.max_iter = max_iter
    if restart is not None:
        o.restart = restart
    if only_converged is not None:
        o.only_converged = int(only_converged)
    if single_colour is not None:
        o.single_colour = int(single_colour)
    if precond_cycles is not None:
        o.precond_cycles = int(precond_cycles)
    return o


def objectives(solver, b: int = 0):
    J = np.zeros(2)
    _check(solver.h, _bind().mgf_objectives(solver.h, b, _dptr(J)))
    return J


def adjoint(solver, objective: int):
    """mgf_adjoint on a single-problem solver: (g_inlet[nfy], g_amuf,
    g_prsc)."""
    L = _bind()
    npx, npy = solver.dims(0)
    gi = np.zeros(npy - 2)
    ga, gp = ctypes.c_double(), ctypes.c_double()
    _check(solver.h, L.mgf_adjoint(solver.h, objective, _dptr(gi), ctypes.byref(ga),
                                   ctypes.byref(gp)))
    return gi, ga.value, gp.value


def batch_adjoint(batch, objective: int, options: Optional[AdjointOpts] = None,
                  allow_unconverged: bool = False):
    """mgf_batch_adjoint on a Batch (or a Solver): per problem (J, g_inlet,
    g_amuf, g_prsc, iterations, relative residual)."""
    L = _bind()
    n = L.mgf_batch_size(batch.h)
    nfy = []
    for b in range(n):
        a, c = ctypes.c_int(), ctypes.c_int()
        L.mgf_batch_level_dims(batch.h, b, 0, ctypes.byref(a), ctypes.byref(c))
        nfy.append(c.value - 2)
    J, ga, gp, rr = (np.zeros(n) for _ in range(4))
    gi = np.zeros(sum(nfy))
    it = np.zeros(n, dtype=np.int32)
    o = options if options is not None else opts()
    rc = L.mgf_batch_adjoint(batch.h, objective, ctypes.byref(o), _dptr(J), _dptr(gi), _dptr(ga),
                             _dptr(gp), it.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dptr(rr))
    _check(batch.h, rc, allow_nonfinite=allow_unconverged)
    out, off = [], 0
    for b in range(n):
        out.append((J[b], gi[off:off + nfy[b]].copy(), ga[b], gp[b], int(it[b]), rr[b]))
        off += nfy[b]
    return out


def state(solver, b: int = 0):
    L = _bind()
    a, c = ctypes.c_int(), ctypes.c_int()
    L.mgf_batch_level_dims(solver.h, b, 0, ctypes.byref(a), ctypes.byref(c))
    npx, npy = a.value, c.value
    arrs = [np.zeros(npx * npy) for _ in range(4)]
    _check(solver.h, L.mgf_adjoint_get_state(solver.h, b, *[_dptr(x) for x in arrs]))
    return [x.reshape(npy, npx) for x in arrs]


def apply(solver, b: int, op: int, vec: np.ndarray) -> np.ndarray:
    """op 0: K^T vec (exact transposed Jacobian on level 0); op 1: one
    preconditioner V-cycle on the right-hand side vec (4 * npx * npy)."""
    L = _bind()
    v = np.ascontiguousarray(vec, dtype=np.float64)
    out = np.zeros_like(v)
    _check(solver.h, L.mgf_adjoint_apply(solver.h, b, op, _dptr(v), _dptr(out)))
    return out


# ----------------------------------------------------------------------
# torch.autograd.Function
# ----------------------------------------------------------------------
def torch_function():
    """A torch.autograd.Function: forward(inlet_u, amuf, prsc, problem,
    objective) solves the steady problem (an mgfvm.inputs.Problem with a
    plug east boundary) with that inlet profile and viscosity / Prandtl
    number on the device and returns J (0-d tensor); backward runs the
    reverse adjoint on the converged state kept from forward and returns
    the vector-Jacobian products for the tensor inputs (None for the
    others).  The east boundary u follows the inlet's mean (a mass-
    consistent change), so the inlet gradient is that of J along
    mass-consistent profile changes."""
    import torch

    class SteadyObjective(torch.autograd.Function):
        @staticmethod
        def forward(ctx, inlet_u, amuf, prsc, problem, objective):
            from ._lib import Solver
            prof = inlet_u.detach().to("cpu", torch.float64).numpy()
            a = float(amuf) if not torch.is_tensor(amuf) else float(amuf.detach().cpu())
            p = float(prsc) if not torch.is_tensor(prsc) else float(prsc.detach().cpu())
            q = problem.replace(amuf=a, prsc=p,
                                bc_xp=(float(prof.mean()),) + tuple(problem.bc_xp[1:]))
            with tempfile.TemporaryDirectory() as d:
                pin, ppat = q.write(d)
                s = Solver(pin, ppat, inlet_u=prof)
            h, n, c = s.solve()
            if not c:
                s.close()
                raise MgfError(-8, "the primal solve did not converge")
            J = objectives(s)[objective]
            ctx.solver = s
            ctx.objective = objective
            ctx.inputs = (inlet_u, amuf, prsc)
            return torch.tensor(J, dtype=inlet_u.dtype, device=inlet_u.device)

        @staticmethod
        def backward(ctx, gJ):
            s = ctx.solver
            gi, ga, gp = adjoint(s, ctx.objective)
            s.close()
            ctx.solver = None
            inlet_u, amuf, prsc = ctx.inputs
            g = gJ.detach().to("cpu", torch.float64)
            out = [None, None, None, None, None]
            if ctx.needs_input_grad[0]:
                out[0] = (g * torch.from_numpy(gi)).to(inlet_u.dtype).to(inlet_u.device)
            if torch.is_tensor(amuf) and ctx.needs_input_grad[1]:
                out[1] = (g * ga).to(amuf.dtype).to(amuf.device)
            if torch.is_tensor(prsc) and ctx.needs_input_grad[2]:
                out[2] = (g * gp).to(prsc.dtype).to(prsc.device)
            return tuple(out)

    return SteadyObjective

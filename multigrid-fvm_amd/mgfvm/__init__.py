"""MI355X-native SIMPLE / FAS-multigrid solver (Python side).

The solver itself is native: host set-up in C and HIP kernels for gfx950 in
libmgfvm.so, reached through the C ABI of include/mgfvm.h.  This package
holds the ctypes binding (`Solver`) and the input-file tools (`inputs`).
"""
from . import adjoint, inputs
from ._lib import (EXPORTS, KERNEL_RELAX_PP, KERNEL_RELAX_U, KERNEL_RELAX_UVS, KERNEL_SWEEP_PP, KERNEL_SWEEP_U,
                   LIB_PATH, PREC_FP64, PREC_MIXED, Batch, MgfError, Solver, lib,
                   pattern_bernoulli, pattern_particles, pattern_staggered)

__all__ = ["adjoint", "inputs", "Solver", "Batch", "MgfError", "lib", "LIB_PATH", "EXPORTS", "PREC_FP64",
           "PREC_MIXED", "KERNEL_SWEEP_PP", "KERNEL_SWEEP_U", "KERNEL_RELAX_U",
           "KERNEL_RELAX_PP", "KERNEL_RELAX_UVS", "pattern_bernoulli", "pattern_staggered", "pattern_particles"]

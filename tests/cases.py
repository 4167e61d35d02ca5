"""Named input cases shared by the oracle, golden and GPU parity tests.

Each case is a pair of reference-format input files (InputData,
pattern.txt).  Together they exercise every branch of the solved path:
Jacobi and red-black smoothers, odd sweep counts (Jacobi copy-back),
obstacles that touch and overlap (ordered pressure copies and diffusion
adds), dx != dy (unequal x/y diffusion adds), discrete inlets/outlets,
p' Dirichlet rows (ibc 2), periodic branches (ibc 3), zero-gradient outflow
(ibc 4), an obstacle-pattern offset, random masks and a staggered pattern.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "multigrid-fvm_amd"))

from mgfvm.inputs import Problem, channel_128_inline, random_mask, staggered_pattern  # noqa: E402

_base = Problem(nfx=32, nfy=16, ngrid=3, ig_obs=3, xl=2.0, yl=0.5, nvcycle=30, nswpm=2,
                nswpp=3, sc_obs=0.5,
                obstacles=[(2, 2, 1, 1), (2, 3, 2, 2), (3, 3, 2, 3), (5, 5, 3, 3),
                           (5, 5, 3, 3), (6, 6, 2, 2)])

# eight two-row bands of a piecewise-constant inlet u profile (mean 1, the
# east boundary u) as discrete inlets (jf, jl, u, v, sc)
PROFILE_INLETS = [(1, 2, 0.6, 0.0, 1.0), (3, 4, 0.9, 0.0, 1.0), (5, 6, 1.2, 0.0, 1.0),
                  (7, 8, 1.3, 0.0, 1.0), (9, 10, 1.1, 0.0, 1.0), (11, 12, 0.8, 0.0, 1.0),
                  (13, 14, 1.05, 0.0, 1.0), (15, 16, 1.05, 0.0, 1.0)]


def profile_of(inlets, nfy):
    """the level-0 inlet u profile (rows 1..nfy) the discrete inlets set"""
    prof = [None] * nfy
    for jf, jl, u, _, _ in inlets:
        for j in range(jf, jl + 1):
            prof[j - 1] = u
    assert None not in prof
    return prof


CASES = {
    "small_jac": _base,
    "small_rb": _base.replace(isolve=2),
    "ibc2": _base.replace(ibc=(2, 2, 2, 2), isolve=2),
    "ibc3": _base.replace(ibc=(3, 1, 3, 1)),
    "ibc4": _base.replace(ibc=(1, 4, 1, 1), isolve=2),
    "odd_sweeps": _base.replace(nswpm=3, nswpp=5),
    "inlets": _base.replace(ig_inlt=2, ig_outlt=3,
                            inlets=[(1, 2, 0.5, 0.1, 2.0), (3, 3, 1.5, 0.0, 1.0)],
                            outlets=[(1, 1, 0.7, 0.0, 0.3)]),
    "ibeg": _base.replace(ibeg_obs=1, ig_obs=2, ngrid=2),
    "rand64": Problem(nfx=64, nfy=64, ngrid=3, ig_obs=3, isolve=2, nvcycle=40, nswpp=6,
                      amuf=0.005, obstacles=random_mask(16, 0.8, 7)),
    "stag128": Problem(nfx=128, nfy=128, ngrid=3, ig_obs=3, isolve=2, nvcycle=20, nswpp=6,
                       obstacles=staggered_pattern(16, 4, "half", 5)[0]),
    # discrete inlets on level 0 only (ig_inlt 1) covering every row, with
    # the west ghost's v and sc: a piecewise-constant inlet u profile in the
    # reference's own input format (pins the inlet-profile path)
    "profile_inlets": _base.replace(isolve=2, ig_inlt=1, inlets=PROFILE_INLETS),
    "cfg0": channel_128_inline(),
    "cfg0_rb": channel_128_inline().replace(isolve=2),
}

# Cases the reference cannot run (ig_obs < ngrid reads uninitialised obstacle
# ranges, Version-6-26-04.c:471, 939-963): GPU vs CPU restatement only.
EXTENSION_CASES = {
    "coarse_free": Problem(nfx=64, nfy=64, ngrid=4, ig_obs=2, isolve=2, nvcycle=40, nswpp=4,
                           amuf=0.01, obstacles=random_mask(32, 0.85, 11)),
    "coarse_free_jac": Problem(nfx=64, nfy=32, ngrid=3, ig_obs=2, isolve=1, nvcycle=25,
                               xl=2.0, yl=0.5,
                               obstacles=[r for r in random_mask(32, 0.8, 12) if r[3] <= 16][:40]),
}
CASES.update(EXTENSION_CASES)

SMALL = ["coarse_free", "coarse_free_jac", "small_jac", "small_rb", "ibc2", "ibc3", "ibc4", "odd_sweeps", "inlets", "ibeg",
         "rand64"]


def write_case(name: str, directory: str, **overrides):
    p = CASES[name]
    if overrides:
        p = p.replace(**overrides)
    return p.write(os.path.join(directory, name))

"""Problem descriptions in the reference's own file formats.

`Problem` holds everything the reference reads from ``InputData`` and
``pattern.txt`` (Version-6-26-04.c:280-401 readInput, 2367-2400
read_obstacle_data; grammar in README.md:7-78) and writes/reads those files
verbatim, so the same pair of files can be handed to the reference, to the CPU
restatement and to the HIP solver.

The pattern generators follow the algorithms of the reference's generators
(create_patterns_staggered.c:71-188, create_patterns_random.c:83-149) with a
project-owned seeded RNG (splitmix64 seeding xoshiro256**) instead of
``srand(time(NULL))``, so every pattern is reproducible from its seed.
Rectangles are emitted one per blocked base cell, i-major then j ascending,
which is the order create_patterns_staggered.c:47-53 writes; that order is
part of the input's meaning (the obstacle-edge pressure copies run in list
order, Version-6-26-04.c:1153-1177).
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import List, Optional, Sequence, Tuple

MASK64 = (1 << 64) - 1


@dataclasses.dataclass
class Problem:
    nfx: int = 128
    nfy: int = 128
    ngrid: int = 4
    ig_obs: int = 4
    ibeg_obs: int = 0
    ig_inlt: int = 0
    ig_outlt: int = 0
    nvcycle: int = 200
    nswpm: int = 2
    nswpp: int = 4
    isolve: int = 1
    omega: float = 1.0
    convergence: float = 1e-6
    xl: float = 1.0
    yl: float = 1.0
    amuf: float = 0.01
    prsc: float = 1.0
    relxm: float = 0.7
    relxp: float = 0.3
    ibc: Tuple[int, int, int, int] = (1, 1, 1, 1)
    bc_xm: Tuple[float, float, float, float] = (1.0, 0.0, 0.0, 1.0)
    bc_xp: Tuple[float, float, float, float] = (1.0, 0.0, 0.0, 0.0)
    bc_ym: Tuple[float, float, float, float] = (0.0, 0.0, 0.0, 0.0)
    bc_yp: Tuple[float, float, float, float] = (0.0, 0.0, 0.0, 0.0)
    guess: Tuple[float, float, float, float] = (1.0, 0.0, 0.0, 0.0)
    sc_obs: float = 0.0
    obstacles: List[Tuple[int, int, int, int]] = dataclasses.field(default_factory=list)
    inlets: List[Tuple[int, int, float, float, float]] = dataclasses.field(default_factory=list)
    outlets: List[Tuple[int, int, float, float, float]] = dataclasses.field(default_factory=list)

    # ------------------------------------------------------------------
    def inputdata_text(self) -> str:
        r = repr
        lines = [
            f"{self.nfx} {self.nfy} {self.ngrid} {self.ig_obs} {self.ibeg_obs} "
            f"{self.ig_inlt} {self.ig_outlt} {self.nvcycle} {self.nswpm} "
            f"{self.nswpp} {self.isolve} {r(float(self.omega))} {r(float(self.convergence))}",
            " ".join(r(float(x)) for x in (self.xl, self.yl, self.amuf, self.prsc,
                                            self.relxm, self.relxp)),
            " ".join(str(int(x)) for x in self.ibc),
        ]
        for side in (self.bc_xm, self.bc_xp, self.bc_ym, self.bc_yp, self.guess):
            lines.append(" ".join(r(float(x)) for x in side))
        lines.append(r(float(self.sc_obs)))
        lines.append(str(len(self.inlets)))
        for jf, jl, u, v, s in self.inlets:
            lines.append(f"{jf} {jl} {r(float(u))} {r(float(v))} {r(float(s))}")
        lines.append(str(len(self.outlets)))
        for jf, jl, u, v, s in self.outlets:
            lines.append(f"{jf} {jl} {r(float(u))} {r(float(v))} {r(float(s))}")
        return "\n".join(lines) + "\n"

    def pattern_text(self) -> str:
        out = [str(len(self.obstacles))]
        out += [f"{a} {b} {c} {d}" for a, b, c, d in self.obstacles]
        return "\n".join(out) + "\n"

    def write(self, directory: str) -> Tuple[str, str]:
        """Write ``InputData`` and ``pattern.txt`` into *directory*."""
        os.makedirs(directory, exist_ok=True)
        p_in = os.path.join(directory, "InputData")
        p_pat = os.path.join(directory, "pattern.txt")
        with open(p_in, "w") as f:
            f.write(self.inputdata_text())
        with open(p_pat, "w") as f:
            f.write(self.pattern_text())
        return p_in, p_pat

    def level_dims(self, ig: int) -> Tuple[int, int]:
        fac = 1 << ig
        return self.nfx // fac + 2, self.nfy // fac + 2

    def replace(self, **kw) -> "Problem":
        return dataclasses.replace(self, **kw)


# ----------------------------------------------------------------------
# seeded RNG: splitmix64 -> xoshiro256**
# ----------------------------------------------------------------------
class Rng:
    def __init__(self, seed: int):
        x = seed & MASK64
        s = []
        for _ in range(4):
            x = (x + 0x9E3779B97F4A7C15) & MASK64
            z = x
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
            s.append(z ^ (z >> 31))
        self.s = s

    @staticmethod
    def _rotl(x: int, k: int) -> int:
        return ((x << k) | (x >> (64 - k))) & MASK64

    def next_u64(self) -> int:
        s = self.s
        result = (self._rotl((s[1] * 5) & MASK64, 7) * 9) & MASK64
        t = (s[1] << 17) & MASK64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = self._rotl(s[3], 45)
        return result

    def uniform(self) -> float:
        """Uniform in [0, 1) with 53 random bits."""
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def below(self, n: int) -> int:
        return self.next_u64() % n

    def normal(self) -> float:
        """Standard normal (Box-Muller, first branch only)."""
        u1 = self.uniform()
        while u1 <= 0.0:
            u1 = self.uniform()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


# ----------------------------------------------------------------------
# pattern generators (rectangles are 1-based, i-range then j-range, on the
# base level ig_obs-1)
# ----------------------------------------------------------------------
def cells_to_rects(blocked: Sequence[Sequence[bool]]) -> List[Tuple[int, int, int, int]]:
    """One ``i i j j`` rectangle per blocked cell, i-major then j ascending.
    ``blocked[i][j]`` with 0-based i (x) and j (y)."""
    rects = []
    for i, col in enumerate(blocked):
        for j, b in enumerate(col):
            if b:
                rects.append((i + 1, i + 1, j + 1, j + 1))
    return rects


def inline_blocks(n_base: int = 16, positions: Sequence[int] = (3, 7, 11, 15)):
    """Inline array of single-cell blocks at ``positions`` x ``positions``
    (1-based), i outer, j inner."""
    return [(i, i, j, j) for i in positions for j in positions if i <= n_base and j <= n_base]


def random_mask(n_base: int, porosity: float, seed: int, free_edge_columns: bool = True):
    """Bernoulli mask: each base cell blocked with probability 1-porosity.
    The first and last base columns stay free so the plug inlet/outlet fluxes
    balance (a blocked boundary column zeroes an inlet/outlet face)."""
    rng = Rng(seed)
    blocked = [[False] * n_base for _ in range(n_base)]
    for i in range(n_base):
        for j in range(n_base):
            b = rng.uniform() < (1.0 - porosity)
            if free_edge_columns and (i == 0 or i == n_base - 1):
                b = False
            blocked[i][j] = b
    return cells_to_rects(blocked)


def staggered_pattern(n_base: int, spacing: int, shift: str, seed: int):
    """Staggered lattice of 1x1 base-cell obstacles with ``spacing`` free
    cells between them, each obstacle moved by one sub-cell unit in a random
    one of {stay, -i, +i, -j, +j}; the sub-cell unit is a quarter, half or
    full base cell (shift in {"quarter", "half", "full"}).  Follows
    create_patterns_staggered.c:71-188.  Returns (rects, sub_factor) where the
    rectangles live on an (n_base*sub_factor)^2 grid."""
    scale = {"quarter": 4, "half": 2, "full": 1}[shift]
    n = n_base * scale
    obj = scale
    sp = spacing * scale
    pad = 5 * scale
    step = obj + sp
    npad = n + 2 * pad
    center = obj // 2
    obs = []
    r = -center
    while r <= npad:
        c = -center
        while c <= npad:
            obs.append([r - pad, r + obj - 1 - pad, c - pad, c + obj - 1 - pad])
            c += step
        r += step
    voff = step // 2
    for o in obs:
        col_idx = (o[2] + pad + center) // step
        if col_idx % 2 == 1:
            o[0] += voff
            o[1] += voff
    rng = Rng(seed)
    for o in obs:
        mv = rng.below(5)
        dr = -1 if mv == 1 else (1 if mv == 2 else 0)
        dc = -1 if mv == 3 else (1 if mv == 4 else 0)
        o[0] += dr
        o[1] += dr
        o[2] += dc
        o[3] += dc
    grid = [[False] * n for _ in range(n)]
    for r1, r2, c1, c2 in obs:
        for rr in range(r1, r2 + 1):
            for cc in range(c1, c2 + 1):
                if 0 <= rr < n and 0 <= cc < n:
                    grid[rr][cc] = True
    # the generator's first index is the solver's i (x); keep the outermost
    # columns free so the plug outlet stays mass-consistent
    for j in range(n):
        grid[0][j] = False
        grid[n - 1][j] = False
    return cells_to_rects(grid), scale


# ----------------------------------------------------------------------
# named workloads
# ----------------------------------------------------------------------
def channel_128_inline() -> Problem:
    """128x128 channel, 4 levels (coarsest 16x16), inline 4x4 array of
    single-cell blocks on the 16x16 base grid, Re=100, Jacobi 2/4,
    200 V-cycles, tolerance 1e-6."""
    return Problem(obstacles=inline_blocks())


def random_1024(ngrid: int = 7, seed: int = 1, porosity: float = 0.8) -> Problem:
    """1024x1024 channel, Bernoulli(porosity) mask on a 64x64 base grid (each
    blocked base cell is 16x16 fine cells), Re=200, red-black 2/6."""
    rects = random_mask(64, porosity, seed)
    return Problem(nfx=1024, nfy=1024, ngrid=ngrid, ig_obs=5, nvcycle=2999,
                   nswpm=2, nswpp=6, isolve=2, amuf=0.005, obstacles=rects)


def sensitivity_256(seed: int = 1):
    """Differentiation workload: 256x256 channel, 5 levels, Bernoulli(0.85)
    mask on a 32x32 base grid (8x8 fine cells per base cell), Re 100,
    red-black 2/6, converged to sum|div| < 1e-12; inlet u profile parabolic
    (mean 1) plus seeded N(0, 0.05) noise, east boundary u at its mean.
    Returns (problem, inlet profile)."""
    nfy = 256
    y = [(j + 0.5) / nfy for j in range(nfy)]
    rng = Rng(seed + 7919)
    inlet = [6.0 * t * (1.0 - t) + 0.05 * rng.normal() for t in y]
    mean = sum(inlet) / nfy
    p = Problem(nfx=256, nfy=nfy, ngrid=5, ig_obs=4, nvcycle=3000, convergence=1e-12,
                nswpm=2, nswpp=6, isolve=2, amuf=0.01, bc_xp=(mean, 0.0, 0.0, 0.0),
                obstacles=random_mask(32, 0.85, seed))
    return p, inlet


def parse(inputdata: str, pattern: str) -> Problem:
    """Read an InputData / pattern.txt pair (whitespace separated, in the
    reference's reading order)."""
    with open(inputdata) as f:
        tok = f.read().split()
    with open(pattern) as f:
        ptok = f.read().split()
    it = iter(tok)

    def ni():
        return int(next(it))

    def nf():
        return float(next(it))

    p = Problem()
    (p.nfx, p.nfy, p.ngrid, p.ig_obs, p.ibeg_obs, p.ig_inlt, p.ig_outlt,
     p.nvcycle, p.nswpm, p.nswpp, p.isolve) = [ni() for _ in range(11)]
    p.omega, p.convergence = nf(), nf()
    p.xl, p.yl, p.amuf, p.prsc, p.relxm, p.relxp = [nf() for _ in range(6)]
    p.ibc = tuple(ni() for _ in range(4))
    p.bc_xm, p.bc_xp, p.bc_ym, p.bc_yp, p.guess = [tuple(nf() for _ in range(4)) for _ in range(5)]
    pit = iter(ptok)
    nobs = int(next(pit))
    p.obstacles = [tuple(int(next(pit)) for _ in range(4)) for _ in range(max(nobs, 0))]
    p.sc_obs = nf()
    n = ni()
    p.inlets = [(ni(), ni(), nf(), nf(), nf()) for _ in range(max(n, 0))]
    n = ni()
    p.outlets = [(ni(), ni(), nf(), nf(), nf()) for _ in range(max(n, 0))]
    return p


def staggered_batch(count: int = 256, seed0: int = 1000):
    """Many channel problems for one batched solve: grid n drawn uniformly
    from {128, 256, 512} with the level count that makes the coarsest level
    16x16, a staggered obstacle array (spacing 4 base cells on a 16x16 base)
    with every obstacle moved by a quarter, half or full base cell, and
    Re drawn from {10, 50, 100, 200, 400}; red-black 2/6, tolerance 1e-6.
    Instance k uses seed seed0 + k for every draw."""
    out = []
    for k in range(count):
        rng = Rng(seed0 + k)
        n = (128, 256, 512)[rng.below(3)]
        shift = ("quarter", "half", "full")[rng.below(3)]
        re = (10, 50, 100, 200, 400)[rng.below(5)]
        rects, scale = staggered_pattern(16, 4, shift, seed0 + k)
        ngrid = int(round(math.log2(n))) - 3
        ig_obs = ngrid - int(round(math.log2(scale)))
        out.append(Problem(nfx=n, nfy=n, ngrid=ngrid, ig_obs=ig_obs, nvcycle=2999, nswpm=2,
                           nswpp=6, isolve=2, amuf=1.0 / re, obstacles=rects))
    return out


def adjoint_batch(count: int = 64, seed0: int = 4000, convergence: float = 1e-8,
                  n_base: int = 16):
    """The batched-adjoint workload: 512x512 channels, 6 levels (coarsest
    16x16), each with a Bernoulli mask on an n_base x n_base base grid
    (obstacles on the coarsest level, ig_obs = 6 for the 16x16 base) whose
    porosity is drawn uniformly from [0.7, 0.95], Re drawn uniformly from
    [20, 300], red-black 2/6, plug inlet and outlet (u = 1), the primal
    converged to the FP64 sum |div| < `convergence` (the adjoint is taken at
    the converged state).  Instance k uses seed seed0 + k for every draw."""
    out = []
    for k in range(count):
        rng = Rng(seed0 + k)
        por = 0.7 + 0.25 * rng.uniform()
        re = 20.0 + 280.0 * rng.uniform()
        ig_obs = 6 - int(round(math.log2(n_base // 16))) if n_base >= 16 else 6
        out.append(Problem(nfx=512, nfy=512, ngrid=6, ig_obs=ig_obs, nvcycle=2999,
                           convergence=convergence, nswpm=2, nswpp=6, isolve=2, amuf=1.0 / re,
                           obstacles=random_mask(n_base, por, seed0 + k)))
    return out

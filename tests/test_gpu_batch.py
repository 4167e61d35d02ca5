"""Batched solves (one launch per same-depth level of many problems) against
one CPU-restatement solve per problem: same cycle count and converged flag,
residual history within 1e-12 relative, every state array bit-identical.
"""
import sys

import numpy as np
import pytest

import orc
from mgfvm.inputs import Problem, random_mask, staggered_pattern

pytestmark = pytest.mark.gpu

STATE = ["u", "v", "p", "pp", "sc", "resu", "resv", "ressc", "apu", "apv"]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _mixed_problems():
    """Different grid sizes and depths (same coarsest shape), different
    Reynolds numbers, obstacle sets and boundary types; one with a different
    coarsest shape (its own groups)."""
    probs = []
    for k, (n, ng) in enumerate([(32, 3), (64, 4), (128, 5), (64, 4), (32, 3)]):
        rects = random_mask(8, 0.8, 100 + k)
        probs.append(Problem(nfx=n, nfy=n, ngrid=ng, ig_obs=ng, isolve=2, nswpm=2, nswpp=3,
                             nvcycle=25, amuf=[0.1, 0.02, 0.01, 0.005, 0.0025][k],
                             obstacles=rects))
    probs[3] = probs[3].replace(ibc=(1, 4, 1, 1))
    probs.append(Problem(nfx=48, nfy=24, ngrid=3, ig_obs=3, isolve=2, nswpm=2, nswpp=3,
                         nvcycle=20, xl=2.0, yl=0.5, obstacles=[(2, 2, 1, 1), (3, 3, 2, 3)]))
    rects, _ = staggered_pattern(16, 4, "half", 77)
    probs.append(Problem(nfx=64, nfy=64, ngrid=4, ig_obs=2, isolve=2, nswpm=2, nswpp=3,
                         nvcycle=30, amuf=0.01, obstacles=rects))
    return probs


def test_batch_matches_single_solves(tmp_path):
    import mgfvm
    probs = _mixed_problems()
    pairs = [p.write(str(tmp_path / f"p{k}")) for k, p in enumerate(probs)]
    _say(f"[gpu] batch of {len(probs)}")
    bt = mgfvm.Batch(pairs)
    ncyc, conv = bt.solve()
    for k, (p, (pin, ppat)) in enumerate(zip(probs, pairs)):
        o = orc.Oracle(pin, ppat)
        ho, no, co = o.solve(p.nvcycle, p.convergence)
        assert ncyc[k] == no and conv[k] == co, k
        np.testing.assert_allclose(bt.history(k), ho, rtol=1e-12, atol=0.0)
        for ig in range(o.ngrid):
            for n in STATE:
                g, r = bt.get(k, ig, n), o.get(ig, n)
                assert np.array_equal(g.view(np.uint64), r.view(np.uint64)), (k, ig, n)
        o.close()


def test_batch_converged_members_stop(tmp_path):
    """Members that converge early stop at their own cycle while the rest
    keep going; a member's own cycle limit is honoured exactly."""
    import mgfvm
    from mgfvm.inputs import channel_128_inline
    a = channel_128_inline().replace(isolve=2, nswpp=4)             # converges in 68
    b = a.replace(obstacles=[], nvcycle=50)                          # stops at its limit
    c = a.replace(amuf=0.02)
    probs = [a, b, c]
    pairs = [p.write(str(tmp_path / f"q{k}")) for k, p in enumerate(probs)]
    _say("[gpu] batch early stop")
    bt = mgfvm.Batch(pairs)
    ncyc, conv = bt.solve()
    for k, (p, (pin, ppat)) in enumerate(zip(probs, pairs)):
        o = orc.Oracle(pin, ppat)
        ho, no, co = o.solve(p.nvcycle, p.convergence)
        assert (ncyc[k], conv[k]) == (no, co), k
        for n in ("u", "v", "p", "sc"):
            assert np.array_equal(bt.get(k, 0, n).view(np.uint64), o.get(0, n).view(np.uint64))
        o.close()
    assert ncyc[1] == 50


def test_batch_rejects_mixed_smoothers(tmp_path):
    import mgfvm
    p = Problem(nfx=32, nfy=32, ngrid=3, ig_obs=3, isolve=2)
    q = p.replace(isolve=1)
    pairs = [p.write(str(tmp_path / "a")), q.write(str(tmp_path / "b"))]
    with pytest.raises(mgfvm.MgfError):
        mgfvm.Batch(pairs)


# ---------------------------------------------------------------------------
# mixed precision (FP32 correction sweeps, FP64 residuals and state)
# ---------------------------------------------------------------------------
MIXED_FIELD_RTOL = 1e-6     # max |mixed - fp64| / max |fp64| per field, measured <= 1.1e-7
# The seeded random batches stop after 8 V-cycles from the initial guess, far
# from convergence, and some members' scalar is then in an unstable transient
# (|sc| up to 3e3 against an inlet value of 1): such states amplify any change
# in the rounding of the FP32 corrections.  Measured worst over the nine
# batches: sc 2.2e-6 (6.6e-7 with FP64-rounded reciprocals in the FP32
# equations), u and v 3e-7.  A mishandled face or cell moves a field by 1e-4
# or more.
MIXED_TRANSIENT_RTOL = 1e-5


@pytest.mark.parametrize("name", ["cfg0_rb", "rand64", "stag128", "coarse_free", "small_jac"])
def test_mixed_precision_tracks_fp64(tmp_path, name):
    """Mixed precision keeps the FP64 stopping test: same cycle count and
    converged flag as the FP64 oracle, history and every level-0 field within
    MIXED_FIELD_RTOL (relative to the field's magnitude)."""
    import cases
    import mgfvm
    pin, ppat = cases.write_case(name, str(tmp_path))
    p = cases.CASES[name]
    _say(f"[gpu] mixed {name}")
    m = mgfvm.Solver(pin, ppat, precision=mgfvm.PREC_MIXED)
    hm, nm, cm = m.solve()
    o = orc.Oracle(pin, ppat)
    ho, no, co = o.solve(p.nvcycle, p.convergence)
    assert (nm, cm) == (no, co)
    np.testing.assert_allclose(hm, ho, rtol=1e-5, atol=1e-12)
    for n in ("u", "v", "p", "pp", "sc"):
        g, r = m.get(0, n), o.get(0, n)
        scale = max(np.abs(r).max(), 1e-300)
        assert np.abs(g - r).max() / scale < MIXED_FIELD_RTOL, n
    o.close()


def test_mixed_batch_matches_mixed_singles(tmp_path):
    """Batching does not change mixed-precision results: every member is
    bit-identical to its own single mixed solve."""
    import mgfvm
    probs = _mixed_problems()[:4]
    pairs = [p.write(str(tmp_path / f"m{k}")) for k, p in enumerate(probs)]
    _say("[gpu] mixed batch vs singles")
    bt = mgfvm.Batch(pairs, precision=mgfvm.PREC_MIXED)
    ncyc, conv = bt.solve()
    for k, (pin, ppat) in enumerate(pairs):
        s = mgfvm.Solver(pin, ppat, precision=mgfvm.PREC_MIXED)
        h, n, c = s.solve()
        assert (n, c) == (ncyc[k], conv[k])
        assert np.array_equal(h, bt.history(k))
        for name in ("u", "v", "p", "sc"):
            assert np.array_equal(s.get(0, name), bt.get(k, 0, name)), (k, name)
        s.close()


@pytest.mark.parametrize("spl,threads,split", [(2, 256, False), (3, 256, False), (3, 512, False),
                                               (6, 1024, False), (1, 256, False), (3, 512, True)])
def test_sweep_launch_shapes_bit_identical(tmp_path, spl, threads, split):
    """The launch shape of the FP32 red-black smoothing is result-neutral:
    several sweeps per launch (temporal blocking, a halo of two points per
    sweep), the tile height, and assembling the equation inside the first
    sweep launch (halo recomputed per workgroup, sum |div| partials in the
    assembly kernel's order) all give every member's history and fields bit
    for bit as one sweep per launch after a separate assembly kernel."""
    import mgfvm
    probs = [p.replace(nswpp=6) for p in _mixed_problems()]
    pairs = [p.write(str(tmp_path / f"b{k}")) for k, p in enumerate(probs)]
    _say(f"[gpu] sweep launch shape spl={spl} threads={threads} split={split}")
    base = mgfvm.Batch(pairs, precision=mgfvm.PREC_MIXED, sweeps_per_launch=1,
                       split_assembly=True)
    nb, cb = base.solve()
    bt = mgfvm.Batch(pairs, precision=mgfvm.PREC_MIXED, sweeps_per_launch=spl,
                     sweep_threads=threads, split_assembly=split)
    ncyc, conv = bt.solve()
    assert (ncyc, conv) == (nb, cb)
    for k in range(len(probs)):
        assert np.array_equal(base.history(k), bt.history(k)), k
        for name in ("u", "v", "p", "pp", "sc", "apu", "apv"):
            assert np.array_equal(base.get(k, 0, name).view(np.uint64),
                                  bt.get(k, 0, name).view(np.uint64)), (k, name)
    base.close()
    bt.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_compaction_result_neutral(tmp_path, precision):
    """Re-capturing the V-cycle over the still-running members (members stop
    at cycles 2 .. 90, so the schedule is rebuilt several times, down to one
    member) gives every member the history and fields of a batch that keeps
    launching over all of them; a second solve after a reset rebuilds the
    schedules again and repeats the first bit for bit."""
    import mgfvm
    prec = mgfvm.PREC_FP64 if precision == "fp64" else mgfvm.PREC_MIXED
    base = _mixed_problems()
    limits = [2, 5, 9, 17, 33, 60, 90]
    probs = [p.replace(nvcycle=lim, convergence=1e-14) for p, lim in zip(base, limits)]
    pairs = [p.write(str(tmp_path / f"c{k}")) for k, p in enumerate(probs)]
    _say(f"[gpu] compaction {precision}")
    ref = mgfvm.Batch(pairs, precision=prec, compact=False)
    nr, cr = ref.solve()
    assert list(nr) == limits
    bt = mgfvm.Batch(pairs, precision=prec)
    for rep in range(2):
        if rep:
            bt.reset()
        nc, cv = bt.solve()
        assert (list(nc), list(cv)) == (list(nr), list(cr)), rep
        for k in range(len(probs)):
            assert np.array_equal(ref.history(k), bt.history(k)), (rep, k)
            for ig in range(probs[k].ngrid):
                for name in STATE if prec == mgfvm.PREC_FP64 else ("u", "v", "p", "sc"):
                    assert np.array_equal(ref.get(k, ig, name).view(np.uint64),
                                          bt.get(k, ig, name).view(np.uint64)), (rep, k, ig, name)
    ref.close()
    bt.close()


@pytest.mark.parametrize("precision,lanes,graph", [("mixed", 2, True), ("mixed", 3, True),
                                                   ("mixed", 4, False), ("fp64", 2, True)])
def test_lanes_result_neutral(tmp_path, precision, lanes, graph):
    """Splitting the running problems into lanes (one V-cycle graph per lane,
    each on its own stream, rebuilt as problems stop) gives every problem
    the history and fields of a single-lane batch, bit for bit."""
    import mgfvm
    prec = mgfvm.PREC_FP64 if precision == "fp64" else mgfvm.PREC_MIXED
    base = _mixed_problems()
    limits = [3, 7, 12, 20, 33, 50, 70]
    probs = [p.replace(nvcycle=lim, convergence=1e-14) for p, lim in zip(base, limits)]
    pairs = [p.write(str(tmp_path / f"l{k}")) for k, p in enumerate(probs)]
    _say(f"[gpu] lanes {lanes} {precision} graph={graph}")
    ref = mgfvm.Batch(pairs, precision=prec, lanes=1)
    nr, cr = ref.solve()
    bt = mgfvm.Batch(pairs, precision=prec, lanes=lanes, use_graph=graph)
    for rep in range(2):
        if rep:
            bt.reset()
        nc, cv = bt.solve()
        assert (list(nc), list(cv)) == (list(nr), list(cr)), rep
        for k in range(len(probs)):
            assert np.array_equal(ref.history(k), bt.history(k)), (rep, k)
            for name in ("u", "v", "p", "pp", "sc"):
                assert np.array_equal(ref.get(k, 0, name).view(np.uint64),
                                      bt.get(k, 0, name).view(np.uint64)), (rep, k, name)
    ref.close()
    bt.close()


@pytest.mark.parametrize("names", [("ibc2", "ibc3", "ibc4", "inlets", "rand64", "ibeg")])
def test_merged_relaxation_boundary_variants(tmp_path, names):
    """Boundary variants in one mixed batch, red-black: p' Dirichlet rows
    (ibc 2), periodic copies (ibc 3: merged relaxation, its update without
    the fused restriction), outflow (ibc 4: field by field), discrete inlets
    and outlets, a random mask on another grid size, an obstacle offset.  Each member bit for
    bit as the field-by-field split-assembly batch, same cycle counts and
    flags as the FP64 oracle, fields within the mixed tolerance."""
    import cases
    import mgfvm
    probs, pairs = [], []
    for k, name in enumerate(names):
        p = cases.CASES[name].replace(isolve=2, nswpm=2, nswpp=4, nvcycle=30)
        probs.append(p)
        pairs.append(p.write(str(tmp_path / f"v{k}")))
    _say(f"[gpu] merged relaxation boundary variants {names}")
    split = mgfvm.Batch(pairs, precision=mgfvm.PREC_MIXED, split_assembly=True)
    ns, cs = split.solve()
    bt = mgfvm.Batch(pairs, precision=mgfvm.PREC_MIXED)
    nb, cb = bt.solve()
    assert (list(nb), list(cb)) == (list(ns), list(cs))
    for k, (p, (pin, ppat)) in enumerate(zip(probs, pairs)):
        assert np.array_equal(split.history(k), bt.history(k)), k
        for name in ("u", "v", "p", "pp", "sc", "apu", "apv"):
            assert np.array_equal(split.get(k, 0, name).view(np.uint64),
                                  bt.get(k, 0, name).view(np.uint64)), (k, name)
        o = orc.Oracle(pin, ppat)
        ho, no, co = o.solve(p.nvcycle, p.convergence)
        assert (nb[k], cb[k]) == (no, co), k
        np.testing.assert_allclose(bt.history(k), ho, rtol=1e-5, atol=1e-12)
        for name in ("u", "v", "p", "sc"):
            g, r = bt.get(k, 0, name), o.get(0, name)
            assert np.abs(g - r).max() / max(np.abs(r).max(), 1e-300) < MIXED_FIELD_RTOL, (k, name)
        o.close()
    split.close()
    bt.close()


def _random_batch(seed):
    """Six seeded problems sharing one smoother and sweep counts (a batch's
    requirement), each with its own size, depth, obstacle level (coarsened
    obstacles below it when ig_obs < ngrid), Bernoulli mask, Reynolds number
    and outflow boundary choice."""
    rng = np.random.default_rng(seed)
    isolve = int(rng.integers(1, 3))
    nswpm, nswpp = int(rng.integers(1, 4)), int(rng.integers(1, 7))
    probs = []
    for k in range(6):
        n = int(rng.choice([32, 64]))
        ng = int(rng.integers(2, 5))
        ig_obs = int(rng.integers(1, ng + 1))
        base = n >> (ig_obs - 1)
        mask = random_mask(base, float(rng.uniform(0.75, 0.9)), int(rng.integers(1, 10**6)))
        ibc = (1, 4, 1, 1) if rng.uniform() < 0.3 else (1, 1, 1, 1)
        probs.append(Problem(nfx=n, nfy=n, ngrid=ng, ig_obs=ig_obs, isolve=isolve, nswpm=nswpm,
                             nswpp=nswpp, nvcycle=10, convergence=1e-30,
                             amuf=float(rng.choice([0.1, 0.05, 0.02])), ibc=ibc,
                             obstacles=mask))
    return probs


def _random_batch_io(seed):
    """Discrete inlets and outlets given on a random level (ig_inlt,
    ig_outlt), and an obstacle offset (ibeg_obs)."""
    rng = np.random.default_rng(seed)
    isolve = int(rng.integers(1, 3))
    nswpm, nswpp = int(rng.integers(1, 4)), int(rng.integers(1, 7))
    probs = []
    for k in range(6):
        n = int(rng.choice([32, 64]))
        ng = int(rng.integers(2, 5))
        ig_obs = int(rng.integers(1, ng + 1))
        ibeg = int(rng.integers(0, 3))
        b = n >> (ig_obs - 1)
        blocked = rng.uniform(size=(b, b)) < 0.15
        rects = [(i + 1, i + 1, j + 1, j + 1) for i in range(1, b - 1 - ibeg) for j in range(b)
                 if blocked[i, j]]

        def ports(ig):
            rows = n >> (ig - 1)
            out, j = [], 1
            while j <= rows and len(out) < 3:
                jf = j + int(rng.integers(0, max(1, rows // 4)))
                jl = min(rows, jf + int(rng.integers(0, max(1, rows // 4))))
                if jf > rows:
                    break
                out.append((jf, jl, float(rng.uniform(0.3, 1.5)), float(rng.uniform(-0.1, 0.1)),
                            float(rng.uniform(0.0, 2.0))))
                j = jl + 2
            return out

        ig_in, ig_out = int(rng.integers(1, ng + 1)), int(rng.integers(1, ng + 1))
        probs.append(Problem(nfx=n, nfy=n, ngrid=ng, ig_obs=ig_obs, ibeg_obs=ibeg,
                             ig_inlt=ig_in, ig_outlt=ig_out, isolve=isolve, nswpm=nswpm,
                             nswpp=nswpp, nvcycle=8, convergence=1e-30,
                             amuf=float(rng.choice([0.1, 0.05])), inlets=ports(ig_in),
                             outlets=ports(ig_out), obstacles=rects))
    return probs


def _family(name, seed):
    return {"mask": _random_batch, "bc": _random_batch_bc, "io": _random_batch_io}[name](seed)


@pytest.mark.parametrize("family,seed", [("mask", 9), ("mask", 38), ("mask", 4), ("mask", 29),
                                         ("mask", 50), ("bc", 6), ("bc", 35), ("bc", 31),
                                         ("bc", 33), ("bc", 70), ("io", 9), ("io", 23),
                                         ("io", 10), ("io", 15), ("io", 29)])
def test_random_batches_bitwise(tmp_path, family, seed):
    """Seeded random batches in FP64: every member's cycle count, history and
    every state array of every level as the oracle's.  "mask" seeds 9 and 38
    draw Jacobi, 4, 29 and 50 red-black; all mix depths 2-4, coarsened
    obstacles and outflow members.  "bc" seeds 6 and 35 draw Jacobi, 31, 33
    and 70 red-black, each with at least three of the four boundary sets on
    rectangular grids.  "io" seeds 9 and 23 draw Jacobi, 10, 15 and 29
    red-black, with discrete inlets and outlets on random levels and obstacle
    offsets.  Every state array of every member stays finite (a
    NaN's bits differ between the host's and the GPU's arithmetic: seed 26,
    whose scalar diverges on a periodic member, is left out for that)."""
    import mgfvm
    probs = _family(family, seed)
    p0 = probs[0]
    _say(f"[gpu] random batch {family} seed {seed}: isolve {p0.isolve}, "
         f"sweeps {p0.nswpm}/{p0.nswpp}, "
         + ", ".join(f"{p.nfx}x{p.nfy} ng{p.ngrid} obs{p.ig_obs} ibc{p.ibc}" for p in probs))
    pairs = [p.write(str(tmp_path / f"r{k}")) for k, p in enumerate(probs)]
    bt = mgfvm.Batch(pairs)
    ncyc, conv = bt.solve()
    for k, (p, (pin, ppat)) in enumerate(zip(probs, pairs)):
        o = orc.Oracle(pin, ppat)
        ho, no, co = o.solve(p.nvcycle, p.convergence)
        assert np.all(np.isfinite(ho)), (k, "pick parameters that stay finite")
        assert ncyc[k] == no and conv[k] == co, k
        np.testing.assert_allclose(bt.history(k), ho, rtol=1e-12, atol=0.0)
        for ig in range(o.ngrid):
            for n in STATE:
                g, r = bt.get(k, ig, n), o.get(ig, n)
                assert np.all(np.isfinite(r)), (k, ig, n, "pick parameters that stay finite")
                assert np.array_equal(g.view(np.uint64), r.view(np.uint64)), (k, ig, n)
        o.close()
    bt.close()


def _connected_cells(o):
    """Level-0 cells reachable from a fluid cell of the inlet column through
    open (unpenalised) faces.  The others are fluid pockets cut off by
    obstacles, and obstacle cells: their p' rows are ~1e-33 times the
    identity, so their p is not determined by the equations (it differs
    between any two arithmetics, FP64 runs included)."""
    tag, spu, spv = o.get(0, "tag"), o.get(0, "spu"), o.get(0, "spv")
    ny, nx = tag.shape
    seen = np.zeros((ny, nx), bool)
    stack = [(j, 1) for j in range(1, ny - 1) if tag[j, 1] != 0]
    for j, i in stack:
        seen[j, i] = True
    while stack:
        j, i = stack.pop()
        # neighbours through the u faces (i-1 | i, i | i+1) and v faces
        for jj, ii, open_ in ((j, i - 1, spu[j, i - 1] < 1e10), (j, i + 1, spu[j, i] < 1e10),
                              (j - 1, i, spv[j - 1, i] < 1e10), (j + 1, i, spv[j, i] < 1e10)):
            if (1 <= ii <= nx - 2 and 1 <= jj <= ny - 2 and open_ and not seen[jj, ii]
                    and tag[jj, ii] != 0):
                seen[jj, ii] = True
                stack.append((jj, ii))
    return seen


@pytest.mark.parametrize("family,seed", [("mask", 4), ("mask", 29), ("mask", 50), ("bc", 31),
                                         ("bc", 33), ("bc", 70), ("io", 10), ("io", 15),
                                         ("io", 29)])
def test_random_batches_mixed(tmp_path, family, seed):
    """The red-black random batches in mixed precision: bit for bit as the
    field-by-field split-assembly batch (members with an outflow boundary take
    the field-by-field sequence in both), and within the mixed tolerance of
    the FP64 oracle: u, v, sc everywhere, p on the cells connected to the
    flow, the residual history to 1e-6 of its largest entry.  Random masks
    make fluid cells enclosed by obstacles and start penalised faces at the
    guess, which the FP32 corrections must not leave at their rounding."""
    import mgfvm
    probs = _family(family, seed)
    _say(f"[gpu] random mixed batch {family} seed {seed}")
    pairs = [p.write(str(tmp_path / f"m{k}")) for k, p in enumerate(probs)]
    split = mgfvm.Batch(pairs, precision=mgfvm.PREC_MIXED, split_assembly=True)
    ns, cs = split.solve()
    bt = mgfvm.Batch(pairs, precision=mgfvm.PREC_MIXED)
    nb, cb = bt.solve()
    assert (list(nb), list(cb)) == (list(ns), list(cs))
    for k, (p, (pin, ppat)) in enumerate(zip(probs, pairs)):
        assert np.array_equal(split.history(k), bt.history(k)), k
        for ig in range(p.ngrid):
            for name in STATE:
                assert np.array_equal(split.get(k, ig, name).view(np.uint64),
                                      bt.get(k, ig, name).view(np.uint64)), (k, ig, name)
        o = orc.Oracle(pin, ppat)
        ho, no, co = o.solve(p.nvcycle, p.convergence)
        assert (nb[k], cb[k]) == (no, co), k
        dev = {name: np.abs(bt.get(k, 0, name) - o.get(0, name)).max()
               / max(np.abs(o.get(0, name)).max(), 1e-300) for name in ("u", "v", "sc")}
        # p where the equations determine it (cells connected to the flow)
        conn = _connected_cells(o)
        gp, rp = bt.get(k, 0, "p")[conn], o.get(0, "p")[conn]
        dev["p"] = np.abs(gp - rp).max() / max(np.abs(rp).max(), 1e-300)
        hdev = np.abs(bt.history(k) - ho).max() / np.abs(ho).max()
        _say(f"  member {k} ({p.nfx}x{p.nfy} ng{p.ngrid} obs{p.ig_obs} ibc{p.ibc}): field dev "
             + " ".join(f"{n} {v:.2e}" for n, v in dev.items()) + f", history dev {hdev:.2e}, "
             + f"{int((o.get(0, 'tag')[1:-1, 1:-1] != 0).sum() - conn.sum())} cut-off fluid cells")
        for name, v in dev.items():
            assert v < MIXED_TRANSIENT_RTOL, (k, name, v)
        assert hdev < 1e-6, (k, hdev)
        o.close()
    split.close()
    bt.close()


def _random_batch_bc(seed):
    """Like _random_batch, with rectangular grids (dx != dy for some), p'
    Dirichlet rows (ibc 2), periodic copies (ibc 3) and outflow (ibc 4)
    members, and an obstacle scalar value."""
    rng = np.random.default_rng(seed)
    isolve = int(rng.integers(1, 3))
    nswpm, nswpp = int(rng.integers(1, 4)), int(rng.integers(1, 7))
    bcs = [(1, 1, 1, 1), (1, 4, 1, 1), (2, 2, 2, 2), (3, 1, 3, 1)]
    probs = []
    for k in range(6):
        nfx = int(rng.choice([32, 64]))
        nfy = int(rng.choice([16, 32, 64]))
        ng = int(rng.integers(2, int(np.log2(min(nfx, nfy))) - 1))
        ig_obs = int(rng.integers(1, ng + 1))
        bx, by = nfx >> (ig_obs - 1), nfy >> (ig_obs - 1)
        blocked = rng.uniform(size=(bx, by)) < rng.uniform(0.1, 0.25)
        blocked[0, :] = blocked[-1, :] = False
        rects = [(i + 1, i + 1, j + 1, j + 1) for i in range(bx) for j in range(by) if blocked[i, j]]
        probs.append(Problem(nfx=nfx, nfy=nfy, ngrid=ng, ig_obs=ig_obs, isolve=isolve,
                             nswpm=nswpm, nswpp=nswpp, nvcycle=8, convergence=1e-30,
                             xl=float(rng.choice([1.0, 2.0])) * nfx / 64,
                             yl=float(rng.choice([0.5, 1.0])) * nfy / 32,
                             amuf=float(rng.choice([0.1, 0.05])), ibc=bcs[int(rng.integers(0, 4))],
                             sc_obs=float(rng.choice([0.0, 0.5])), obstacles=rects))
    return probs


def test_random_problems_mixed_cycle_counts_equal_fp64(tmp_path):
    """540 seeded random problems (the three families above, seeds 200-229,
    red-black 2/6) run to sum|div| < 1e-6 or 400 cycles as one FP64 batch
    and one mixed batch: every problem stops after the same number of cycles
    with the same converged flag (the FP64 batch is the oracle bit for bit,
    by the tests above)."""
    import mgfvm
    probs = [p.replace(isolve=2, nswpm=2, nswpp=6, nvcycle=400, convergence=1e-6)
             for fam in ("mask", "bc", "io") for seed in range(200, 230)
             for p in _family(fam, seed)]
    pairs = [p.write(str(tmp_path / f"c{k}")) for k, p in enumerate(probs)]
    _say(f"[gpu] {len(pairs)} random problems to convergence, FP64 and mixed")
    f64 = mgfvm.Batch(pairs)
    n64, c64 = f64.solve()
    f64.close()
    mx = mgfvm.Batch(pairs, precision=mgfvm.PREC_MIXED)
    nmx, cmx = mx.solve()
    mx.close()
    assert sum(c64) >= 60, "the set should hold converging problems"
    diff = [k for k in range(len(pairs)) if (n64[k], c64[k]) != (nmx[k], cmx[k])]
    assert not diff, [(k, n64[k], nmx[k]) for k in diff[:10]]

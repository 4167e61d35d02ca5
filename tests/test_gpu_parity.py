"""Parity of the HIP solver (through the C ABI) with the CPU restatement.

FP64 bar: every state array (u, v, p, pp, sc, the FAS terms resu/resv/ressc
and the momentum diagonals apu/apv) bit-identical on every level, after
single steps, after whole V-cycles and after a graph-replayed solve.  The
recorded residual sum(|div|) is summed in a different order on the GPU, so it
is compared with a relative tolerance of 1e-12 (the bound for summing
non-negative terms is n*eps ~ 2e-12 at 1024^2 points); the cycle count and
the converged flag must be equal.
"""
import sys

import numpy as np
import pytest

import cases
import orc

pytestmark = pytest.mark.gpu

STATE = ["u", "v", "p", "pp", "sc", "resu", "resv", "ressc", "apu", "apv"]
HIST_RTOL = 1e-12


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _pair(tmp_path, name, **kw):
    import mgfvm
    pin, ppat = cases.write_case(name, str(tmp_path), **kw)
    return mgfvm.Solver(pin, ppat), orc.Oracle(pin, ppat)


def _diff_report(gpu, ora, levels, names=STATE):
    bad = []
    for ig in levels:
        for n in names:
            g, o = gpu.get(ig, n), ora.get(ig, n)
            if not np.array_equal(g.view(np.uint64), o.view(np.uint64)):
                d = np.abs(g - o)
                idx = np.unravel_index(np.nanargmax(d), d.shape) if np.isfinite(d).any() else None
                bad.append(f"L{ig}.{n}: max|d|={np.nanmax(d):.3e} at {idx} "
                           f"(gpu {g[idx] if idx else 'nan'}, ref {o[idx] if idx else 'nan'}), "
                           f"{int((g.view(np.uint64) != o.view(np.uint64)).sum())} differ")
    return bad


@pytest.mark.parametrize("name", cases.SMALL + ["cfg0"])
def test_initial_state(tmp_path, name):
    _say(f"[gpu] initial state {name}")
    gpu, ora = _pair(tmp_path, name)
    bad = _diff_report(gpu, ora, range(ora.ngrid), ["u", "v", "p", "sc", "apu", "apv", "pp"])
    assert not bad, "\n".join(bad)


@pytest.mark.parametrize("name", cases.SMALL + ["cfg0_rb"])
def test_step_by_step_vcycle(tmp_path, name):
    """Walk two V-cycles step by step (relax / restrict / prolong) and compare
    after every step, so a mismatch names the step that caused it."""
    _say(f"[gpu] steps {name}")
    gpu, ora = _pair(tmp_path, name)
    ng = ora.ngrid
    for cyc in range(2):
        for ig in range(ng):
            eg, eo = gpu.relax(ig, 0), ora.relax(ig, 0)
            bad = _diff_report(gpu, ora, [ig])
            assert not bad, f"cycle {cyc} relax down L{ig}:\n" + "\n".join(bad)
            assert eg == pytest.approx(eo, rel=HIST_RTOL, abs=0.0)
            if ig != ng - 1:
                gpu.restrict(ig)
                ora.restrict(ig)
                bad = _diff_report(gpu, ora, [ig, ig + 1])
                assert not bad, f"cycle {cyc} restrict L{ig}:\n" + "\n".join(bad)
        for ig in range(ng - 1, 0, -1):
            gpu.prolong(ig)
            ora.prolong(ig)
            bad = _diff_report(gpu, ora, [ig - 1, ig])
            assert not bad, f"cycle {cyc} prolong L{ig}:\n" + "\n".join(bad)
            if ig != 1:
                gpu.relax(ig - 1, 1)
                ora.relax(ig - 1, 1)
                bad = _diff_report(gpu, ora, [ig - 1])
                assert not bad, f"cycle {cyc} relax up L{ig-1}:\n" + "\n".join(bad)


@pytest.mark.parametrize("name", ["small_rb", "coarse_free"])
def test_prolong_after_fine_level_changed(tmp_path, name):
    """Step calls in an order no V-cycle uses: the fine level relaxes again
    between its restriction and the prolongation onto it.  mgf_prolong then
    recomputes R x_f from the changed fine level, as the reference does (the
    V-cycle's prolongation reuses the copy its restriction kept, valid only
    because a V-cycle never changes the fine level in between); a V-cycle
    afterwards must still match."""
    _say(f"[gpu] prolong after a fine relax {name}")
    gpu, ora = _pair(tmp_path, name)
    for g in (gpu, ora):
        g.relax(0, 0)
        g.restrict(0)
        g.relax(1, 0)
        g.relax(0, 1)
        g.prolong(1)
    bad = _diff_report(gpu, ora, [0, 1])
    assert not bad, "\n".join(bad)
    eg, eo = gpu.vcycle(0), ora.vcycle(0)
    assert eg == pytest.approx(eo, rel=HIST_RTOL, abs=0.0)
    bad = _diff_report(gpu, ora, range(ora.ngrid))
    assert not bad, "\n".join(bad)


@pytest.mark.parametrize("name", ["small_jac", "small_rb", "inlets", "ibc3", "rand64"])
def test_vcycles(tmp_path, name):
    _say(f"[gpu] vcycles {name}")
    gpu, ora = _pair(tmp_path, name)
    for icyc in range(5):
        eg, eo = gpu.vcycle(icyc), ora.vcycle(icyc)
        assert eg == pytest.approx(eo, rel=HIST_RTOL, abs=0.0), f"cycle {icyc}"
    bad = _diff_report(gpu, ora, range(ora.ngrid))
    assert not bad, "\n".join(bad)


@pytest.mark.parametrize("name", cases.SMALL + ["cfg0", "cfg0_rb", "stag128"])
def test_solve_matches_oracle(tmp_path, name):
    """Whole solve with the captured V-cycle graph and the device-side
    stopping test: same cycle count, same converged flag, same fields."""
    _say(f"[gpu] solve {name}")
    gpu, ora = _pair(tmp_path, name)
    p = cases.CASES[name]
    hg, ng_, cg = gpu.solve(p.nvcycle, p.convergence)
    ho, no, co = ora.solve(p.nvcycle, p.convergence)
    assert ng_ == no and cg == co
    np.testing.assert_allclose(hg, ho, rtol=HIST_RTOL, atol=0.0)
    bad = _diff_report(gpu, ora, range(ora.ngrid))
    assert not bad, "\n".join(bad)


def test_solve_without_graph_equals_graph(tmp_path):
    import mgfvm
    pin, ppat = cases.write_case("cfg0_rb", str(tmp_path))
    _say("[gpu] graph vs eager cfg0_rb")
    a = mgfvm.Solver(pin, ppat, use_graph=True)
    b = mgfvm.Solver(pin, ppat, use_graph=False)
    ha, na, ca = a.solve()
    hb, nb, cb = b.solve()
    assert (na, ca) == (nb, cb)
    assert np.array_equal(ha, hb)
    for n in STATE:
        assert np.array_equal(a.get(0, n), b.get(0, n)), n


def test_solve_twice_restarts_counting(tmp_path):
    """A second mgf_solve continues from the current state with a fresh
    cycle counter; replaying the graph again must not corrupt anything."""
    import mgfvm
    pin, ppat = cases.write_case("small_rb", str(tmp_path))
    _say("[gpu] solve twice small_rb")
    gpu = mgfvm.Solver(pin, ppat)
    ora = orc.Oracle(pin, ppat)
    for _ in range(2):
        hg, ng_, _ = gpu.solve(7, 0.0)
        ho, no, _ = ora.solve(7, 0.0)
        assert ng_ == no == 7
        np.testing.assert_allclose(hg, ho, rtol=HIST_RTOL, atol=0.0)
    bad = _diff_report(gpu, ora, range(ora.ngrid))
    assert not bad, "\n".join(bad)


def test_reset_reproduces_fresh_solve(tmp_path):
    """mgf_reset restores exactly the state mgf_create built."""
    import mgfvm
    pin, ppat = cases.write_case("coarse_free", str(tmp_path))
    _say("[gpu] reset coarse_free")
    a = mgfvm.Solver(pin, ppat)
    h1, n1, c1 = a.solve(15, 0.0)
    f1 = {(ig, n): a.get(ig, n) for ig in range(a.ngrid) for n in STATE}
    a.reset()
    h2, n2, c2 = a.solve(15, 0.0)
    assert (n1, c1) == (n2, c2) and np.array_equal(h1, h2)
    for (ig, n), want in f1.items():
        assert np.array_equal(a.get(ig, n).view(np.uint64), want.view(np.uint64)), (ig, n)
    b = mgfvm.Solver(pin, ppat)
    b.solve(15, 0.0)
    for ig in range(a.ngrid):
        for n in STATE:
            assert np.array_equal(a.get(ig, n).view(np.uint64), b.get(ig, n).view(np.uint64)), (ig, n)


def test_set_inlet_profile_equals_creation_with_it(tmp_path):
    """mgf_set_inlet_profile: a solver given an inlet profile after creation
    solves to the bits of one created with it, on every level; NULL brings
    back the InputData inlet; a batch member takes its own profile."""
    import mgfvm
    from mgfvm.inputs import Problem, random_mask
    nf = 32
    j = (np.arange(nf) + 0.5) / nf
    prof = 1.0 + 0.3 * np.cos(np.pi * j) + 0.1 * np.sin(5.0 * np.pi * j)
    prob = Problem(nfx=nf, nfy=nf, ngrid=3, ig_obs=3, isolve=2, nvcycle=40, nswpp=6,
                   bc_xp=(float(prof.mean()), 0.0, 0.0, 0.0), obstacles=random_mask(8, 0.85, 3))
    pin, ppat = prob.write(str(tmp_path / "ip"))
    _say("[gpu] set_inlet_profile")

    def state(s):
        return {(ig, n): s.get(ig, n).copy() for ig in range(prob.ngrid) for n in STATE}

    def same(x, y):
        return all(np.array_equal(x[k].view(np.uint64), y[k].view(np.uint64)) for k in x)

    a = mgfvm.Solver(pin, ppat, inlet_u=prof)
    ha = a.solve()
    b = mgfvm.Solver(pin, ppat)
    b.solve()                                   # a solve with the InputData inlet first
    b.set_inlet_profile(prof)
    hb = b.solve()
    assert (ha[1], ha[2]) == (hb[1], hb[2]) and np.array_equal(ha[0], hb[0])
    assert same(state(a), state(b))
    c = mgfvm.Solver(pin, ppat)
    hc = c.solve()
    b.set_inlet_profile(None)
    hb2 = b.solve()
    assert (hc[1], hc[2]) == (hb2[1], hb2[2]) and np.array_equal(hc[0], hb2[0])
    assert same(state(c), state(b))
    bt = mgfvm.Batch([(pin, ppat), (pin, ppat)])
    bt.set_inlet_profile(1, prof)
    nb, cb = bt.solve()
    assert (nb[1], cb[1]) == (ha[1], ha[2]) and (nb[0], cb[0]) == (hc[1], hc[2])
    for ig in range(prob.ngrid):
        for n in STATE:
            assert np.array_equal(bt.get(1, ig, n).view(np.uint64), a.get(ig, n).view(np.uint64))
            assert np.array_equal(bt.get(0, ig, n).view(np.uint64), c.get(ig, n).view(np.uint64))
    # the GPU's inlet-profile path against the oracle's, bit for bit (the
    # oracle's is pinned to the reference by test_oracle_golden)
    o = orc.Oracle(pin, ppat, inlet_u=prof)
    ho, no, co = o.solve(prob.nvcycle, prob.convergence)
    assert (no, co) == (ha[1], ha[2])
    np.testing.assert_allclose(ha[0], ho, rtol=HIST_RTOL, atol=0)
    for ig in range(prob.ngrid):
        for n in STATE:
            want = o.get(ig, n).view(np.uint64)
            assert np.array_equal(a.get(ig, n).view(np.uint64), want), (ig, n)
            assert np.array_equal(bt.get(1, ig, n).view(np.uint64), want), (ig, n)
    o.close()
    for s in (a, b, c, bt):
        s.close()


def test_inlet_profile_bitwise_with_oracle_on_reference_case(tmp_path):
    """The piecewise-constant profile of the reference-pinned case
    (discrete inlets on level 0 covering every row, fixture
    e2e_profile_inlets): the GPU given it as an inlet profile, at creation
    and through mgf_set_inlet_profile, equals the oracle given the same
    profile, and the GPU given the discrete inlets, bit for bit on every
    level."""
    import mgfvm
    base = cases.CASES["profile_inlets"]
    prof = np.array(cases.profile_of(base.inlets, base.nfy))
    plain = base.replace(ig_inlt=0, inlets=[])
    pin, ppat = plain.write(str(tmp_path / "plain"))
    din, dpat = base.write(str(tmp_path / "discrete"))
    _say("[gpu] inlet profile vs oracle (reference-pinned case)")
    o = orc.Oracle(pin, ppat, inlet_u=prof)
    ho, no, co = o.solve(plain.nvcycle, plain.convergence)
    a = mgfvm.Solver(pin, ppat, inlet_u=prof)
    b = mgfvm.Solver(pin, ppat)
    b.set_inlet_profile(prof)
    d = mgfvm.Solver(din, dpat)
    for s in (a, b, d):
        h, n, c = s.solve()
        assert (n, c) == (no, co)
        np.testing.assert_allclose(h, ho, rtol=HIST_RTOL, atol=0)
        bad = _diff_report(s, o, range(o.ngrid))
        assert not bad, "\n".join(bad)
        s.close()
    o.close()

"""The lexicographic Gauss-Seidel variant (mgf_options.gs_order = 1; the
reference has only red-black): the GPU's anti-diagonal wavefront sweeps
against the oracle's sequential j-outer, i-inner loop.  Every state array
bit-identical, same cycle count and converged flag, residual history within
1e-12 relative (the GPU sums |div| as a fixed two-level tree)."""
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


def _cases():
    rects, _ = staggered_pattern(16, 4, "half", 7)
    return {
        "rand64": Problem(nfx=64, nfy=64, ngrid=3, ig_obs=3, isolve=2, nswpm=2, nswpp=6,
                          nvcycle=30, amuf=0.01, obstacles=random_mask(16, 0.85, 21)),
        "rect48x24": Problem(nfx=48, nfy=24, ngrid=3, ig_obs=3, isolve=2, nswpm=2, nswpp=3,
                             nvcycle=20, xl=2.0, yl=0.5, obstacles=[(2, 2, 1, 1), (3, 3, 2, 3)]),
        "stag128": Problem(nfx=128, nfy=128, ngrid=4, ig_obs=2, isolve=2, nswpm=2, nswpp=6,
                           nvcycle=12, amuf=0.005, obstacles=rects),
        "outflow32": Problem(nfx=32, nfy=32, ngrid=3, ig_obs=3, isolve=2, nswpm=1, nswpp=4,
                             nvcycle=25, ibc=(1, 4, 1, 1), obstacles=random_mask(8, 0.8, 3)),
    }


def _check(k, g_get, o, p, ncyc, conv, hist):
    ho, no, co = o.solve(p.nvcycle, p.convergence)
    assert (ncyc, conv) == (no, co), k
    np.testing.assert_allclose(hist, ho, rtol=1e-12, atol=0.0)
    for ig in range(o.ngrid):
        for n in STATE:
            g, r = g_get(ig, n), o.get(ig, n)
            assert np.array_equal(g.view(np.uint64), r.view(np.uint64)), (k, ig, n)


@pytest.mark.parametrize("name", ["rand64", "rect48x24", "stag128", "outflow32"])
def test_lexicographic_single_matches_oracle(tmp_path, name):
    import mgfvm
    p = _cases()[name]
    pin, ppat = p.write(str(tmp_path / name))
    _say(f"[gpu] lexicographic GS single {name}")
    s = mgfvm.Solver(pin, ppat, gs_lex=True)
    h, n, c = s.solve()
    o = orc.Oracle(pin, ppat, gs_lex=True)
    _check(name, s.get, o, p, n, c, h)
    o.close()
    s.close()


def test_lexicographic_batch_matches_oracle(tmp_path):
    import mgfvm
    probs = list(_cases().values())[:3]
    probs = [q.replace(nswpp=6) for q in probs]
    pairs = [q.write(str(tmp_path / f"b{k}")) for k, q in enumerate(probs)]
    _say("[gpu] lexicographic GS batch")
    bt = mgfvm.Batch(pairs, gs_lex=True)
    ncyc, conv = bt.solve()
    for k, (q, (pin, ppat)) in enumerate(zip(probs, pairs)):
        o = orc.Oracle(pin, ppat, gs_lex=True)
        _check(k, lambda ig, n, k=k: bt.get(k, ig, n), o, q, ncyc[k], conv[k], bt.history(k))
        o.close()
    bt.close()


def test_lexicographic_differs_from_red_black_and_rejects_mixed(tmp_path):
    import mgfvm
    p = _cases()["rand64"].replace(nvcycle=3)
    pin, ppat = p.write(str(tmp_path / "d"))
    _say("[gpu] lexicographic vs red-black")
    a = mgfvm.Solver(pin, ppat, gs_lex=True)
    b = mgfvm.Solver(pin, ppat)
    a.solve()
    b.solve()
    assert not np.array_equal(a.get(0, "u"), b.get(0, "u"))
    a.close()
    b.close()
    with pytest.raises(mgfvm.MgfError):
        mgfvm.Solver(pin, ppat, precision=mgfvm.PREC_MIXED, gs_lex=True)

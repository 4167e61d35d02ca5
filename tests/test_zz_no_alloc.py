"""mgf_solve_noalloc: the solve in the library's static device scratch gives
what mgf_create + mgf_solve + mgf_get_field give, bit for bit (FP64) or to
the last bit of the same mixed-precision arithmetic, on two calls in a row,
the second with a smaller problem (the scratch keeps the first call's
contents, so anything that must start at zero has to be cleared); a
problem too large for the scratch is rejected with MGF_ERR_LIMIT."""
import sys

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "p", "sc")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _allocating(pin, ppat, precision):
    import mgfvm
    with mgfvm.Solver(pin, ppat, precision=precision) as s:
        h, nc, cv = s.solve()
        return np.asarray(h), nc, cv, {k: s.get(0, k).ravel() for k in FIELDS}


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_matches_allocating_solve_twice(tmp_path, precision):
    from mgfvm import _lib
    # first the larger problem, then a smaller one on the same scratch
    for name, over in (("stag128", dict(nvcycle=25)), ("small_rb", dict(nvcycle=30))):
        _say(f"[gpu] noalloc precision={precision} case={name}")
        pin, ppat = cases.write_case(name, str(tmp_path / f"{name}_{precision}"), **over)
        h0, n0, c0, f0 = _allocating(pin, ppat, precision)
        h1, n1, c1, f1 = _lib.solve_noalloc(pin, ppat, precision=precision)
        assert (n1, c1) == (n0, c0)
        np.testing.assert_array_equal(h1.view(np.uint64), h0.view(np.uint64))
        for k in FIELDS:
            assert np.array_equal(f1[k].view(np.uint64), f0[k].view(np.uint64)), k


def test_noalloc_rejects_problem_beyond_scratch(tmp_path):
    from mgfvm import _lib
    from mgfvm.inputs import Problem
    _say("[gpu] noalloc size limit")
    big = Problem(nfx=512, nfy=512, ngrid=6, ig_obs=6, isolve=2, nvcycle=5, obstacles=[])
    pin, ppat = big.write(str(tmp_path / "big"))
    with pytest.raises(_lib.MgfError) as e:
        _lib.solve_noalloc(pin, ppat)
    assert e.value.code == -6 and "scratch" in str(e.value)
    # the scratch still serves a problem that fits after the rejection
    pin, ppat = cases.write_case("small_rb", str(tmp_path / "after"), nvcycle=10)
    h0, n0, c0, f0 = _allocating(pin, ppat, 0)
    h1, n1, c1, f1 = _lib.solve_noalloc(pin, ppat)
    assert n1 == n0 and np.array_equal(f1["u"].view(np.uint64), f0["u"].view(np.uint64))

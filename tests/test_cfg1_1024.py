"""The 1024^2 single-problem workload (64x64 Bernoulli(0.8) mask, seed 1,
Re 200, red-black 2/6) as specified, with 7 levels (coarsest 16x16), and
with the 5 levels the reference can run.

7 levels: the method diverges, identically in the CPU restatement and on the
GPU.  In the first V-cycle's coarse-grid correction from the 16x16 level,
the up-leg relaxation of the 32x32 level (level 5) meets an unpenalised
u face whose hybrid coefficients aw, ae, as, an are all zero: its two
neighbouring cells carry no flux and no diffusion reaches it.  The
reference's red-black sweep divides by ap without a guard
(Version-6-26-04.c:1316-1326, unlike Jacobi's +1e-30 at 1279), so that face
becomes non-finite (0/0), the prolongations carry it to level 0, and the residual
recorded in the second V-cycle is NaN.  The GPU reproduces every bit up to
there and reports the non-finite residual as MGF_ERR_NONFINITE (no fault:
non-finite values are data).

5 levels: the mixed-precision solve stays within 1e-6 of the FP64 oracle
after the first cycles and stops after the same number of V-cycles as the
FP64 GPU solve (which is bit-identical to the oracle).
"""
import ctypes
import sys

import numpy as np
import pytest

import orc
from mgfvm.inputs import random_1024

STATE = ["u", "v", "p", "pp", "sc", "resu", "resv", "ressc", "apu", "apv"]
# the unpenalised u face of level 5 (32x32 cells) with ap = 0 in the first
# cycle's up leg, 1-based (i, j) as the reference indexes it
ZERO_AP_FACE = (31, 23)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _walk_cycle0_to_level5_up(o):
    """the first V-cycle's steps (Version-6-26-04.c:197-233) up to the
    prolongation onto level 5"""
    ng = o.ngrid
    for ig in range(ng):
        o.relax(ig, 0)
        if ig < ng - 1:
            o.restrict(ig)
    o.prolong(ng - 1)


def test_cfg1_seven_levels_diverges_in_oracle(tmp_path):
    p = random_1024(ngrid=7)
    assert (p.ngrid, p.ig_obs, p.nswpm, p.nswpp, p.isolve) == (7, 5, 2, 6, 2)
    pin, ppat = p.write(str(tmp_path / "c1"))
    o = orc.Oracle(pin, ppat)
    _walk_cycle0_to_level5_up(o)
    ig = 5
    nx, ny = o.dims(ig)
    assert (nx, ny) == (34, 34)
    # u coefficients of level 5 as the up-leg relaxation assembles them
    o.call("fluxes", ig)
    o.call("coefu", ig)
    ap, spu = o.get(ig, "ap"), o.get(ig, "spu")
    inner = (slice(1, ny - 1), slice(1, nx - 2))
    zero = np.argwhere((ap[inner] == 0.0) & (spu[inner] < 1e10))
    assert [(int(i) + 1, int(j) + 1) for j, i in zero] == [ZERO_AP_FACE]
    i, j = ZERO_AP_FACE
    for n in ("aw", "ae", "as", "an"):
        assert o.get(ig, n)[j, i] == 0.0, n
    for k in range(ig):
        assert all(np.isfinite(o.get(k, n)).all() for n in ("u", "v", "p", "sc"))
    o.relax(ig, 1)
    assert not np.isfinite(o.get(ig, "u")[j, i])
    # the rest of the first cycle; its recorded residual was finite, the
    # second cycle's is not
    for igc in range(ig, 0, -1):
        o.prolong(igc)
        if igc - 1 != 0:
            o.relax(igc - 1, 1)
    assert not np.isfinite(o.get(0, "u")).all()
    assert not np.isfinite(o.vcycle(1))
    o.close()

    o = orc.Oracle(pin, ppat)
    h, n, c = o.solve(3, 1e-6)
    assert np.isfinite(h[0]) and abs(h[0] - 20.986794646572463) < 1e-9
    assert not np.isfinite(h[1]) and n == 3 and not c
    o.close()


@pytest.mark.gpu
def test_cfg1_seven_levels_gpu_same_divergence(tmp_path):
    """FP64 GPU: bit-identical to the oracle through the first V-cycle on every
    level (the non-finite face and the NaNs it spreads in the same places), the
    same first non-finite cycle, and MGF_ERR_NONFINITE from mgf_solve."""
    import mgfvm
    from mgfvm import _lib
    p = random_1024(ngrid=7).replace(nvcycle=3)
    pin, ppat = p.write(str(tmp_path / "c1"))
    _say("[gpu] 1024^2, 7 levels: first V-cycle vs oracle, then the non-finite cycle")
    s = mgfvm.Solver(pin, ppat)
    o = orc.Oracle(pin, ppat)
    eg = s.vcycle(0)
    eo = o.vcycle(0)
    assert np.isfinite(eg) and abs(eg - eo) <= 1e-12 * abs(eo)
    for ig in range(o.ngrid):
        for n in STATE:
            g, r = s.get(ig, n), o.get(ig, n)
            # NaN sign and payload bits are not IEEE-specified: the
            # non-finite entries must sit in the same places, every other
            # value must be bit-identical
            assert np.array_equal(np.isnan(g), np.isnan(r)), (ig, n)
            fin = ~np.isnan(r)
            assert np.array_equal(g[fin].view(np.uint64), r[fin].view(np.uint64)), (ig, n)
    assert not np.isfinite(s.get(5, "u")[ZERO_AP_FACE[1], ZERO_AP_FACE[0]])
    o.close()
    s.close()

    s = mgfvm.Solver(pin, ppat)
    hist = np.zeros(3)
    nc, cv = ctypes.c_int(), ctypes.c_int()
    rc = _lib.lib().mgf_solve(s.h, 3, 1e-6, _lib._dptr(hist), ctypes.byref(nc), ctypes.byref(cv))
    assert rc == -8, rc                      # MGF_ERR_NONFINITE
    assert "non-finite" in _lib.lib().mgf_last_error(s.h).decode()
    assert nc.value == 3 and cv.value == 0
    assert np.isfinite(hist[0]) and abs(hist[0] - eo) <= 1e-12 * abs(eo)
    assert not np.isfinite(hist[1])
    s.close()


@pytest.mark.gpu
def test_cfg1_five_levels_mixed_tracks_fp64(tmp_path):
    """Mixed precision at 1024^2 (5 levels): within 1e-6 of the FP64 oracle
    (relative to each field's magnitude) after 8 V-cycles, and the whole
    solve stops after the same number of V-cycles as the FP64 solve."""
    import mgfvm
    short = random_1024(ngrid=5).replace(nvcycle=8, convergence=0.0)
    pin, ppat = short.write(str(tmp_path / "short"))
    _say("[gpu] 1024^2, 5 levels: 8 mixed V-cycles vs the FP64 oracle")
    s = mgfvm.Solver(pin, ppat, precision=mgfvm.PREC_MIXED)
    h, n, c = s.solve()
    o = orc.Oracle(pin, ppat)
    ho, no, co = o.solve(8, 0.0)
    assert (n, c) == (no, co) == (8, False)
    np.testing.assert_allclose(h, ho, rtol=1e-5, atol=1e-12)
    for name in ("u", "v", "p", "sc"):
        g, r = s.get(0, name), o.get(0, name)
        scale = max(np.abs(r).max(), 1e-300)
        assert np.abs(g - r).max() / scale < 1e-6, name
    o.close()
    s.close()

    full = random_1024(ngrid=5)
    pin, ppat = full.write(str(tmp_path / "full"))
    _say("[gpu] 1024^2, 5 levels: full FP64 and mixed solves")
    counts = {}
    fields = {}
    for prec in (mgfvm.PREC_FP64, mgfvm.PREC_MIXED):
        with mgfvm.Solver(pin, ppat, precision=prec) as s:
            h, n, c = s.solve()
            counts[prec] = (n, c)
            fields[prec] = {k: s.get(0, k) for k in ("u", "v", "sc")}
    assert counts[mgfvm.PREC_FP64] == counts[mgfvm.PREC_MIXED]
    assert counts[mgfvm.PREC_FP64][1]
    for k in ("u", "v", "sc"):
        r = fields[mgfvm.PREC_FP64][k]
        err = np.abs(fields[mgfvm.PREC_MIXED][k] - r).max() / np.abs(r).max()
        assert err < 1e-6, (k, err)

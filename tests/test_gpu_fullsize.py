"""Parity at the workloads' full sizes, on the first V-cycles (the oracle
runs ~0.4 s per 1024^2 cycle, so whole solves stay with the smaller cases):

* the 1024^2 single problem (64x64 Bernoulli(0.8) mask, Re 200, red-black
  2/6, 5 levels): FP64 state bit-identical to the oracle after 3 V-cycles on
  every level, history within 1e-12 relative;
* the largest member shape of the 256-problem batch (512^2, 6 levels, a
  staggered array with quarter-cell shifts, so its coarse levels use the
  coarsened obstacles), solved inside a batch with 128^2 and 256^2 members:
  FP64 bit-identical to the oracle after 4 cycles, and the mixed-precision
  batch within the mixed tolerance of the FP64 oracle after 12 cycles.
"""
import sys

import numpy as np
import pytest

import orc
from mgfvm.inputs import random_1024, staggered_batch

pytestmark = pytest.mark.gpu

STATE = ["u", "v", "p", "pp", "sc", "resu", "resv", "ressc", "apu", "apv"]
MIXED_FIELD_RTOL = 1e-6


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _bitwise(get, o, levels, tag):
    for ig in levels:
        for n in STATE:
            g, r = get(ig, n), o.get(ig, n)
            assert np.array_equal(g.view(np.uint64), r.view(np.uint64)), (tag, ig, n)


def test_1024_fp64_first_cycles_bitwise(tmp_path):
    import mgfvm
    p = random_1024(ngrid=5).replace(nvcycle=3, convergence=0.0)
    pin, ppat = p.write(str(tmp_path / "c1"))
    _say("[gpu] 1024^2, 5 levels, 3 FP64 V-cycles vs oracle")
    s = mgfvm.Solver(pin, ppat)
    h, n, c = s.solve()
    o = orc.Oracle(pin, ppat)
    ho, no, co = o.solve(3, 0.0)
    assert (n, c) == (no, co) == (3, False)
    np.testing.assert_allclose(h, ho, rtol=1e-12, atol=0.0)
    _bitwise(s.get, o, range(o.ngrid), "1024")
    o.close()
    s.close()


def _batch_members():
    """one problem of each grid size from the benchmark's batch generator,
    the 512^2 one with quarter-cell shifts (coarsened obstacles)"""
    probs = staggered_batch(64)
    pick = {}
    for q in probs:
        if q.nfx not in pick and (q.nfx != 512 or q.ig_obs < q.ngrid):
            pick[q.nfx] = q
    return [pick[n] for n in (128, 256, 512)]


def test_batch_512_member_fp64_bitwise_and_mixed_tracks(tmp_path):
    import mgfvm
    members = _batch_members()
    assert members[2].ig_obs < members[2].ngrid
    fp = [q.replace(nvcycle=4, convergence=0.0) for q in members]
    pairs = [q.write(str(tmp_path / f"f{k}")) for k, q in enumerate(fp)]
    _say("[gpu] batch 128/256/512, 4 FP64 V-cycles vs oracle")
    bt = mgfvm.Batch(pairs)
    bt.solve()
    for k, (pin, ppat) in enumerate(pairs):
        o = orc.Oracle(pin, ppat)
        ho, _, _ = o.solve(4, 0.0)
        np.testing.assert_allclose(bt.history(k), ho, rtol=1e-12, atol=0.0)
        _bitwise(lambda ig, n, k=k: bt.get(k, ig, n), o, range(o.ngrid), k)
        o.close()
    bt.close()
    mx = [q.replace(nvcycle=12, convergence=0.0) for q in members]
    pairs = [q.write(str(tmp_path / f"m{k}")) for k, q in enumerate(mx)]
    _say("[gpu] batch 128/256/512, 12 mixed V-cycles vs FP64 oracle")
    bt = mgfvm.Batch(pairs, precision=mgfvm.PREC_MIXED)
    bt.solve()
    for k, (pin, ppat) in enumerate(pairs):
        o = orc.Oracle(pin, ppat)
        ho, _, _ = o.solve(12, 0.0)
        np.testing.assert_allclose(bt.history(k), ho, rtol=1e-5, atol=1e-12)
        for n in ("u", "v", "p", "sc"):
            g, r = bt.get(k, 0, n), o.get(0, n)
            scale = max(np.abs(r).max(), 1e-300)
            assert np.abs(g - r).max() / scale < MIXED_FIELD_RTOL, (k, n)
        o.close()
    bt.close()

"""The project-written autotuner (mgf_autotune): it measures and applies
launch shapes and a lane count, writes a table that mgf_options.tuning
reads back, and -- every shape being result-neutral -- a tuned batch gives
the same bits as an untuned one (histories and every state array of every
level of every problem)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATE = ["u", "v", "p", "pp", "sc", "resu", "resv", "ressc", "apu", "apv"]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _pairs(tmp_path):
    from mgfvm.inputs import staggered_batch
    probs = staggered_batch(24)
    pick, seen = [], {}
    for p in probs:
        if seen.get(p.nfx, 0) < 2:
            seen[p.nfx] = seen.get(p.nfx, 0) + 1
            pick.append(p.replace(nvcycle=30, convergence=0.0))
    return [p.write(str(tmp_path / f"p{k}")) for k, p in enumerate(pick)]


def _state(b, n):
    out = {}
    for k in range(n):
        out[("h", k)] = b.history(k)
        for ig in range(b.num_levels(k)):
            for f in STATE:
                out[(k, ig, f)] = b.get(k, ig, f)
    return out


def _same(a, b):
    for key in a:
        assert np.array_equal(np.asarray(a[key]).view(np.uint64),
                              np.asarray(b[key]).view(np.uint64)), key


def test_autotune_result_neutral_and_table_round_trip(tmp_path):
    import mgfvm
    pairs = _pairs(tmp_path)
    _say("[gpu] autotune: untuned reference batch")
    with mgfvm.Batch(pairs, precision=mgfvm.PREC_MIXED) as b0:
        b0.solve()
        ref = _state(b0, len(pairs))
    _say("[gpu] autotune: tune, then solve")
    with mgfvm.Batch(pairs, precision=mgfvm.PREC_MIXED) as b1:
        table = b1.autotune(reps=5)
        lines = [ln for ln in table.splitlines() if ln and not ln.startswith("#")]
        assert any(ln.startswith("level 514 514 ") for ln in lines)
        assert all(ln.split()[0] in ("lanes", "level") for ln in lines)
        b1.solve()
        _same(ref, _state(b1, len(pairs)))
    _say("[gpu] autotune: batch created with the table")
    with mgfvm.Batch(pairs, precision=mgfvm.PREC_MIXED, tuning=table) as b2:
        b2.solve()
        _same(ref, _state(b2, len(pairs)))
    # every shape forced on every level gives the same bits too
    for uvs, pp, nt in ((1, 1, 512), (2, 2, 256), (2, 1, 512)):
        forced = "lanes 2\n" + "".join(
            f"level {n + 2} {n + 2} {uvs} {pp} {nt}\n" for n in (16, 32, 64, 128, 256, 512))
        with mgfvm.Batch(pairs, precision=mgfvm.PREC_MIXED, tuning=forced) as b3:
            b3.solve()
            _same(ref, _state(b3, len(pairs)))


def test_autotune_refuses_fp64(tmp_path):
    import mgfvm
    from mgfvm import _lib
    pairs = _pairs(tmp_path)[:2]
    with mgfvm.Batch(pairs) as b:
        with pytest.raises(_lib.MgfError) as e:
            b.autotune(reps=2)
        assert e.value.code == -9

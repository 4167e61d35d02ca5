"""Seeded pattern generators and the reference file formats (host only)."""
import numpy as np

from mgfvm.inputs import (Problem, Rng, channel_128_inline, parse, random_1024, random_mask,
                          staggered_batch, staggered_pattern)


def test_rng_is_deterministic_and_uniform():
    a, b = Rng(7), Rng(7)
    xs = [a.uniform() for _ in range(20000)]
    assert xs == [b.uniform() for _ in range(20000)]
    assert abs(np.mean(xs) - 0.5) < 0.01
    assert Rng(8).uniform() != xs[0]


def test_random_mask_properties():
    rects = random_mask(64, 0.8, 1)
    cells = {(r[0], r[2]) for r in rects}
    assert all(r[0] == r[1] and r[2] == r[3] for r in rects)
    assert all(2 <= i <= 63 for i, _ in cells)          # edge columns kept free
    assert abs(len(cells) / (62 * 64) - 0.2) < 0.03
    # i-major then j ascending, the order the reference's generator writes
    assert rects == sorted(rects, key=lambda r: (r[0], r[2]))
    assert random_mask(64, 0.8, 1) == rects


def test_staggered_pattern_properties():
    for shift, scale in (("quarter", 4), ("half", 2), ("full", 1)):
        rects, s = staggered_pattern(16, 4, shift, 3)
        assert s == scale
        n = 16 * scale
        assert all(1 <= r[0] <= n and 1 <= r[2] <= n for r in rects)
        assert all(r[0] not in (1, n) for r in rects)
        assert len(set(rects)) == len(rects)


def test_staggered_batch_mix():
    probs = staggered_batch(256)
    sizes = {p.nfx for p in probs}
    assert sizes == {128, 256, 512}
    for p in probs:
        assert p.nfx >> (p.ngrid - 1) == 16                 # coarsest 16x16
        assert 1 <= p.ig_obs <= p.ngrid
        assert round(1 / p.amuf) in (10, 50, 100, 200, 400)
    assert staggered_batch(3)[2] == probs[2]


def test_named_workloads_roundtrip(tmp_path):
    for k, p in enumerate((channel_128_inline(), random_1024(ngrid=5))):
        pin, ppat = p.write(str(tmp_path / str(k)))
        assert parse(pin, ppat) == p

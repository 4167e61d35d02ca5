"""Seeded pattern generators of the native library (host only): the same
rectangles as the Python generators used to build the benchmark and test
inputs, the reference generator's own inline and staggered arrays (fixtures from
create_patterns_staggered.c:113-162), and the particle
placement's invariants (create_patterns_random.c:58-138)."""
import os
import subprocess

import pytest

import mgfvm
from mgfvm.inputs import random_mask, staggered_pattern

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "multigrid-fvm_amd")


@pytest.mark.parametrize("n,por,seed", [(8, 0.8, 1), (16, 0.5, 7), (64, 0.8, 1), (32, 0.95, 123)])
def test_bernoulli_matches_python_generator(n, por, seed):
    assert mgfvm.pattern_bernoulli(n, por, seed) == random_mask(n, por, seed)


@pytest.mark.parametrize("shift,scale", [("quarter", 4), ("half", 2), ("full", 1)])
@pytest.mark.parametrize("seed", [3, 1000, 1255])
def test_perturbed_staggered_matches_python_generator(shift, scale, seed):
    rects, s = staggered_pattern(16, 4, shift, seed)
    assert s == scale
    assert mgfvm.pattern_staggered(16, 4, 4, scale, 3, seed) == rects


def _cells(rects):
    return {(i, j) for a, b, c, d in rects for i in range(a, b + 1) for j in range(c, d + 1)}


def _fixture_params():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "patterns_staggered.json")) as f:
        return json.load(f)


def _parse_pattern(text):
    lines = text.split("\n")
    n = int(lines[0])
    rects = [tuple(int(t) for t in ln.split()) for ln in lines[1:1 + n]]
    assert len(rects) == n
    return rects


@pytest.mark.parametrize("rec", _fixture_params(),
                         ids=lambda r: f"n{r['n_base']}_s{r['scale']}_x{r['spacing_x']}y{r['spacing_y']}")
def test_inline_and_staggered_arrays_equal_reference_generator(rec):
    """Variants 1 and 2 (the inline and the staggered array) equal the
    reference generator's pattern_1 and pattern_2 line for line: the files
    create_patterns_staggered.c writes for the same base grid, scale and
    spacings (fixture written by tests/golden/make_pattern_golden.py from
    the reference's generator).  With free edge columns, the same minus the
    rectangles in the first and last i row."""
    n = rec["n_base"] * rec["scale"]
    for variant in (1, 2):
        want = _parse_pattern(rec[f"pattern_{variant}"])
        got = mgfvm.pattern_staggered(rec["n_base"], rec["spacing_x"], rec["spacing_y"],
                                      rec["scale"], variant, 0, False)
        assert got == want, variant
        free = mgfvm.pattern_staggered(rec["n_base"], rec["spacing_x"], rec["spacing_y"],
                                       rec["scale"], variant, 0, True)
        assert free == [r for r in want if r[0] not in (1, n)], variant


def test_perturbed_moves_each_obstacle_at_most_one_unit():
    base = _cells(mgfvm.pattern_staggered(16, 4, 4, 4, 2, 0, False))
    pert = _cells(mgfvm.pattern_staggered(16, 4, 4, 4, 3, 99, False))
    # same number of blocked cells up to obstacles moved across the border
    assert abs(len(base) - len(pert)) <= 16 * 4
    assert pert != base
    assert mgfvm.pattern_staggered(16, 4, 4, 4, 3, 99, False) == \
        mgfvm.pattern_staggered(16, 4, 4, 4, 7, 99, False)      # variants >= 3 differ by seed only


@pytest.mark.parametrize("n,ratio,seed", [(32, 0.2, 1), (64, 0.3, 5), (50, 0.1, 9)])
def test_particles_invariants(n, ratio, seed):
    rects = mgfvm.pattern_particles(n, ratio, 50, 30, 20, seed)
    sizes = [b - a + 1 for a, b, c, d in rects]
    assert all(b - a == d - c for a, b, c, d in rects)
    assert set(sizes) <= {1, 2, 3}
    assert sizes == sorted(sizes, reverse=True)              # largest placed first
    cells = [(i, j) for a, b, c, d in rects for i in range(a, b + 1) for j in range(c, d + 1)]
    assert len(cells) == len(set(cells))                      # no overlap
    assert all(2 <= i <= n - 1 and 1 <= j <= n for i, j in cells)   # edge columns free
    total = round(n * n * ratio / (0.5 + 0.3 * 4 + 0.2 * 9))
    assert sizes.count(3) <= round(total * 0.2) and sizes.count(1) <= round(total * 0.5)
    assert abs(len(cells) / (n * n) - ratio) < 0.02
    assert mgfvm.pattern_particles(n, ratio, 50, 30, 20, seed) == rects


def test_generator_argument_errors():
    with pytest.raises(mgfvm.MgfError):
        mgfvm.pattern_particles(6, 0.2, 50, 30, 20, 1)
    with pytest.raises(mgfvm.MgfError):
        mgfvm.pattern_particles(16, 0.2, 50, 30, 30, 1)
    with pytest.raises(mgfvm.MgfError):
        mgfvm.pattern_staggered(16, 2, 2, 3, 1, 1)


def test_patterns_cli_writes_readable_files(tmp_path):
    exe = os.path.join(PKG, "mgfvm_patterns")
    out = tmp_path / "pats"
    r = subprocess.run([exe, "staggered", "--n-base", "16", "--spacing", "4", "--shift", "half",
                        "--count", "4", "--seed", "10", "--out", str(out)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    man = (out / "manifest.txt").read_text().splitlines()
    assert len(man) == 5 and man[1].startswith("pattern_1.txt staggered")
    from mgfvm.inputs import channel_128_inline
    pin, _ = channel_128_inline().replace(ig_obs=3).write(str(tmp_path / "case"))
    for k in range(1, 5):
        txt = (out / f"pattern_{k}.txt").read_text().split()
        rects = [tuple(map(int, txt[1 + 4 * q:5 + 4 * q])) for q in range(int(txt[0]))]
        assert rects == mgfvm.pattern_staggered(16, 4, 4, 2, k, 10 + k)
        assert (out / f"pattern_{k}.pgm").read_text().startswith("P2\n32 32\n255\n")
        p = mgfvm._lib.read_problem(pin, str(out / f"pattern_{k}.txt"))
        assert p.contents.nobs == len(rects)
        mgfvm._lib.free_problem(p)
    bad = subprocess.run([exe, "staggered", "--count", "0"], capture_output=True, text=True)
    assert bad.returncode == 1


def test_batch_cli_input_errors(tmp_path):
    exe = os.path.join(PKG, "mgfvm_batch")
    r = subprocess.run([exe], cwd=str(tmp_path), capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr
    (tmp_path / "patterns").mkdir()
    r = subprocess.run([exe, "--re", "20"], cwd=str(tmp_path), capture_output=True, text=True)
    assert r.returncode == 1 and "no pattern_" in r.stderr
    r = subprocess.run([exe, "--re", "x"], cwd=str(tmp_path), capture_output=True, text=True)
    assert r.returncode == 1

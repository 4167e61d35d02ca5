"""The command-line tools on the GPU: mgfvm_batch (the batched replacement of
run_patterns_and_save_vtk.sh) writes, for every (Re, pattern) pair, the same
VTK and convergence files as mgfvm_run (the drop-in for the reference
program) run on that case alone."""
import os
import subprocess
import sys

import pytest

from mgfvm.inputs import Problem

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "multigrid-fvm_amd")


def test_batch_driver_matches_single_runs(tmp_path):
    print("[gpu] mgfvm_batch vs mgfvm_run", flush=True)
    sys.stdout.flush()
    pats = tmp_path / "patterns"
    r = subprocess.run([os.path.join(PKG, "mgfvm_patterns"), "staggered", "--n-base", "16",
                        "--spacing", "4", "--shift", "half", "--count", "3", "--seed", "5",
                        "--out", str(pats)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    tmpl = Problem(nfx=64, nfy=64, ngrid=3, ig_obs=2, nvcycle=40, nswpm=2, nswpp=3, isolve=2,
                   amuf=0.01)
    tmpl.write(str(tmp_path))
    r = subprocess.run([os.path.join(PKG, "mgfvm_batch"), "--re", "20,50", "--input", "InputData",
                        "--patterns", "patterns", "--out", "solution"],
                       cwd=str(tmp_path), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("-> pattern_") == 6
    for re_, amuf in (("20", 0.05), ("50", 0.02)):
        for k in (1, 2, 3):
            case = tmp_path / f"single_{re_}_{k}"
            tmpl.replace(amuf=amuf).write(str(case))
            os.replace(str(case / "pattern.txt"), str(case / "unused.txt"))
            with open(pats / f"pattern_{k}.txt") as f, open(case / "pattern.txt", "w") as g:
                g.write(f.read())
            s = subprocess.run([os.path.join(PKG, "mgfvm_run")], cwd=str(case),
                               capture_output=True, text=True, timeout=300)
            assert s.returncode == 0, s.stderr
            sol = tmp_path / "solution" / f"Re_{re_}"
            assert (sol / "vtk_files" / f"vtk_{k}.vtk").read_bytes() == \
                (case / "vtk_acc.vtk").read_bytes(), (re_, k)
            assert (sol / "convergence_files" / f"convergence_{k}.csv").read_bytes() == \
                (case / "convergence.csv").read_bytes(), (re_, k)


def _out_fixtures():
    import glob
    return sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "out_*.json")))


@pytest.mark.parametrize("path", _out_fixtures(),
                         ids=[os.path.basename(p)[4:-5] for p in _out_fixtures()])
def test_driver_outputs_equal_reference_program(tmp_path, path):
    """mgfvm_run (the drop-in for the reference program) writes the same
    vtk_acc.vtk and convergence.csv, byte for byte, and the same stdout
    (input echo, one `icycle = %d error = %e` line per cycle, averages,
    record count) as the reference program on the same inputs; only the
    timing line's value differs.  Fixtures written by
    tests/golden/make_output_golden.py from the unmodified reference."""
    import hashlib
    import json
    with open(path) as f:
        fx = json.load(f)
    print(f"[gpu] mgfvm_run outputs vs reference: {fx['case']}", flush=True)
    (tmp_path / "InputData").write_text(fx["inputdata"])
    (tmp_path / "pattern.txt").write_text(fx["pattern"])
    r = subprocess.run([os.path.join(PKG, "mgfvm_run")], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from make_output_golden import mask_timing
    assert mask_timing(r.stdout) == fx["stdout"]
    for name, key in (("vtk_acc.vtk", "vtk_sha256"), ("convergence.csv", "csv_sha256")):
        got = hashlib.sha256((tmp_path / name).read_bytes()).hexdigest()
        assert got == fx[key], name

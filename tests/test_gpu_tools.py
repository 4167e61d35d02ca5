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

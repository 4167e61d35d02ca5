"""Write the fixtures that pin the drop-in program's outputs to the
reference program's.

Runs the unmodified reference program (oracle/_ref/ref_plain, built from the
reference's own source by oracle/ref/Makefile) on named cases and records
what it writes: the SHA-256 of vtk_acc.vtk and of convergence.csv
(Version-6-26-04.c:2261-2363, 2544-2558) and its stdout, with the value of
the timing line (V6:256-258) masked.  Fixtures out_<case>.json (data only)
are written next to this script.

usage: python tests/golden/make_output_golden.py [--reference /root/reference]
"""
import argparse
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "multigrid-fvm_amd"))

import cases  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "ref_plain")
# a converging obstacle case, a non-converged one, discrete inlets/outlets
# with Jacobi, and red-black on the small grid with touching obstacles
OUTPUT_CASES = ["cfg0_rb", "stag128", "inlets", "small_rb"]
TIMING = re.compile(r"^simple solution took [0-9.]+ seconds to execute $")


def mask_timing(stdout: str):
    return ["simple solution took <t> seconds to execute " if TIMING.match(ln) else ln
            for ln in stdout.splitlines()]


def sha256(path: str) -> str:
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    args = ap.parse_args()
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle", "ref"), f"REF={args.reference}"],
                   check=True)
    with tempfile.TemporaryDirectory() as work:
        for name in OUTPUT_CASES:
            pin, ppat = cases.write_case(name, work)
            d = os.path.dirname(pin)
            r = subprocess.run([REF], cwd=d, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"reference failed on {name}: {r.stderr[-500:]}")
            fx = {"case": name,
                  "inputdata": open(pin).read(), "pattern": open(ppat).read(),
                  "vtk_sha256": sha256(os.path.join(d, "vtk_acc.vtk")),
                  "csv_sha256": sha256(os.path.join(d, "convergence.csv")),
                  "stdout": mask_timing(r.stdout)}
            with open(os.path.join(HERE, f"out_{name}.json"), "w") as f:
                json.dump(fx, f, indent=1)
            print(f"out {name}: {len(fx['stdout'])} stdout lines")


if __name__ == "__main__":
    main()

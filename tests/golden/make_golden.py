"""Write the golden fixtures that pin the CPU restatement to the reference.

Runs in a container where the reference's source is mounted (it is read
there, compiled by oracle/ref/Makefile into oracle/_ref/, never copied):

  ref_a  the unmodified reference program, state dumped after its V-cycle
         loop (link-time wrap of malloc/clock, oracle/ref/harness_a.c);
  ref_b  the reference's own routines called one at a time on a staged
         state (oracle/ref/harness_b.c).

Fixtures written next to this script (data only: the inputs, residual
histories as exact hex floats, SHA-256 digests of every array of every
level, and for the per-routine cases the staged state):

  e2e_<case>.json   whole solve (ref_a)
  kat_<case>.json   per-routine digests after each routine (ref_b)
  kat_<case>.npz    the staged state those routines start from

usage: python tests/golden/make_golden.py [--reference /root/reference] [--only e2e_x ...]
"""
import argparse
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "multigrid-fvm_amd"))

import cases  # noqa: E402
import orc  # noqa: E402
from mgfvm.inputs import random_1024, random_mask, Problem  # noqa: E402

REF_DIR = os.path.join(ROOT, "oracle", "_ref")

# whole-solve cases: every named case plus the large oracle-safe random mask
E2E = {k: v for k, v in cases.CASES.items() if k not in cases.EXTENSION_CASES}
E2E["rand1024_l5"] = random_1024(ngrid=5).replace(nvcycle=3)
E2E["rand256_l3"] = Problem(nfx=256, nfy=256, ngrid=3, ig_obs=3, isolve=2, nvcycle=15,
                            nswpp=6, amuf=0.005, obstacles=random_mask(64, 0.8, 1))
# per-routine cases (small grids, every branch)
KAT = ["small_jac", "small_rb", "ibc2", "ibc3", "ibc4", "inlets"]


def digest(a: np.ndarray) -> str:
    """First 128 bits of the SHA-256 of the array's float64 bytes."""
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()[:32]


def level_digests(levels):
    return [{name: digest(lv[name]) for name in orc.DUMP_NAMES} for lv in levels]


def run_e2e(name, prob, workdir):
    d = os.path.join(workdir, name)
    prob.write(d)
    env = dict(os.environ, MGF_REF_DUMP=os.path.join(d, "ref.bin"))
    r = subprocess.run([os.path.join(REF_DIR, "ref_a")], cwd=d, env=env, capture_output=True,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError(f"reference failed on {name}: {r.stderr[-500:]}")
    levels, hist = orc.read_dump(os.path.join(d, "ref.bin"))
    icyc_lines = [ln for ln in r.stdout.splitlines() if ln.startswith("icycle = ")]
    return {
        "case": name,
        "inputdata": prob.inputdata_text(),
        "pattern": prob.pattern_text(),
        "ngrid": len(levels),
        "dims": [list(lv["u"].shape[::-1]) for lv in levels],
        "ncycles": int(len(hist)),
        "converged": bool(len(hist) > 0 and hist[-1] < prob.convergence),
        "history_hex": [float(h).hex() for h in hist],
        "stdout_icycle_first_last": [icyc_lines[0], icyc_lines[-1]] if icyc_lines else [],
        "digests": level_digests(levels),
    }


def run_kat(name, workdir, nstage=3):
    prob = cases.CASES[name]
    d = os.path.join(workdir, "kat_" + name)
    prob.write(d)
    r = subprocess.run([os.path.join(REF_DIR, "ref_b"), str(nstage), "kat.bin"], cwd=d,
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"reference harness B failed on {name}: {r.stderr[-500:]}")
    recs = orc.read_kat(os.path.join(d, "kat.bin"), prob.ngrid)
    stage = recs[0][3]
    npz = {f"L{ig}_{n}": lv[n] for ig, lv in enumerate(stage) for n in orc.DUMP_NAMES}
    stage_d = {f"L{ig}.{n}": d for ig, lv in enumerate(level_digests(stage)) for n, d in lv.items()}
    routines = []
    for rname, lvl, ret, levels in recs[1:]:
        after = {f"L{ig}.{n}": d for ig, lv in enumerate(level_digests(levels)) for n, d in lv.items()}
        changed = {k: d for k, d in after.items() if d != stage_d[k]}
        routines.append({"routine": rname, "level": lvl, "ret_hex": float(ret).hex(),
                         "changed": changed})
    meta = {"case": name, "inputdata": prob.inputdata_text(), "pattern": prob.pattern_text(),
            "ngrid": prob.ngrid, "stage_vcycles": nstage, "stage_digests": stage_d,
            "routines": routines}
    return meta, npz


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    ap.add_argument("--only", nargs="*", default=None,
                    help="write only these fixtures (e2e_<case> / kat_<case> names)")
    args = ap.parse_args()
    want = (lambda tag: args.only is None or tag in args.only)
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle", "ref"), f"REF={args.reference}"],
                   check=True)
    with tempfile.TemporaryDirectory() as work:
        for name, prob in E2E.items():
            if not want(f"e2e_{name}"):
                continue
            fx = run_e2e(name, prob, work)
            with open(os.path.join(HERE, f"e2e_{name}.json"), "w") as f:
                json.dump(fx, f, indent=1)
            print(f"e2e {name}: {fx['ncycles']} cycles, converged={fx['converged']}")
        for name in KAT:
            if not want(f"kat_{name}"):
                continue
            meta, npz = run_kat(name, work)
            with open(os.path.join(HERE, f"kat_{name}.json"), "w") as f:
                json.dump(meta, f, indent=1)
            np.savez_compressed(os.path.join(HERE, f"kat_{name}.npz"), **npz)
            print(f"kat {name}: {len(meta['routines'])} routine records")


if __name__ == "__main__":
    main()

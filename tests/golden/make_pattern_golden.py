"""Write the fixtures that pin the seeded pattern generator to the
reference's staggered generator.

create_patterns_staggered.c draws no random numbers for its first two
patterns (the inline array, pattern_1, and the staggered array, pattern_2;
create_patterns_staggered.c:127-162): they depend only on the base grid,
the shift scale and the spacings it reads from stdin.  This script runs the
reference generator (built from its own source by oracle/ref/Makefile into
oracle/_ref/ref_patterns_staggered) with zero perturbed patterns and stores
both pattern files' text for each parameter set in patterns_staggered.json
(data only).

usage: python tests/golden/make_pattern_golden.py [--reference /root/reference]
"""
import argparse
import glob
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
GEN = os.path.join(ROOT, "oracle", "_ref", "ref_patterns_staggered")

# (n_base, shift choice 1/2/3 = scale 4/2/1, spacing x, spacing y)
PARAMS = [(16, 3, 4, 4), (16, 2, 4, 4), (16, 1, 4, 4), (16, 3, 3, 3), (12, 2, 2, 3),
          (20, 1, 3, 2)]
SCALE = {1: 4, 2: 2, 3: 1}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    args = ap.parse_args()
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle", "ref"), f"REF={args.reference}"],
                   check=True)
    out = []
    for n, choice, sx, sy in PARAMS:
        with tempfile.TemporaryDirectory() as d:
            r = subprocess.run([GEN], cwd=d, input=f"{n}\n{choice}\n{sx}\n{sy}\n0\n",
                               capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"generator failed for {(n, choice, sx, sy)}")
            (sub,) = glob.glob(os.path.join(d, "n*"))
            rec = {"n_base": n, "scale": SCALE[choice], "spacing_x": sx, "spacing_y": sy,
                   "directory": os.path.basename(sub)}
            for k in (1, 2):
                with open(os.path.join(sub, f"pattern_{k}.txt")) as f:
                    rec[f"pattern_{k}"] = f.read()
            out.append(rec)
            print(f"n={n} scale={SCALE[choice]} sx={sx} sy={sy}: "
                  f"{rec['pattern_1'].split()[0]} / {rec['pattern_2'].split()[0]} pixels")
    with open(os.path.join(HERE, "patterns_staggered.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()

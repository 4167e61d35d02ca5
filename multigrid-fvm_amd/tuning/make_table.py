#!/usr/bin/env python3
"""Run the library's autotuner (mgf_autotune) on a workload and write the
table it measures.  The batch benchmark's table is made with

    python multigrid-fvm_amd/tuning/make_table.py --workload batch256 \
        > multigrid-fvm_amd/tuning/batch256_mi355x.txt

on one MI355X; bench.py passes that file's text as mgf_options.tuning.
Shapes are result-neutral: the table changes timings, never results.
"""
import argparse
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="batch256", choices=["batch256", "adjoint64"])
    ap.add_argument("--reps", type=int, default=20)
    args = ap.parse_args()
    import torch
    torch.cuda.is_available()
    import mgfvm
    from mgfvm.inputs import adjoint_batch, staggered_batch
    probs = staggered_batch(256) if args.workload == "batch256" else adjoint_batch(64)
    with tempfile.TemporaryDirectory() as d:
        pairs = [p.write(os.path.join(d, f"p{k}")) for k, p in enumerate(probs)]
        with mgfvm.Batch(pairs, precision=mgfvm.PREC_MIXED) as b:
            table = b.autotune(args.reps)
    sys.stdout.write(f"# workload {args.workload}, reps {args.reps}\n" + table)


if __name__ == "__main__":
    main()

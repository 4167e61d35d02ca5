#!/usr/bin/env python3
"""Converged-solve throughput of the MI355X SIMPLE/FAS-multigrid solver.

A step is one complete steady solve from the initial state to the stopping
test (level-0 sum |div| < 1e-6, FP64), inputs already resident in HBM: the
solver is created once, and every step restores its initial state on the
device (mgf_reset) and runs graph-replayed V-cycles until the device-side
stopping test fires (mgf_solve).

Default workload: 1024x1024 channel, random obstacle mask on a 64x64 base
grid (Bernoulli, porosity 0.8, seed 1; each blocked base cell is 16x16 fine
cells), Re 200, red-black Gauss-Seidel 2/6 sweeps, 5 multigrid levels
(coarsest 64x64, the obstacle level).  With more levels the coarse grids
cannot represent the obstacles and the reference's algorithm diverges
(7 levels) -- see DESIGN.md.

Printed: one JSON line with the throughput, the roofline of the dominant
kernel (one red-black sweep of the p' equation on level 0, both colours in
one launch, HIP events on the solver's stream) and a CPU baseline (the reference program built from
its own source, single core, a bounded number of V-cycles of the same
solve, scaled to the cycles the solve needs).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "multigrid-fvm_amd"))

METRIC = ("converged solves/s at FP64 residual 1e-6, batched-mixed vs plain GPU; "
          "% HBM peak")
HBM_PEAK_GBS = 8000.0


def workloads():
    from mgfvm.inputs import channel_128_inline, random_1024
    return {
        "rand1024_l5": (random_1024(ngrid=5),
                        "1024x1024 channel, 64x64 Bernoulli(porosity 0.8, seed 1) obstacle mask, "
                        "Re 200, red-black 2/6, 5 levels (coarsest = obstacle level 64x64)"),
        "rand1024_l6": (random_1024(ngrid=6),
                        "1024x1024 channel, 64x64 Bernoulli(porosity 0.8, seed 1) obstacle mask, "
                        "Re 200, red-black 2/6, 6 levels (one level below the obstacle level)"),
        "inline128_rb": (channel_128_inline().replace(isolve=2),
                         "128x128 channel, inline 4x4 single-cell blocks on 16x16, Re 100, "
                         "red-black 2/4, 4 levels"),
    }


def cpu_baseline(prob, ncycles_full, workdir, target_s=12.0):
    """Time the reference program (oracle/_ref/ref_plain, built from the
    reference's own source) or, where it was not built, the CPU restatement
    (oracle/orc_run), single core, on the same inputs for a bounded number of
    V-cycles; scale per-cycle time to the cycles the full solve takes."""
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_plain")
    port = os.path.join(ROOT, "oracle", "orc_run")
    kind = "reference" if os.path.exists(ref) else "port"
    exe = ref if kind == "reference" else port
    if not os.path.exists(exe):
        return None

    def run(ncyc):
        d = os.path.join(workdir, f"cpu_{ncyc}")
        shutil.rmtree(d, ignore_errors=True)
        prob.replace(nvcycle=ncyc, convergence=0.0).write(d)
        taskset = shutil.which("taskset")
        pre = [taskset, "-c", "0"] if taskset else []
        t0 = time.perf_counter()
        if kind == "reference":
            r = subprocess.run(pre + [exe], cwd=d, capture_output=True, text=True)
            wall = time.perf_counter() - t0
            m = re.search(r"simple solution took ([0-9.eE+-]+) seconds", r.stdout)
            secs = float(m.group(1)) if m else wall
        else:
            r = subprocess.run(pre + [exe, "InputData", "pattern.txt", os.path.join(d, "o.bin")],
                               cwd=d, capture_output=True, text=True)
            m = re.search(r"seconds ([0-9.eE+-]+)", r.stderr)
            secs = float(m.group(1)) if m else time.perf_counter() - t0
        return secs

    probe = 2
    t_probe = run(probe)
    per = t_probe / probe
    n = max(probe, min(ncycles_full, int(target_s / max(per, 1e-6))))
    t = run(n) if n != probe else t_probe
    per_cycle = t / n
    solve_s = per_cycle * ncycles_full
    return {"value": 1.0 / solve_s, "unit": "solves/s", "cores": 1, "kind": kind,
            "sample": f"{n} of the {ncycles_full} V-cycles of the same solve on 1 core "
                      f"({per_cycle * 1e3:.1f} ms/V-cycle, V-cycle loop time only); "
                      f"solve time = per-cycle time x {ncycles_full}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rand1024_l5")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cycles", type=int, default=0,
                    help="cap V-cycles per solve (profiling runs only; the line is then not "
                         "a converged-solve figure)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(instances are independent and shard by instance; see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import mgfvm

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)
    prob, desc = workloads()[args.workload]
    work = tempfile.mkdtemp(prefix="mgf_bench_")
    try:
        pin, ppat = prob.write(os.path.join(work, "case"))
        s = mgfvm.Solver(pin, ppat, device=0)
        nv, tol = prob.nvcycle, prob.convergence
        if args.cycles > 0:
            nv, tol = args.cycles, 0.0
        ncyc = conv = None
        for _ in range(args.warmup):
            s.reset()
            _, ncyc, conv = s.solve(nv, tol)
        torch.cuda.synchronize()
        s.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            s.reset()
            hist, ncyc, conv = s.solve(nv, tol)
        s.synchronize()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        if not conv:
            print(f"bench.py: workload did not converge in {nv} V-cycles", file=sys.stderr)
        # dominant kernel: one red-black sweep (both colours) of the p' equation on level 0
        us, nbytes = s.time_kernel(mgfvm.KERNEL_SWEEP_PP, 0, 200)
        achieved = nbytes / (us * 1e-6) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_rb_sweep (one red-black sweep of the p' equation, level 0)",
                "mean_us": round(us, 3), "bytes_per_launch": nbytes}
        s.close()
        cpu = None if args.no_cpu else cpu_baseline(prob, ncyc, work)
        value = args.steps / elapsed
        line = {
            "metric": METRIC, "value": round(value, 4), "unit": "solves/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(1e3 * elapsed / args.steps, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": desc, "name": args.workload, "v_cycles_per_solve": ncyc,
                       "converged": bool(conv), "precision": "fp64 (bit-exact)",
                       "ms_per_v_cycle": round(1e3 * elapsed / args.steps / max(ncyc, 1), 4)},
            "roofline": roof, "cpu_baseline": cpu,
        }
        print(json.dumps(line))
    finally:
        shutil.rmtree(work, ignore_errors=True)


if __name__ == "__main__":
    main()

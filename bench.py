#!/usr/bin/env python3
"""Converged-solve throughput of the MI355X SIMPLE/FAS-multigrid solver.

Metric: converged solves per second at the FP64 stopping test (level-0 sum
of |div| < 1e-6, evaluated in FP64), batched mixed precision vs the plain
FP64 GPU path, with the dominant kernel's fraction of HBM peak.

A step is one batched solve of a fixed set of 256 independent channel
problems (grids 128^2/256^2/512^2 with 16x16 coarsest level, staggered
obstacle arrays moved by quarter/half/full cells, Re in {10,50,100,200,400},
red-black 2/6; seeds 1000..1255), inputs already resident in HBM: the batch
is created once, each step restores the initial state on the device
(mgf_reset) and runs graph-replayed batched V-cycles until every problem has
met its stopping test or its 2999-cycle limit (mgf_batch_solve).  `value`
counts the problems that converged, per second.  Problems the reference's
algorithm does not converge (low Re on 512^2 with a 16^2 coarsest level,
identical in the CPU restatement) run their full cycle limit and count as
not converged.

Also reported: the plain FP64 GPU path (one problem at a time, the FP64
kernels, no batching), estimated over all problems by stratified sampling
(--plain-per-class solves per grid size and convergence class), and the
ratio; the same batch in FP64 (bit-exact with the reference) and the
mixed batch's per-field error against it; the converging problems solved
alone (the share of the step the non-converging ones take); a single
1024^2 solve (64x64 random mask, Re 200, 5 levels) in FP64 and mixed; the
reverse-mode gradients of the 256^2 differentiation workload (per
objective, against the primal solve and against the forward-mode batch of
all 258 directions); the batched adjoint of 64 512^2 problems (primal and
reverse solve times); the roofline of the dominant kernel (the merged u, v,
sc relaxation launch over the 512^2 members, HIP events on the batch's
stream; its HBM traffic read from the committed counter summary); and a
CPU baseline (the reference program built from its own source, one
single-threaded copy per core on up to 16 cores, per-V-cycle time per grid
size measured under that load on bounded runs, times the cycle counts of
the batch; the one-core figure beside it).

--roofline-only runs just the dominant kernel's launches and a 1 GiB device
copy, for the rocprofv3 counter passes that give `traffic`.  --tail-scaling
times V-cycles of 1 .. 32 of the 512^2 problems at Re 10/50 solved together
(the fixed latency chain and the per-problem part of the non-converging
tail that sets the step).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
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
# HBM bytes per launch of the roofline kernel (k_rb32x_uvs<2,512,32> over the 84
# 512^2 members): rocprofv3 FETCH_SIZE and WRITE_SIZE passes over
# `bench.py --roofline-only --reps 10`, each scaled by what the counter reads
# for a 1 GiB device copy, read from the committed summary of those passes
TRAFFIC_PROFILE = os.path.join("profiles", "traffic_relax_uvs_512x84.md")
# launch shapes and lane count of the batch, measured by the library's
# autotuner (multigrid-fvm_amd/tuning/make_table.py --workload batch256)
TUNING_TABLE = os.path.join("multigrid-fvm_amd", "tuning", "batch256_mi355x.txt")


def tuning_table():
    path = os.path.join(ROOT, TUNING_TABLE)
    return (open(path).read(), TUNING_TABLE) if os.path.exists(path) else (None, None)


def profiled_traffic():
    """(bytes per launch, source) from the counter summary's '**total X MB**'"""
    path = os.path.join(ROOT, TRAFFIC_PROFILE)
    try:
        txt = open(path).read()
    except OSError:
        return None, None
    m = re.search(r"\*\*total ([0-9.]+) MB\*\*", txt)
    return (float(m.group(1)) * 1e6, TRAFFIC_PROFILE) if m else (None, None)
WORKLOAD = ("256 independent channel problems solved as one batch: n in {128,256,512} "
            "(coarsest level 16x16), staggered obstacle array (spacing 4 on a 16x16 base, "
            "each obstacle moved by a seeded quarter/half/full cell), Re in "
            "{10,50,100,200,400}, red-black 2/6, FP32 correction sweeps with FP64 "
            "residuals and state, stop at FP64 sum|div| < 1e-6 or 2999 V-cycles")


def _sync(obj):
    import torch
    obj.synchronize()
    torch.cuda.synchronize()


def cpu_info():
    """(model name, logical CPUs) of the host"""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def time_reference_cycle(prob, workdir, procs=1, target_s=4.0):
    """Seconds per V-cycle of the reference program (oracle/_ref/ref_plain,
    built from the reference's own source) -- or of the CPU restatement
    (oracle/orc_run) where that was not built -- measured on a bounded run
    of the problem (V-cycle loop time only), with `procs` copies running at
    once, one per core (the aggregate of independent single-threaded solves
    on the host).  Returns (seconds per V-cycle of one copy, kind, cycles)."""
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_plain")
    port = os.path.join(ROOT, "oracle", "orc_run")
    kind = "reference" if os.path.exists(ref) else "port"
    exe = ref if kind == "reference" else port
    if not os.path.exists(exe):
        return None, None, 0
    taskset = shutil.which("taskset")

    def run(ncyc):
        ps = []
        for c in range(procs):
            d = os.path.join(workdir, f"cpu_{prob.nfx}_{ncyc}_{c}")
            shutil.rmtree(d, ignore_errors=True)
            prob.replace(nvcycle=ncyc, convergence=0.0).write(d)
            pre = [taskset, "-c", str(c)] if taskset else []
            cmd = pre + ([exe] if kind == "reference"
                         else [exe, "InputData", "pattern.txt", os.path.join(d, "o.bin")])
            ps.append(subprocess.Popen(cmd, cwd=d, stdout=subprocess.PIPE,
                                       stderr=subprocess.PIPE, text=True))
        ts = []
        for pr in ps:
            out, err = pr.communicate()
            m = (re.search(r"simple solution took ([0-9.eE+-]+) seconds", out) if kind == "reference"
                 else re.search(r"seconds ([0-9.eE+-]+)", err))
            if m:
                ts.append(float(m.group(1)))
        return sum(ts) / len(ts) if ts else None

    probe = 2
    t = run(probe)
    if t is None:
        return None, None, 0
    per = t / probe
    n = max(probe, int(target_s / max(per, 1e-6)))
    t = run(n) if n != probe else t
    return (t / n if t is not None else per), kind, n


def mixed_field_error(bm, bf, ks):
    """max over the problems of max |mixed - FP64| / max |FP64| per field on
    level 0; p on the cells of the flow's connected region (a pocket closed
    off by obstacles has an undetermined pressure)"""
    import numpy as np
    import scipy.sparse as sp
    import scipy.sparse.csgraph as csg
    err = {"u": 0.0, "v": 0.0, "sc": 0.0, "p": 0.0}
    for b in ks:
        for f in ("u", "v", "sc"):
            r = bf.get(b, 0, f)
            err[f] = max(err[f], float(np.abs(bm.get(b, 0, f) - r).max() / np.abs(r).max()))
        apu, apv = bf.get(b, 0, "apu"), bf.get(b, 0, "apv")
        ny, nx = apu.shape
        fluid = np.zeros((ny, nx), bool)
        fluid[1:-1, 1:-1] = True
        open_e = apu < 1e29
        open_n = apv < 1e29
        # connect cells through open faces: label a graph on cells
        idx = np.arange(nx * ny).reshape(ny, nx)
        a = np.concatenate([idx[:, :-1][open_e[:, :-1] & fluid[:, :-1] & fluid[:, 1:]].ravel(),
                            idx[:-1, :][open_n[:-1, :] & fluid[:-1, :] & fluid[1:, :]].ravel()])
        c = np.concatenate([idx[:, 1:][open_e[:, :-1] & fluid[:, :-1] & fluid[:, 1:]].ravel(),
                            idx[1:, :][open_n[:-1, :] & fluid[:-1, :] & fluid[1:, :]].ravel()])
        g = sp.coo_matrix((np.ones(len(a)), (a, c)), shape=(nx * ny, nx * ny))
        _, lab = csg.connected_components(g, directed=False)
        main = lab.reshape(ny, nx) == lab[idx[ny // 2, 1]]
        pr, pm = bf.get(b, 0, "p"), bm.get(b, 0, "p")
        scale = np.abs(pr[main]).max()
        if scale > 0:
            err["p"] = max(err["p"], float(np.abs(pm[main] - pr[main]).max() / scale))
    return {k: float(f"{v:.3e}") for k, v in err.items()}


def kernel_roofline(probs, pairs, reps, cycles=None):
    """The dominant kernel of a batched mixed V-cycle is the launch that
    relaxes u, v and sc together (k_rb32x_uvs<2,512,32>: for each field, FP64
    assembly of its equation on the tile and halo, two FP32 red-black sweeps,
    FP32 correction out).  Its mean duration is timed with HIP events on the
    batch's own stream over every member of the largest grid size at level 0,
    after that batch has run (graph-replayed V-cycles cannot be bracketed per
    kernel, so the launch is repeated on its own).  Also reported: the FP32
    sweep kernel (k_rb32x<3,1024,0,-1>, three p' sweeps on the stored
    equation)."""
    import mgfvm
    nmax = max(p.nfx for p in probs)
    big = [k for k in range(len(probs)) if probs[k].nfx == nmax]
    big_pairs = [pairs[k] for k in big]
    if cycles is not None:
        work = os.path.dirname(os.path.dirname(pairs[0][0]))
        big_pairs = [probs[k].replace(nvcycle=cycles, convergence=0.0).write(
            os.path.join(work, f"roof{k}")) for k in big]
    bb = mgfvm.Batch(big_pairs, precision=mgfvm.PREC_MIXED, device=0)
    bb.reset()
    bb.solve()          # leaves an assembled state behind
    us, nbytes = bb.time_kernel(mgfvm.KERNEL_RELAX_UVS, 0, reps)
    us_s, nbytes_s = bb.time_kernel(mgfvm.KERNEL_SWEEP_PP, 0, reps)
    bb.close()
    achieved = nbytes / (us * 1e-6) / 1e9
    ncell = nmax * nmax * len(big_pairs)
    traffic, source = profiled_traffic()
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_source": source,
            "kernel": f"k_rb32x_uvs<2,512,32>: u, v and sc relaxed in one launch (FP64 assembly + "
                      f"2 FP32 red-black sweeps each) over the {len(big_pairs)} {nmax}^2 "
                      f"members, level 0",
            "mean_us": round(us, 2), "bytes_per_launch": nbytes,
            "bytes_per_cell": round(nbytes / ncell, 2),
            "sweep_kernel": {"kernel": "k_rb32x<3,1024,0,-1>: 3 FP32 red-black p' sweeps on "
                                       "a stored equation (split assembly, or p' beyond six "
                                       "sweeps), same members",
                             "mean_us": round(us_s, 2), "bytes_per_launch": nbytes_s,
                             "compulsory_bytes_per_launch": 32 * ncell,
                             "compulsory_frac": round(32 * ncell / (us_s * 1e-6) / 1e9
                                                      / HBM_PEAK_GBS, 4),
                             "compulsory_note": "the equation (6 FP32 words) and d read once "
                                                "and d written once per launch, against the "
                                                "HBM peak",
                             "per_sweep_effective_GBps": round(nbytes_s / (us_s * 1e-6) / 1e9, 1),
                             "note": "8 words per unknown and sweep over 3 temporally blocked "
                                     "sweeps whose data stay on chip: an effective rate, not "
                                     "HBM bandwidth"}}


def plain_estimate(probs, pairs, ncyc, conv, per_class):
    """Converged solves/s of the plain FP64 GPU path (one problem at a time,
    the FP64 kernels, no batching) over the whole set, estimated by
    stratified sampling: the problems are split by grid size and by whether
    the batch converged them; per_class problems of each class are solved to
    their stopping test, and each class's total time is its count times the
    class's mean solve time.  Returns (value, detail)."""
    import mgfvm
    classes = {}
    for k, p in enumerate(probs):
        classes.setdefault((p.nfx, bool(conv[k])), []).append(k)
    total_t, detail, mismatch = 0.0, [], 0
    for key in sorted(classes):
        ks = classes[key]
        sample = ks[:per_class]
        ts = []
        for k in sample:
            s = mgfvm.Solver(*pairs[k], device=0)
            s.solve()
            s.reset()
            _sync(s)
            t0 = time.perf_counter()
            _, n, c = s.solve()
            _sync(s)
            ts.append(time.perf_counter() - t0)
            mismatch += int(bool(c) != key[1])
            s.close()
        mean_t = sum(ts) / len(ts)
        total_t += mean_t * len(ks)
        detail.append(f"{key[0]}^2 {'converging' if key[1] else 'not converging'}: "
                      f"{len(ks)} problems, {len(sample)} timed, {mean_t * 1e3:.1f} ms/solve")
    nconv = sum(bool(c) for c in conv)
    return (nconv / total_t if total_t > 0 else None), {
        "method": "stratified by grid size and convergence; class time = count x mean of "
                  "the timed solves", "classes": detail,
        "sampled_convergence_mismatches": mismatch, "estimated_total_s": round(total_t, 2)}


def roofline_only(args):
    """Counter-profiling run: the dominant kernel's launches and nothing else
    of weight (a 3-cycle batch of the 512^2 members, then --reps sweeps),
    followed by one device-to-device copy of 1 GiB whose counters scale
    FETCH_SIZE / WRITE_SIZE to bytes."""
    import torch
    from mgfvm.inputs import staggered_batch
    torch.cuda.set_device(0)
    work = tempfile.mkdtemp(prefix="mgf_roof_")
    try:
        probs = staggered_batch(args.count)
        pairs = [p.write(os.path.join(work, f"p{k}")) for k, p in enumerate(probs)]
        if args.fp64_only:
            bf = mgfvm.Batch(pairs, precision=mgfvm.PREC_FP64, device=0)
            t0 = time.perf_counter()
            nf, cf = bf.solve()
            _sync(bf)
            tf = time.perf_counter() - t0
            bf.close()
            print(f"[bench] batched fp64: {sum(cf)} converged in {tf:.2f}s "
                  f"({sum(cf) / tf:.2f} solves/s)", file=sys.stderr, flush=True)
            return
        roof = kernel_roofline(probs, pairs, args.reps, cycles=3)
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        roof["copy_bytes"] = a.numel() * a.element_size()
        print(json.dumps(roof))
    finally:
        shutil.rmtree(work, ignore_errors=True)


def tail_scaling(args):
    """V-cycle time of n of the batch's 512^2 problems at Re 10/50 (the set
    whose 24 non-converging members set the step), solved together for a
    fixed number of V-cycles, for n = 1 .. 32 with 1 and 3 lanes: the split
    of the tail's V-cycle into a fixed latency chain and a per-problem part.
    Prints one JSON line."""
    import torch
    import mgfvm
    from mgfvm.inputs import staggered_batch
    torch.cuda.set_device(0)
    ncyc = 100
    work = tempfile.mkdtemp(prefix="mgf_tail_")
    try:
        probs = [p.replace(nvcycle=ncyc, convergence=0.0) for p in staggered_batch(args.count)
                 if p.nfx == 512 and p.amuf >= 1.0 / 50 - 1e-12]
        pairs = [p.write(os.path.join(work, f"t{k}")) for k, p in enumerate(probs)]
        rows = []
        for n in (1, 2, 4, 8, 16, 24, 32):
            if n > len(pairs):
                break
            for lanes in (1, 3):
                if lanes > n:
                    continue
                b = mgfvm.Batch(pairs[:n], precision=mgfvm.PREC_MIXED, device=0, lanes=lanes)
                b.solve()
                b.reset()
                _sync(b)
                t0 = time.perf_counter()
                b.solve()
                _sync(b)
                dt = time.perf_counter() - t0
                b.close()
                rows.append({"problems": n, "lanes": lanes,
                             "ms_per_v_cycle": round(dt / ncyc * 1e3, 4)})
                print(f"[bench] tail: {n} problems, {lanes} lanes: {dt / ncyc * 1e3:.3f} ms "
                      "per V-cycle", file=sys.stderr, flush=True)
        # least squares of ms per V-cycle against n, 3 lanes, n >= 4
        pts = [(r["problems"], r["ms_per_v_cycle"]) for r in rows
               if r["lanes"] == 3 and r["problems"] >= 4]
        fit = None
        if len(pts) >= 2:
            mx = sum(x for x, _ in pts) / len(pts)
            my = sum(y for _, y in pts) / len(pts)
            slope = (sum((x - mx) * (y - my) for x, y in pts) /
                     sum((x - mx) ** 2 for x, _ in pts))
            fit = {"fixed_ms": round(my - slope * mx, 4), "per_problem_us": round(slope * 1e3, 2)}
        print(json.dumps({"workload": "512^2 problems of the batch at Re 10 and 50, mixed "
                                      f"precision, {ncyc} V-cycles from the initial state",
                          "rows": rows, "fit_3_lanes_n_ge_4": fit}))
    finally:
        shutil.rmtree(work, ignore_errors=True)


# GMRES restart of the batched-adjoint figure: 100 Krylov vectors (plus 4
# work vectors) of 4 x 514^2 doubles per 512^2 problem, 0.88 GB each
ADJ_RESTART = 100


def adjoint_figure(count, work):
    """The batched adjoint: `count` 512^2 problems (random masks, Re in
    [20, 300]) solved in mixed precision to FP64 sum|div| < 1e-8, then the
    reverse-mode dJ1/d(inlet profile, amuf) of every converged one in one
    batched solve (GMRES restarted after ADJ_RESTART vectors); both timed
    after an untimed set-up run."""
    import mgfvm
    from mgfvm import adjoint as gadj
    from mgfvm.inputs import adjoint_batch
    aprobs = adjoint_batch(count)
    apairs = [p.write(os.path.join(work, f"a{k}")) for k, p in enumerate(aprobs)]
    ba = mgfvm.Batch(apairs, precision=mgfvm.PREC_MIXED, device=0)
    ba.solve()
    aopt = gadj.opts(restart=ADJ_RESTART)
    gadj.batch_adjoint(ba, 0, aopt, allow_unconverged=True)      # set-up, untimed
    ba.reset()
    _sync(ba)
    t0 = time.perf_counter()
    anc, acv = ba.solve()
    _sync(ba)
    t_ap = time.perf_counter() - t0
    t0 = time.perf_counter()
    res = gadj.batch_adjoint(ba, 0, aopt, allow_unconverged=True)
    t_aa = time.perf_counter() - t0
    ba.close()
    its = [r[4] for r in res if r[4] >= 0]
    ok = sum(1 for r in res if r[4] >= 0 and r[5] <= 1e-10)
    print(f"[bench] batched adjoint: primal {t_ap:.2f}s ({sum(acv)} converged), adjoint "
          f"{t_aa:.2f}s ({ok} gradients)", file=sys.stderr, flush=True)
    return {
        "workload": f"{len(aprobs)} channels of 512x512, 6 levels, Bernoulli masks on a "
                    "16x16 base with porosity U[0.7, 0.95], Re U[20, 300], seeds 4000+k; "
                    "mixed-precision primal to FP64 sum|div| < 1e-8, then dJ1/d(inlet "
                    "profile, amuf) of every converged problem in one batched reverse "
                    f"solve (FP64, GMRES({ADJ_RESTART}) to 1e-11 relative residual)",
        "gmres_restart": ADJ_RESTART,
        "krylov_workspace_gb_per_problem": round((ADJ_RESTART + 4) * 4 * 514 * 514 * 8 / 1e9, 2),
        "primal_s": round(t_ap, 3), "primal_converged": int(sum(acv)),
        "adjoint_s": round(t_aa, 3), "adjoint_converged": ok,
        "adjoint_over_primal": round(t_aa / t_ap, 3),
        "gmres_iterations_mean": round(sum(its) / max(len(its), 1), 1),
        "gmres_iterations_max": max(its) if its else None,
        "gradients_per_s": round(ok / t_aa, 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--count", type=int, default=256, help="problems in the batch")
    ap.add_argument("--plain-per-class", type=int, default=3,
                    help="problems per class timed one at a time for the plain FP64 figure")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-single", action="store_true", help="skip the 1024^2 single solve")
    ap.add_argument("--no-plain", action="store_true", help="skip the one-at-a-time figure")
    ap.add_argument("--no-gradient", action="store_true",
                    help="skip the 256^2 gradient (forward sensitivities) figure")
    ap.add_argument("--reps", type=int, default=100, help="launches timed for the roofline")
    ap.add_argument("--adjoint-count", type=int, default=64,
                    help="problems of the batched-adjoint figure (0: skip)")
    ap.add_argument("--cpu-procs", type=int, default=16,
                    help="concurrent single-threaded reference runs of the CPU baseline")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the batched V-cycles eagerly instead of replaying a hipGraph")
    ap.add_argument("--roofline-only", action="store_true",
                    help="run only the dominant kernel and a 1 GiB copy (for counter profiles)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the batched-adjoint figure (for kernel traces of that path)")
    ap.add_argument("--fp64-only", action="store_true",
                    help="solve only the batch in FP64, once (for kernel traces of that path)")
    ap.add_argument("--tail-scaling", action="store_true",
                    help="time V-cycles of 1 .. 32 of the 512^2 problems at Re 10/50 together "
                         "(the fixed and per-problem parts of the non-converging tail)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU; --gpus {args.gpus} is not supported "
              "(the problems of a batch are independent and would be split across GPUs "
              "by problem, each GPU solving its share as one batch)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import mgfvm
    from mgfvm.inputs import random_1024, staggered_batch

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.roofline_only:
        roofline_only(args)
        return
    if args.tail_scaling:
        tail_scaling(args)
        return
    if args.adjoint_only:
        print(json.dumps(adjoint_figure(args.adjoint_count, tempfile.mkdtemp(prefix="mgf_adj_"))))
        return
    torch.cuda.set_device(0)
    work = tempfile.mkdtemp(prefix="mgf_bench_")
    try:
        probs = staggered_batch(args.count)
        pairs = [p.write(os.path.join(work, f"p{k}")) for k, p in enumerate(probs)]
        if args.fp64_only:
            bf = mgfvm.Batch(pairs, precision=mgfvm.PREC_FP64, device=0)
            t0 = time.perf_counter()
            nf, cf = bf.solve()
            _sync(bf)
            tf = time.perf_counter() - t0
            bf.close()
            print(f"[bench] batched fp64: {sum(cf)} converged in {tf:.2f}s "
                  f"({sum(cf) / tf:.2f} solves/s)", file=sys.stderr, flush=True)
            return

        # ---- batched mixed precision (the measured line) ----------------
        table, table_src = tuning_table()
        bt = mgfvm.Batch(pairs, precision=mgfvm.PREC_MIXED, device=0,
                         use_graph=not args.no_graph, tuning=table)
        for _ in range(args.warmup):
            bt.reset()
            bt.solve()
        _sync(bt)
        t0 = time.perf_counter()
        conv_total = 0
        for _ in range(args.steps):
            bt.reset()
            ncyc, conv = bt.solve()
            conv_total += sum(conv)
        _sync(bt)
        elapsed = time.perf_counter() - t0
        value = conv_total / elapsed
        print(f"[bench] batched mixed: {conv_total} converged in {elapsed:.2f}s", file=sys.stderr,
              flush=True)

        # ---- the converging problems alone: what the non-converging ones cost
        conv_only = None
        if sum(conv) < len(probs):
            keep = [pairs[k] for k in range(len(probs)) if conv[k]]
            bc = mgfvm.Batch(keep, precision=mgfvm.PREC_MIXED, device=0, tuning=table)
            bc.solve()
            bc.reset()
            _sync(bc)
            t0 = time.perf_counter()
            _, cc = bc.solve()
            _sync(bc)
            tc = time.perf_counter() - t0
            bc.close()
            conv_only = {"problems": len(keep), "seconds": round(tc, 3),
                         "solves_per_s": round(sum(cc) / tc, 3),
                         "non_converging_share_of_step": round(1.0 - tc / (elapsed / args.steps), 3)}

        # ---- roofline of the dominant kernel ------------------------------
        roof = kernel_roofline(probs, pairs, args.reps)

        # ---- the same batch in FP64 (bit-exact with the reference) --------
        batched_fp64 = None
        if not args.no_plain:
            bf = mgfvm.Batch(pairs, precision=mgfvm.PREC_FP64, device=0)
            t0 = time.perf_counter()
            nf, cf = bf.solve()
            _sync(bf)
            tf = time.perf_counter() - t0
            batched_fp64 = {"solves_per_s": round(sum(cf) / tf, 3), "seconds": round(tf, 2),
                            "converged": int(sum(cf)),
                            "cycle_counts_equal_mixed": bool(list(nf) == list(ncyc))}
            print(f"[bench] batched fp64: {sum(cf)} converged in {tf:.2f}s", file=sys.stderr,
                  flush=True)
            # the mixed batch's final state against the FP64 batch's (bit-exact
            # with the reference), per field
            both = [k for k in range(len(probs)) if conv[k] and cf[k]]
            rest = [k for k in range(len(probs)) if not (conv[k] and cf[k])]
            batched_fp64["mixed_field_error"] = mixed_field_error(bt, bf, both)
            batched_fp64["mixed_field_error_definition"] = (
                f"max over the {len(both)} converged problems of max|mixed - FP64| / max|FP64| "
                "on level 0; p on the cells of the flow's connected region")
            if rest:
                batched_fp64["non_converging_field_difference"] = mixed_field_error(bt, bf, rest)
                batched_fp64["non_converging_note"] = (
                    f"the {len(rest)} problems the reference's method does not converge run "
                    "2999 cycles in both precisions; their states after that are not a solution "
                    "and differ")
            bf.close()

        bt.close()

        # ---- plain FP64 GPU path: one problem at a time -----------------
        plain_value, plain_detail = None, None
        if not args.no_plain:
            plain_value, plain_detail = plain_estimate(probs, pairs, ncyc, conv,
                                                       args.plain_per_class)
            print(f"[bench] plain fp64: {plain_value:.3f} solves/s (stratified)",
                  file=sys.stderr, flush=True)

        # ---- single large problem ----------------------------------------
        single = None
        if not args.no_single:
            p1 = random_1024(ngrid=5)
            pin, ppat = p1.write(os.path.join(work, "r1024"))
            single = {"workload": "1024x1024 channel, 64x64 Bernoulli(porosity 0.8, seed 1) "
                                  "mask, Re 200, red-black 2/6, 5 levels"}
            for prec, tag in ((mgfvm.PREC_FP64, "fp64"), (mgfvm.PREC_MIXED, "mixed")):
                s = mgfvm.Solver(pin, ppat, precision=prec, device=0)
                s.solve()
                s.reset()
                _sync(s)
                t0 = time.perf_counter()
                _, n, c = s.solve()
                _sync(s)
                dt = time.perf_counter() - t0
                single[f"{tag}_solves_per_s"] = round(1.0 / dt, 4) if c else 0.0
                single[f"{tag}_v_cycles"] = n
                if tag == "fp64":
                    us1, b1 = s.time_kernel(mgfvm.KERNEL_SWEEP_PP, 0, 100)
                    single["fp64_pp_sweep_effective_GBps"] = round(b1 / (us1 * 1e-6) / 1e9, 1)
                    single["fp64_pp_sweep_note"] = (
                        "8 words per unknown and sweep, three temporally blocked sweeps per "
                        "launch: an effective per-sweep rate, not HBM bandwidth")
                s.close()

        # ---- the lexicographic Gauss-Seidel variant ------------------------
        lexv = None
        if not args.no_single:
            from mgfvm.inputs import channel_128_inline
            lexv = {"what": "FP64 solves with isolve 2 as the reference's red-black (the parity "
                            "path) and as the labelled lexicographic variant (anti-diagonal "
                            "wavefronts, one workgroup per problem)"}
            cases = {"128^2 inline array, Re 100, 2/4": channel_128_inline().replace(
                         isolve=2, nvcycle=2999),
                     "256^2 staggered array (batch problem 1001), 2/6": next(
                         p for p in probs if p.nfx == 256 and p.ig_obs == p.ngrid)}
            for name, pq in cases.items():
                lin, lpat = pq.write(os.path.join(work, "lex" + str(len(lexv))))
                row = {}
                for tag, lex in (("red_black", False), ("lexicographic", True)):
                    s = mgfvm.Solver(lin, lpat, device=0, gs_lex=lex)
                    s.solve()
                    s.reset()
                    _sync(s)
                    t0 = time.perf_counter()
                    _, n, c = s.solve()
                    _sync(s)
                    dt = time.perf_counter() - t0
                    s.close()
                    row[tag] = {"seconds": round(dt, 4), "v_cycles": n, "converged": bool(c)}
                lexv[name] = row
            print(f"[bench] lexicographic variant: {lexv}", file=sys.stderr, flush=True)

        # ---- gradient of the 256^2 differentiation workload -----------------
        gradient = None
        if not args.no_gradient:
            import numpy as np
            from mgfvm import adjoint as gadj
            from mgfvm.inputs import sensitivity_256
            from mgfvm.tangent import gradients
            pg, inlet = sensitivity_256(1)
            gin, gpat = pg.write(os.path.join(work, "grad256"))
            s = mgfvm.Solver(gin, gpat, inlet_u=np.array(inlet), device=0)
            s.solve()
            gadj.adjoint(s, 0)                     # workspace set-up, outside the timing
            s.reset()
            _sync(s)
            t0 = time.perf_counter()
            _, ncg, cg = s.solve()
            _sync(s)
            t_primal = time.perf_counter() - t0
            rev = {}
            for obj in (0, 1):
                t0 = time.perf_counter()
                Ja, gi, ga, gp, it, rr = gadj.batch_adjoint(s, obj)[0]
                rev[obj] = (time.perf_counter() - t0, it, rr, Ja, ga, gp)
            s.close()
            t0 = time.perf_counter()
            J, din, dmu, dpr, cgrad = gradients(gin, gpat, inlet_u=np.array(inlet), device=0)
            t_grad = time.perf_counter() - t0
            nd = din.shape[1] + 2
            gradient = {"workload": "256x256 channel, 5 levels, Bernoulli(0.85) mask on a 32x32 "
                                    "base, Re 100, noisy parabolic inlet, FP64 to sum|div| < 1e-12",
                        "directions": nd, "primal_s": round(t_primal, 3),
                        "primal_v_cycles": ncg,
                        "reverse_mode": {
                            "per_objective_s": [round(rev[o][0], 4) for o in (0, 1)],
                            "per_objective_over_primal": [round(rev[o][0] / t_primal, 3)
                                                          for o in (0, 1)],
                            "gmres_iterations": [rev[o][1] for o in (0, 1)],
                            "relative_residual": [float(f"{rev[o][2]:.2e}") for o in (0, 1)],
                            "dJ_damuf": [rev[o][4] for o in (0, 1)],
                            "dJ_dprsc": [rev[o][5] for o in (0, 1)]},
                        "forward_mode_s": round(t_grad, 3),
                        "central_fd_equivalent_s": round(2 * nd * t_primal, 1),
                        "converged": bool(cg and cgrad),
                        "J": [float(x) for x in J], "forward_dJ_damuf": [float(x) for x in dmu],
                        "forward_dJ_dprsc": [float(x) for x in dpr]}
            print(f"[bench] gradient 256^2: reverse {rev[0][0]:.3f}s + {rev[1][0]:.3f}s, "
                  f"forward {t_grad:.2f}s for {nd} directions, primal {t_primal:.3f}s",
                  file=sys.stderr, flush=True)

        # ---- batched adjoint: 64 x 512^2 ------------------------------------
        batched_adjoint = None
        if not args.no_gradient and args.adjoint_count > 0:
            batched_adjoint = adjoint_figure(args.adjoint_count, work)

        # ---- CPU baseline ---------------------------------------------------
        cpu = None
        if not args.no_cpu:
            model, ncpu = cpu_info()
            # independent single-threaded solves, one per core, at most the
            # 16 cores a one-GPU host share has
            procs = max(1, min(ncpu or 1, args.cpu_procs))
            per_cycle, kinds, samples = {}, set(), []
            for n in sorted({p.nfx for p in probs}):
                # a problem of that size whose obstacles sit on the coarsest
                # level (so the reference program itself can run it)
                cand = [p for p in probs if p.nfx == n and p.ig_obs == p.ngrid] or \
                       [p for p in probs if p.nfx == n]
                t, kind, ncyc_s = time_reference_cycle(cand[0], work, procs=procs)
                if t is None:
                    continue
                per_cycle[n] = t
                kinds.add(kind)
                samples.append(f"{n}^2: {ncyc_s} V-cycles at {t * 1e3:.1f} ms/V-cycle")
            if len(per_cycle) == len({p.nfx for p in probs}):
                total = sum(per_cycle[p.nfx] * c for p, c in zip(probs, ncyc))
                total_1 = None
                if procs > 1:
                    one = {}
                    for n in sorted(per_cycle):
                        cand = [p for p in probs if p.nfx == n and p.ig_obs == p.ngrid] or \
                               [p for p in probs if p.nfx == n]
                        one[n] = time_reference_cycle(cand[0], work, procs=1, target_s=2.0)[0]
                    if all(one.values()):
                        total_1 = sum(one[p.nfx] * c for p, c in zip(probs, ncyc))
                # procs independent solves at a time: the batch's CPU time
                # divided over the cores
                cpu = {"value": sum(conv) * procs / total, "unit": "solves/s", "cores": procs,
                       "kind": "reference" if kinds == {"reference"} else "port",
                       "cpu_model": model, "host_logical_cpus": ncpu,
                       "single_core_value": sum(conv) / total_1 if total_1 else None,
                       "sample": f"{procs} single-threaded copies at once, one per core; "
                                 "per-V-cycle time of each grid size under that load from a "
                                 "bounded run of one problem of that size ("
                                 + "; ".join(samples) + "), times the V-cycles each problem of "
                                 f"the batch needs, the problems spread over the {procs} cores; "
                                 f"{sum(conv)} of {len(probs)} problems converge"}

        line = {
            "metric": METRIC, "value": round(value, 3), "unit": "solves/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(1e3 * elapsed / args.steps, 2), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
            "config": {"workload": WORKLOAD, "problems": len(probs),
                       "launch_tuning": table_src,
                       "converged_per_step": sum(conv),
                       "v_cycles_mean": round(sum(ncyc) / len(ncyc), 1),
                       "v_cycles_max": max(ncyc),
                       "plain_fp64_gpu_solves_per_s":
                           round(plain_value, 3) if plain_value else None,
                       "plain_fp64_gpu_estimate": plain_detail,
                       "batched_mixed_vs_plain":
                           round(value / plain_value, 2) if plain_value else None,
                       "batched_fp64": batched_fp64,
                       "single_1024": single,
                       "gradient_256": gradient,
                       "batched_adjoint_512": batched_adjoint,
                       "lexicographic_variant": lexv,
                       "converging_only": conv_only},
            "roofline": roof, "cpu_baseline": cpu,
        }
        print(json.dumps(line))
    finally:
        shutil.rmtree(work, ignore_errors=True)


if __name__ == "__main__":
    main()

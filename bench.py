#!/usr/bin/env python3
"""Throughput benchmark of the RheoCell hot path on one MI355X.

Default workload: the wet (viscous) model, periodic 512x512 box, 1600 cells
(jittered 40x40 lattice, R=8, 41x41 tanh patches), FP64, matrix-free
Jacobi-PCG to 1e-8 relative residual warm-started by the reference rule.
A "step" is one full Heun step (2 force evaluations + 2 field updates +
4 CG solves); value = cells * steps / wall seconds (cell-steps/s) with all
inputs resident in HBM.

Prints ONE JSON line.  Also measures, outside the timed region:
  * roofline: per-kernel device time with HIP events on the library stream;
    achieved bandwidth of the dominant kernel from its algorithmic bytes, and
    of the field-update kernels (field_update);
  * cpu_baseline: the C oracle (reference-faithful port, -O3 -ffast-math,
    1 core) on a bounded sample of the same workload: one untimed step (it
    holds the cold solve of a wet run), then at least 3 timed steps, in a
    child process pinned to one core that runs beside the GPU measurement;
  * cg_warm_start: CG time per step with the previous-solution start and the
    learned start over 500 steps each (the 512^2 / 1600-cell comparison), with
    the CG in its default site form and in the row form.

The wet solve is the reference's Jacobi-PCG (Eigen's ConjugateGradient) in
site form: its iterates carried as scalars times the first residual's
per-site zero-sum part plus box-site fields, one pass over the box sites per
iteration (rc_set_cg_form; include/rheocell.h).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0

# algorithmic bytes per launch, (bytes per patch-site row, bytes per box site)
BYTES = {
    "box_gather": (8, 12),          # read phi; write sum phi^2 and cover count
    "stage_fused_pred": (24, 8),    # read phi_n; write phi*, dphi_old; box window
    "stage_fused_corr": (32, 8),    # read phi*, phi_n, dphi_old; write phi_n+1
    "stage_cf": (24, 8),            # read phi; write f (2 doubles)
    "stage_fs_pred": (40, 8),       # read phi, v; write phi*, dphi_old
    "stage_fs_corr": (48, 8),       # read phi*, v, phi_n, dphi_old; write phi_n+1
    "cg_diag": (0, 12),
    "cg_gather_x": (16, 16),        # read x; write box sums
    "cg_gather_p_first": (32, 40),  # read r; write p; read diag, write box sums of p, p^2
    "cg_gather_p": (48, 40),        # read r, p; write p; read diag, write box sums of p, p^2
    "cg_row_resid": (48, 24),       # read x, b; write r; read N4 and diag
    "cg_box_n4": (0, 32),           # read box sums; write N4
    "cg_box_pap": (0, 56),          # read box sums, P2, diag; write N4
    "cg_row_update": (80, 24),      # read x, p, r; write x, r; read N4 and diag
    "cg_reduce": (0, 0),
    "nem_sums": (8, 16),            # (a separate sums pass; the nematic gather now writes them)
    "nem_force": (24, 24),          # read phi; write f (2 doubles); read sum phi^2, sum phi Q00, sum phi Q01
    # site form of the CG (both components, double2 fields)
    "scg_gather": (32, 32),         # read b, x0 rows; write their box sums
    "scg_init": (0, 96),            # read cover, sum b, sum x0; write D, m|n, c, y, s
    "scg_init_w": (0, 28),          # read cover, D, y (the first partial sums)
    "scg_row_init": (32, 24),       # read b, x0 rows; read c, D
    # read m|n, y, q, write y, q (68 B); odd iterations also read and write s
    # (100 B): 84 B per site averaged over an even and an odd iteration
    "scg_iter": (0, 84),
    "scg_final": (48, 40),          # read b, x0 rows, write x; read c, s, D
}
CG_SWEEP_KINDS = ("cg_box_pap", "cg_row_update", "scg_iter")

# HBM traffic per launch from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (KiB per
# dispatch, mean over the launches that did work), scaled by what the same
# counters read for a plain 256 MiB device copy (bench.py --probe-copy 256):
# FETCH_SIZE 131085 KiB and WRITE_SIZE 262144 KiB for 262144 KiB moved each
# way, i.e. reads x 2.0, writes x 1.0.  Recorded in profiles/.
READ_SCALE, WRITE_SCALE = 262144.0 / 131085.0, 1.0
TRAFFIC_KIB = {
    ("wet512_1600", "scg_iter"): (6040.0, 10252.2),
    ("wet512_1600", "cg_row_update"): (68309.3, 84801.2),
    ("wet512_1600", "cg_gather_p"): (44435.5, 52247.6),
    ("dry512_1024", "stage_fused_pred"): (8607.8, 27347.2),
    ("dry512_1024", "stage_fused_corr"): (22061.8, 13675.4),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20, help="untimed steps (also captures the dry-step graph)")
    ap.add_argument("--workload", default="wet512_1600")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--timing-steps", type=int, default=2)
    ap.add_argument("--cpu-steps", type=int, default=3, help="timed CPU-baseline steps (after one untimed step)")
    ap.add_argument("--cpu-baseline-only", action="store_true",
                    help="only run the CPU baseline of the workload and print its JSON object (the child process)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-timing", action="store_true")
    ap.add_argument("--weights", default=os.path.join(ROOT, "rheocell_amd", "weights", "warmstart_wet256.npz"),
                    help="learned warm-start weights (wet workloads; skipped when absent)")
    ap.add_argument("--cg-steps", type=int, default=500, help="steps of the learned-vs-previous CG comparison")
    ap.add_argument("--no-graphs", action="store_true", help="launch dry steps one by one instead of HIP graph replays")
    ap.add_argument("--cg-form", choices=("sites", "rows"), default="sites",
                    help="form of the wet solve's CG: box-site fields (default) or every row (rc_set_cg_form)")
    ap.add_argument("--stage-shape", default="", metavar="TxR",
                    help="force the field-update kernel shape (threads x strip rows) instead of the autotuner's choice")
    ap.add_argument("--probe-copy", type=int, default=0, metavar="MIB",
                    help="only fill and copy a MIB-sized device buffer 3 times (counter calibration), then exit")
    ap.add_argument("--box-field", type=int, default=0, metavar="SIZE_GRID",
                    help="only time rc_box_field (VelocityField box fields with a SIZE_GRID coarse average) "
                         "on the workload after --warmup steps, --steps calls, then exit")
    ap.add_argument("--probe-bw", action="store_true",
                    help="only time torch's own streaming kernels (copy, sum, add) at 1 GiB and at the wet "
                         "row-vector size, the practical HBM ceiling of this part, then exit")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(conf, keys, steps, budget_s=12.0):
    """One untimed Heun step (a wet run's cold solve), then at least `steps`
    timed steps, more while they fit a ~budget_s window: a bounded sample of
    the same workload on one core."""
    from oracle import OracleSim  # the checker, timed here as the CPU port
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except Exception:
        pass
    wall = None
    if str(keys.get("type", "none")) != "none":   # a wall box (the library takes the same field)
        from rheocell_amd.walls import wall_field
        wall = wall_field(str(keys["type"]), conf.Lx, float(keys["lambda_wall"]))
    s = OracleSim(conf, keys, flavour="fast", walls=wall)
    s.step(1)
    s0 = s.cg_stats()
    t0 = time.perf_counter()
    s.step(1)
    t1 = time.perf_counter() - t0
    more = max(steps - 1, min(int(budget_s / max(t1, 1e-6)) - 1, 2000))
    if more > 0:
        s.step(more)
    steps = 1 + more
    dt = time.perf_counter() - t0
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    st = s.cg_stats()
    solves = st.solves - s0.solves
    out = {
        "value": len(conf.cells) * steps / dt,
        "unit": "cell-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": "%s: %d Heun steps timed after 1 untimed step, oracle/liborc_fast.so (-O3 -ffast-math), "
                  "1 thread on %s; %.1f s" % (keys.get("_name", "workload"), steps, cpu, dt),
    }
    if solves:
        out["cg_iters_per_solve"] = round((st.iters_total - s0.iters_total) / solves, 2)
        out["sample"] += "; %.1f CG iterations per solve over %d solves" % (out["cg_iters_per_solve"], solves)
    return out


def ensemble_boxes(workload, seed):
    """The replica ensembles: 32 dry replicas (ensemble32) or 8 wet ones
    (wet_ensemble8), uneven boxes and patches, one handle; None otherwise."""
    from rheocell_amd import confgen
    if workload == "ensemble32":
        return confgen.replica_ensemble(32, seed=seed)
    if workload == "wet_ensemble8":
        return confgen.wet_replica_ensemble(8, seed=seed)
    return None


def cpu_baseline_child(args):
    """The CPU baseline in a child process (started before this process
    touches the GPU), so that it runs beside the GPU measurement."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-baseline-only", "--workload", args.workload,
           "--seed", str(args.seed), "--cpu-steps", str(args.cpu_steps)]
    return subprocess.Popen(cmd, stdout=subprocess.PIPE, text=True)


def cg_compare(conf, keys, weights, warmup, steps, form):
    """CG wall time per step with the previous-solution warm start and with the
    learned warm start, from the same initial state (same random stream), with
    the CG iterating in the given form (_lib.RC_CG_SITES or RC_CG_ROWS; the
    learned refinement stages run on rows in both)."""
    from rheocell_amd import _lib, learned
    from rheocell_amd.sim import Simulation
    w, b, scale = learned.load(weights)
    stages = learned.trained_stages(weights)
    out = {}
    for mode in ("previous", "learned"):
        sim = Simulation([(conf, keys)], cg_form=form)
        g = sim.gaussians(warmup + steps)
        if mode == "learned":
            learned.attach(sim, w, b, scale, stages=stages)
        else:
            sim.set_warm_start(_lib.RC_WARM_PREVIOUS)
        sim.step(warmup, g[:warmup])
        s0 = sim.cg_stats()
        t0 = time.perf_counter()
        sim.step(steps, g[warmup:])
        dt = time.perf_counter() - t0
        s1 = sim.cg_stats()
        out[mode] = {"cg_ms_per_step": round((s1.total_solve_ms - s0.total_solve_ms) / steps, 4),
                     "ms_per_step": round(1e3 * dt / steps, 4),
                     "cg_iters_per_solve": round((s1.iters_total - s0.iters_total) / max(s1.solves - s0.solves, 1), 2)}
        if mode == "learned":
            out[mode]["net_step_taken"] = "%d/%d" % (s1.learned_used - s0.learned_used, s1.learned_tried - s0.learned_tried)
        sim.close()
    out["cg_time_ratio_previous_over_learned"] = round(out["previous"]["cg_ms_per_step"] /
                                                       max(out["learned"]["cg_ms_per_step"], 1e-9), 3)
    out["weights"] = os.path.relpath(weights, ROOT)
    out["stages"] = stages
    return out


def probe_copy(mib):
    """A plain fill + device-to-device copy of known size: the reference for
    scaling the FETCH_SIZE / WRITE_SIZE counters of the path's kernels."""
    import torch
    n = mib * (1 << 20) // 8
    a = torch.empty(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"probe_copy_bytes_read": n * 8, "probe_copy_bytes_written": n * 8, "copies": 3}))


def probe_bw():
    """Achieved bandwidth of torch's streaming kernels (bytes moved / event-timed
    duration): the practical ceiling next to the 8 TB/s published peak.  At
    43 MB (one double2 row vector of the 512^2 / 1600-cell wet workload) the
    arrays partly stay in the 256 MB last-level cache."""
    import torch
    out = {}

    def rate(name, fn, nbytes, reps):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        t = e0.elapsed_time(e1) * 1e-3 / reps
        out[name] = {"GB/s": round(nbytes / t / 1e9, 1), "us": round(t * 1e6, 2), "bytes": nbytes}

    for label, n, reps in (("1GiB", 128 << 20, 20), ("43MB", 2 * 2689600, 200)):
        a = torch.rand(n, dtype=torch.float64, device="cuda")
        c = torch.rand(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        rate("copy_" + label, lambda: b.copy_(a), 2 * 8 * n, reps)
        rate("sum_" + label, lambda: a.sum(), 8 * n, reps)
        rate("add_" + label, lambda: torch.add(a, c, out=b), 3 * 8 * n, reps)
        del a, b, c
    print(json.dumps({"probe_bw": out, "peak_GB/s": 8000.0}))


def main():
    args = parse()
    if args.probe_copy:
        probe_copy(args.probe_copy)
        return
    if args.probe_bw:
        probe_bw()
        return
    if args.cpu_baseline_only:
        from rheocell_amd import confgen
        boxes = ensemble_boxes(args.workload, args.seed) or [confgen.workload(args.workload, seed=args.seed)]
        conf, keys = boxes[0]
        keys["_name"] = args.workload if len(boxes) == 1 else "%s replica 0 (%d cells, %d^2 box)" % (
            args.workload, len(conf.cells), conf.Lx)
        print(json.dumps(cpu_baseline(conf, keys, args.cpu_steps)), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has one GPU and this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    from rheocell_amd import confgen, _lib
    from rheocell_amd.sim import Simulation

    cpu_child = None
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline: 1 + %d step(s) of the oracle on 1 core, in a child process" % args.cpu_steps)
        cpu_child = cpu_baseline_child(args)

    log("[bench] generating %s (seed %d)" % (args.workload, args.seed))
    boxes = ensemble_boxes(args.workload, args.seed) or [confgen.workload(args.workload, seed=args.seed)]
    conf, keys = boxes[0]
    keys["_name"] = args.workload if len(boxes) == 1 else "%s replica 0 (%d cells, %d^2 box)" % (
        args.workload, len(conf.cells), conf.Lx)
    ncells = sum(len(c.cells) for c, _ in boxes)
    L = conf.Lx
    box_sites = sum(c.Lx * c.Lx for c, _ in boxes)
    wet = float(keys["friction_cell"]) != 0
    sim = Simulation(boxes, cg_form=_lib.RC_CG_ROWS if args.cg_form == "rows" else _lib.RC_CG_SITES)
    if args.no_graphs:
        sim.set_graphs(False)
    if args.stage_shape:
        t, r = (int(v) for v in args.stage_shape.lower().split("x"))
        sim.set_stage_shape(t, r)
    if args.box_field:
        # the VelocityField observable of box 0 on the device: first call after
        # the warm-up steps untimed, then --steps calls (each includes the copy
        # of the four L x L outputs to the host)
        sim.step(max(args.warmup, 1), sim.gaussians(max(args.warmup, 1)))
        sim.box_field(0, args.box_field)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            sim.box_field(0, args.box_field)
        ms = 1e3 * (time.perf_counter() - t0) / max(args.steps, 1)
        print(json.dumps({"box_field_ms_per_call": round(ms, 4), "size_grid": args.box_field, "box": L,
                          "cells": ncells, "workload": args.workload, "calls": args.steps}))
        return
    rows = sim.total_rows()
    if len(boxes) > 1:
        log("[bench] %d boxes, sides %s" % (len(boxes), sorted(set(c.Lx for c, _ in boxes))))
    gauss = sim.gaussians(args.warmup + args.steps + args.timing_steps)
    log("[bench] %d cells, %d rows, warmup %d" % (ncells, rows, args.warmup))
    if args.warmup:
        sim.step(args.warmup, gauss[:args.warmup])
    st0 = sim.cg_stats()
    log("[bench] timed %d steps" % args.steps)
    t0 = time.perf_counter()
    sim.step(args.steps, gauss[args.warmup:args.warmup + args.steps])
    t1 = time.perf_counter()
    st1 = sim.cg_stats()
    secs = t1 - t0
    value = ncells * args.steps / secs
    ms_step = 1e3 * secs / args.steps
    log("[bench] %.3f ms/step, %.4g cell-steps/s" % (ms_step, value))
    groups = sim.stage_groups()
    for g in groups:
        log("[bench] stage group: %d cells, shape %dx%d, tuner us %s" % (g["cells"], g["threads"], g["rows"],
                                                                         g["tune_us"]))
    step_roof = None
    fused = not wet and str(keys.get("interaction_type", "wetpolarmodel")) == "wetpolarmodel" and \
        str(keys.get("type", "none")) == "none"
    if fused:
        # whole fused dry step: read phi_n, write phi*, dphi; read phi*, phi_n, dphi, write phi_n+1;
        # the box sum phi^2 written and read twice (72 B per patch site + 32 B per box site)
        step_bytes = 72 * rows + 32 * box_sites
        step_roof = {"bytes_per_step": step_bytes, "achieved": round(step_bytes / (secs / args.steps) / 1e9, 1),
                     "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(step_bytes / (secs / args.steps) / 1e9 / PEAK_HBM_GBS, 4),
                     "bytes_model": "72 B/patch site x %d + 32 B/box site x %d" % (rows, box_sites)}

    roof = None
    kernels = {}
    field_update = {}
    if not args.no_timing and args.timing_steps > 0:
        sim.enable_timing(True)
        st_a = sim.cg_stats()
        sim.step(args.timing_steps, gauss[args.warmup + args.steps:])
        st_b = sim.cg_stats()
        sweeps = st_b.sweeps_total - st_a.sweeps_total
        for i, name in enumerate(_lib.T_KINDS):
            ms, n = sim.timing(i)
            if n:
                kernels[name] = {"ms_total": ms, "launches": n}
        sim.enable_timing(False)
        best = None
        for name, d in kernels.items():
            if name not in BYTES:
                continue
            if best is None or d["ms_total"] > kernels[best]["ms_total"]:
                best = name
        if best:
            d = kernels[best]
            per_row, per_site = BYTES[best]
            bytes_launch = per_row * rows + per_site * box_sites
            useful = sweeps if best in CG_SWEEP_KINDS else d["launches"]
            avg_us = 1e3 * d["ms_total"] / d["launches"]
            achieved = bytes_launch * useful / (d["ms_total"] * 1e-3) / 1e9
            tr = TRAFFIC_KIB.get((args.workload, best))
            traffic = round(1024 * (tr[0] * READ_SCALE + tr[1] * WRITE_SCALE)) if tr else None
            roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                    "traffic_source": ("recorded: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this kernel and workload "
                                       "(profiles/), x2.0 reads / x1.0 writes from a plain copy; not measured in "
                                       "this run") if tr else None,
                    "kernel": _lib.T_SYMBOL.get(best, best), "kind": best,
                    "bytes_per_launch": bytes_launch, "avg_launch_us": round(avg_us, 3),
                    "launches": d["launches"], "useful_launches": useful,
                    "bytes_model": "%d B/row x %d rows + %d B/site x %d box sites" % (per_row, rows, per_site,
                                                                                      box_sites)}
        log("[bench] kernels: " + json.dumps({k: (round(v["ms_total"], 3), v["launches"]) for k, v in kernels.items()}))
        # the field-update kernels (force evaluation and Heun update), each
        # against the HBM peak with its algorithmic bytes
        for name in ("stage_fused_pred", "stage_fused_corr", "stage_cf", "stage_fs_pred", "stage_fs_corr", "box_gather",
                     "nem_sums", "nem_force"):
            d = kernels.get(name)
            if not d:
                continue
            per_row, per_site = BYTES[name]
            if name == "box_gather" and str(keys.get("interaction_type", "wetpolarmodel")) == "wetmodel":
                per_site += 16   # the nematic gather also writes sum phi Q00, sum phi Q01
            b = per_row * rows + per_site * box_sites
            us = 1e3 * d["ms_total"] / d["launches"]
            field_update[_lib.T_SYMBOL.get(name, name) + " " + name] = {
                "avg_launch_us": round(us, 3), "bytes_per_launch": b,
                "achieved_GB/s": round(b / (us * 1e-6) / 1e9, 1), "frac": round(b / (us * 1e-6) / 1e9 / PEAK_HBM_GBS, 4)}

    cgcmp = None
    polar = str(keys.get("interaction_type", "wetpolarmodel")) == "wetpolarmodel"
    if wet and polar and len(boxes) == 1 and args.weights and os.path.exists(args.weights) and args.cg_steps > 0:
        log("[bench] CG warm starts: previous solution vs learned, %d steps each, site and row forms" % args.cg_steps)
        cgcmp = {"site_form": cg_compare(conf, keys, args.weights, args.warmup, args.cg_steps, _lib.RC_CG_SITES),
                 "row_form": cg_compare(conf, keys, args.weights, args.warmup, args.cg_steps, _lib.RC_CG_ROWS),
                 "note": "site form: the default CG (value above); row form: the same iteration over every row "
                         "(Eigen's form); the learned start's network stages run on rows in both"}
        log("[bench] " + json.dumps(cgcmp))

    cpu = None
    if cpu_child is not None:
        log("[bench] waiting for the CPU baseline")
        out_txt, _ = cpu_child.communicate(timeout=900)
        if cpu_child.returncode != 0:
            raise RuntimeError("CPU baseline child failed (%d)" % cpu_child.returncode)
        cpu = json.loads(out_txt.strip().splitlines()[-1])
        log("[bench] CPU %.4g cell-steps/s" % cpu["value"])

    solves = st1.solves - st0.solves
    iters = st1.iters_total - st0.iters_total
    out = {
        "metric": "cell-steps/s, 512^2 wet, 1600 cells" if args.workload == "wet512_1600" else "cell-steps/s, %s" % args.workload,
        "value": round(value, 2),
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (jittered-lattice tanh cells, seed %d)" % args.seed,
        "config": {"workload": args.workload, "box": L if len(boxes) == 1 else sorted(c.Lx for c, _ in boxes),
                   "boxes": len(boxes), "cells": ncells,
                   "patch": "41x41" if len(boxes) == 1 else "29..49 (uneven)", "rows": rows,
                   "model": str(keys.get("interaction_type", "wetpolarmodel")),
                   "box_type": str(keys.get("type", "none")), "friction_cell": float(keys["friction_cell"]),
                   "cg_tolerance": float(keys["CGtolerance"]) if wet else None,
                   "cg_form": args.cg_form if wet else None,
                   "warm_start": ("previous solution, cold every %s solve (reference rule)" %
                                  ("101st" if keys.get("interaction_type") == "wetmodel" else "11th")) if wet else None},
        "roofline": roof,
        "cpu_baseline": cpu,
        "graphs": (lambda g: {"replays": g[0], "steps_launched_singly": g[1]})(sim.graph_stats()),
        "growths": (lambda g: {"served": g[0], "host_ms": round(g[1], 1)})(sim.growth_stats()),
        "stage_groups": [{"cells": g["cells"], "threads": g["threads"], "rows": g["rows"], "tuner_us": g["tune_us"]}
                         for g in groups],
    }
    if step_roof:
        out["step_roofline"] = step_roof
    if wet:
        out["cg_ms_per_step"] = round((st1.total_solve_ms - st0.total_solve_ms) / args.steps, 4)
        out["cg_iters_per_solve"] = round(iters / max(solves, 1), 2)
    if cgcmp:
        out["cg_warm_start"] = cgcmp
    if field_update:
        out["field_update_roofline"] = field_update
    if kernels:
        out["kernels_ms"] = {k: round(v["ms_total"], 4) for k, v in kernels.items()}
    print(json.dumps(out), flush=True)
    sim.close()


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Trains the learned CG warm start on the GPU and saves its weights.

Recipe: wet-model runs on a 256x256 box with 400 cells (friction_cell = 1,
CG to 1e-8) record training pairs: the box sums of a start's residual and
sum phi^2 as the input, the box sums of the correction to the converged
solution as the target, both in units of the start's residual scale.

  round 0   starts are the previous solutions (RC_WARM_PREVIOUS runs);
  round k   the network of round k-1 refines each start over --stages stages
            and every stage's start is recorded, so the network also learns
            the residuals it leaves behind; it is fine-tuned on all records.

The network is trained with the library's BF16-MFMA forward / backward kernels
and Adam, then applied fully convolutionally on any box.

    python -m rheocell_amd.train_warmstart --seeds 4 --steps 500 --epochs 40 \\
        --rounds 2 --stages 4 --out rheocell_amd/weights/warmstart_wet256.npz
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from rheocell_amd import learned  # noqa: E402


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--L", type=int, default=256)
    ap.add_argument("--nside", type=int, default=20)
    ap.add_argument("--seeds", type=int, default=4)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--burn-in", type=int, default=10)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--rounds", type=int, default=2, help="refinement rounds after round 0")
    ap.add_argument("--stages", type=int, default=4, help="network stages per solve in the refinement rounds")
    ap.add_argument("--round-steps", type=int, default=60, help="steps per seed of each refinement round")
    ap.add_argument("--round-epochs", type=int, default=15)
    ap.add_argument("--out", default=os.path.join(HERE, "weights", "warmstart_wet256.npz"))
    args = ap.parse_args()
    seeds = list(range(101, 101 + args.seeds))
    info = {"box": args.L, "cells": args.nside ** 2, "batch": args.batch, "lr": args.lr, "rounds": []}
    t0 = time.time()
    runs = learned.generate_snapshots(L=args.L, nside=args.nside, seeds=seeds, steps=args.steps,
                                      burn_in=args.burn_in, log=log)
    t1 = time.time()
    w, b, scale, hist = learned.train(runs, epochs=args.epochs, batch=args.batch, lr=args.lr, log=log)
    info["rounds"].append({"records": int(sum(len(r) for r in runs)), "epochs": args.epochs,
                           "loss_first_epoch": hist[0], "loss_last_epoch": hist[-1],
                           "record_s": round(t1 - t0, 1), "train_s": round(time.time() - t1, 1)})
    log("[train] round 0: " + json.dumps(info["rounds"][-1]))
    for k in range(1, args.rounds + 1):
        t0 = time.time()
        new = learned.generate_snapshots(L=args.L, nside=args.nside, seeds=[s + 100 * k for s in seeds],
                                         steps=args.round_steps, burn_in=args.burn_in, log=log,
                                         params=(w, b, scale), stages=args.stages)
        runs = runs + new
        t1 = time.time()
        w, b, scale, hist = learned.train(runs, epochs=args.round_epochs, batch=args.batch, lr=0.5 * args.lr,
                                          log=log, params=(w, b), scale=scale, seed=k)
        info["rounds"].append({"records": int(sum(len(r) for r in new)), "stages": args.stages,
                               "epochs": args.round_epochs, "loss_first_epoch": hist[0],
                               "loss_last_epoch": hist[-1], "record_s": round(t1 - t0, 1),
                               "train_s": round(time.time() - t1, 1)})
        log("[train] round %d: " % k + json.dumps(info["rounds"][-1]))
    stages = args.stages if args.rounds > 0 else 1
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    learned.save(args.out, w, b, scale, stages=stages)
    info.update({"records_total": int(sum(len(r) for r in runs)), "stages": stages,
                 "scale": [float(v) for v in scale]})
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(info, f, indent=1)
    print(json.dumps(info))


if __name__ == "__main__":
    main()

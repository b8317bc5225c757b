#!/usr/bin/env python3
"""Trains the learned CG warm start on the GPU and saves its weights.

Recipe: wet-model runs on a 256x256 box with 400 cells (the wet256_400
workload, friction_cell = 1, CG to 1e-8) record the box fields of every
solve: the box sums of the previous solution's residual and sum phi^2 as the
input, the box sums of the correction to the converged solution as the target;
the network is trained with the library's BF16-MFMA forward / backward kernels
and Adam, then applied fully convolutionally on any box.

    python -m rheocell_amd.train_warmstart --seeds 4 --steps 500 --epochs 40 \
        --out rheocell_amd/weights/warmstart_wet256.npz
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

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
    ap.add_argument("--out", default=os.path.join(HERE, "weights", "warmstart_wet256.npz"))
    args = ap.parse_args()
    t0 = time.time()
    runs = learned.generate_snapshots(L=args.L, nside=args.nside, seeds=list(range(101, 101 + args.seeds)),
                                      steps=args.steps, burn_in=args.burn_in, log=log)
    t1 = time.time()
    nsolves = sum(len(r) for r in runs)
    log("[train] %d solves recorded in %.1f s" % (nsolves, t1 - t0))
    w, b, scale, hist = learned.train(runs, epochs=args.epochs, batch=args.batch, lr=args.lr, log=log)
    t2 = time.time()
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    learned.save(args.out, w, b, scale)
    info = {"solves": nsolves, "pairs": nsolves, "box": args.L, "cells": args.nside ** 2,
            "epochs": args.epochs, "batch": args.batch, "lr": args.lr, "loss_first_epoch": hist[0],
            "loss_last_epoch": hist[-1], "record_s": round(t1 - t0, 1), "train_s": round(t2 - t1, 1),
            "scale": [float(v) for v in scale]}
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(info, f, indent=1)
    print(json.dumps(info))


if __name__ == "__main__":
    main()

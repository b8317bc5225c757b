#!/usr/bin/env python3
"""Writes tests/golden/ref_fields.json, ref_stencils.json and ref_walls.json
from the reference's own code.

Runs oracle/_ref/ref_fields (oracle/ref_driver.cpp linked against the
reference's MultiPhaseField, SquareBox, the wall boxes, BaseInteraction and
Utils sources,
compiled from /root/reference by oracle/build_ref.sh) and, for the trig
tables only, the same driver compiled with the reference's Release flags
(ref_fields_fast).  The fixtures hold inputs and outputs only.

Usage: bash oracle/build_ref.sh && python tests/golden/make_golden.py
"""
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "ref_fields")
REF_FAST = os.path.join(ROOT, "oracle", "_ref", "ref_fields_fast")


def run(binary, cmds):
    out = subprocess.run([binary], input="\n".join(cmds) + "\n", capture_output=True, text=True, check=True).stdout
    # the wall boxes also print a line of their own ("TESTING: Walls empty area")
    return [json.loads(l) for l in out.splitlines() if l.startswith("{")]


def main():
    if not os.path.exists(REF):
        sys.exit("build oracle/_ref first: bash oracle/build_ref.sh")
    rng = np.random.default_rng(2024)
    cmds = ["trig %d" % L for L in range(29, 50)]
    cmds += ["rng %d 64" % s for s in (1, 4982, 123456789)]
    cmds += ["gauss %d 64" % s for s in (1, 4982, 123456789)]
    sub = [(64, 5, 4, 2, 1, 4090), (140, 41, 41, 0, 0, 99 * 140 + 130), (140, 41, 41, 17, 33, 139),
           (512, 41, 41, 40, 40, 511 * 512 + 511), (128, 29, 49, 3, 48, 5000)]
    cmds += ["subindex %d %d %d %d %d %d" % s for s in sub]
    cmds += ["nbrsub 5 4", "nbrsub 41 41", "nbrsub 29 33"]
    setpos_inputs = []
    # recentering: random CoM within the patch, corners near box edges
    for t in range(24):
        Lb = int(rng.choice([64, 140, 512]))
        LX = int(rng.choice([29, 41, 49]))
        LY = LX if t % 3 else int(rng.choice([29, 41]))
        ox, oy = int(rng.integers(LX)), int(rng.integers(LY))
        corner = int(rng.integers(Lb)) + Lb * int(rng.integers(Lb))
        if t % 4 == 0:
            corner = (Lb - 3) + Lb * (Lb - 5)
        ux, uy = float(rng.integers(-50, 50)), float(rng.integers(-50, 50))
        cx, cy = float(rng.uniform(0, LX)), float(rng.uniform(0, LY))
        setpos_inputs.append((Lb, LX, LY, ox, oy, corner, ux, uy, cx, cy, 9, 9, np.zeros(LX * LY)))
    # growth / reshaping branch (x_sub_left or y_sub_bottom < border = 4)
    for t in range(10):
        Lb = 64
        LX = LY = int(rng.choice([12, 16, 21]))
        ox, oy = int(rng.integers(LX)), int(rng.integers(LY))
        corner = int(rng.integers(Lb)) + Lb * int(rng.integers(Lb))
        xsl = int(rng.integers(0, 4)) if t % 2 == 0 else int(rng.integers(4, LX // 2))
        ysb = int(rng.integers(0, 4)) if t % 2 == 1 or t % 3 == 0 else int(rng.integers(4, LY // 2))
        phi = np.round(rng.uniform(0, 1, LX * LY), 12)
        setpos_inputs.append((Lb, LX, LY, ox, oy, corner, 3.0, -7.0, LX / 2 + 0.3, LY / 2 - 0.2, xsl, ysb, phi))
    for (Lb, LX, LY, ox, oy, corner, ux, uy, cx, cy, xsl, ysb, phi) in setpos_inputs:
        cmds.append("setpos %d %d %d %d %d %d %.17g %.17g %.17g %.17g %d %d " % (Lb, LX, LY, ox, oy, corner, ux, uy, cx, cy, xsl, ysb)
                    + " ".join("%.17g" % v for v in phi))
    res = run(REF, cmds)
    fast = run(REF_FAST, ["trig %d" % L for L in range(29, 50)])
    setpos = []
    k = 0
    for r in res:
        if r["cmd"] == "setpos":
            inp = setpos_inputs[k]
            k += 1
            r["input"] = {"Lbox": inp[0], "LX": inp[1], "LY": inp[2], "offset": [inp[3], inp[4]], "corner": inp[5],
                          "unrap": [inp[6], inp[7]], "CoM": [inp[8], inp[9]], "xsl": inp[10], "ysb": inp[11],
                          "phi": [float(v) for v in inp[12]]}
            setpos.append(r)
    out = {
        "source": "reference MultiPhaseField/SquareBox/Utils compiled from source (oracle/build_ref.sh)",
        "trig": [r for r in res if r["cmd"] == "trig"],
        "trig_fast": fast,
        "rng": [r for r in res if r["cmd"] == "rng"],
        "gauss": [r for r in res if r["cmd"] == "gauss"],
        "subindex": [r for r in res if r["cmd"] == "subindex"],
        "nbrsub": [r for r in res if r["cmd"] == "nbrsub"],
        "setpos": setpos,
    }
    with open(os.path.join(HERE, "ref_fields.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", os.path.join(HERE, "ref_fields.json"))
    stencils()
    walls()
    example_input()


def stencils():
    """BaseInteraction::derivX / derivY / Laplacian (BaseInteraction.h:196-263)
    of random patches in a periodic square box, ring offsets included."""
    rng = np.random.default_rng(77)
    cases = [(64, 9, 7, 0, 0, 0), (64, 9, 7, 4, 5, 63 * 64 + 60), (140, 13, 11, 12, 3, 70 * 140 + 139),
             (128, 29, 31, 17, 30, 127), (96, 21, 21, 0, 20, 95 * 96)]
    cmds, inputs = [], []
    for (Lb, LX, LY, ox, oy, corner) in cases:
        phi = np.round(rng.uniform(-0.2, 1.2, LX * LY), 12)
        inputs.append([float(v) for v in phi])
        cmds.append("stencil %d %d %d %d %d %d " % (Lb, LX, LY, ox, oy, corner) + " ".join("%.17g" % v for v in phi))
    res = run(REF, cmds)
    for r, phi in zip(res, inputs):
        r["phi"] = phi
    path = os.path.join(HERE, "ref_stencils.json")
    with open(path, "w") as f:
        json.dump({"source": "reference BaseInteraction stencils on MultiPhaseField/SquareBox (oracle/build_ref.sh)",
                   "stencil": res}, f, separators=(",", ":"))
    print("wrote", path)


def walls():
    """The wall fields of SquareWalls / CircleWalls / TriangleWalls and the
    BaseInteraction stencils of patches that reach their walls."""
    rng = np.random.default_rng(91)
    boxes = [("square_walls", 48, 4.0), ("circle_walls", 48, 3.0), ("triangle_walls", 48, 2.0),
             ("square_walls", 64, 5.0), ("circle_walls", 61, 4.0)]
    cmds = ["walls %s %d %.17g" % b for b in boxes]
    st_cases = [("square_walls", 48, 4.0, 0.5, 15, 13, 3, 2, 0), ("square_walls", 48, 4.0, 0.5, 17, 17, 0, 0, 40 * 48 + 38),
                ("circle_walls", 48, 3.0, 0.5, 21, 21, 5, 9, 1 * 48 + 2), ("triangle_walls", 48, 2.0, 0.5, 19, 15, 0, 4, 30 * 48 + 14),
                ("square_walls", 48, 4.0, 2.0, 15, 13, 3, 2, 0)]
    inputs = []
    for (t, Lb, lw, ws, LX, LY, ox, oy, corner) in st_cases:
        phi = np.round(rng.uniform(-0.2, 1.2, LX * LY), 12)
        inputs.append([float(v) for v in phi])
        cmds.append("stencilw %s %d %.17g %.17g %d %d %d %d %d " % (t, Lb, lw, ws, LX, LY, ox, oy, corner)
                    + " ".join("%.17g" % v for v in phi))
    res = run(REF, cmds)
    st = [r for r in res if r["cmd"] == "stencilw"]
    for r, phi in zip(st, inputs):
        r["phi"] = phi
    path = os.path.join(HERE, "ref_walls.json")
    with open(path, "w") as f:
        json.dump({"source": "reference SquareWalls/CircleWalls/TriangleWalls and BaseInteraction stencils "
                             "(oracle/build_ref.sh)", "walls": [r for r in res if r["cmd"] == "walls"], "stencilw": st},
                  f, separators=(",", ":"))
    print("wrote", path)


def example_input():
    """The key = value pairs of the reference's examples/input (data: the
    parameters its example run uses), parsed as the reference's reader does."""
    sys.path.insert(0, ROOT)
    from rheocell_amd.conf import read_input
    keys = read_input("/root/reference/examples/input")
    path = os.path.join(HERE, "example_input.json")
    with open(path, "w") as f:
        json.dump({"source": "key = value pairs of examples/input of the reference", "keys": keys}, f, indent=0)
    print("wrote", path)


if __name__ == "__main__":
    main()

"""Walls and external forces of a square box, for the host side.

* ``wall_field`` is the getWalls field the reference's wall boxes build in
  ``init`` (src/Boxes/SquareWalls.cpp:19-40, CircleWalls.cpp:19-37,
  TriangleWalls.cpp:19-50): 1 on wall sites, 0 elsewhere, site x + L y.
* ``read_forces`` parses an ``external_forces_file`` of ``{ key = value }``
  blocks (ForceFactory::make_forces, src/Forces/ForceFactory.cpp:61-150) into
  the force list the library takes (rheocell.h ``rc_ext_force``).

The library applies the wall masks and forces on the device; this module only
describes them.
"""
from __future__ import annotations

import math
import re
from typing import Dict, List

import numpy as np

from .conf import RCexception

WALL_BOXES = ("square_walls", "circle_walls", "triangle_walls")

# ForceFactory::add_force (ForceFactory.cpp:31-46): type -> the force class's kind
FORCE_KINDS = {"channel_walls": 0, "moving_walls": 0, "circle_walls": 0, "triangle_walls": 0, "square_walls": 0,
               "shear_flow_channel": 1, "pois_flow_channel": 2, "stretching": 3}


def wall_field(box_type: str, L: int, lambda_wall: float) -> np.ndarray:
    """getWalls of every site of an L x L wall box."""
    w = np.zeros(L * L)
    lw = float(lambda_wall)
    if box_type == "square_walls":
        for y in range(L):
            for x in range(L):
                if x < lw - 1 or x > L - lw or y < lw or y > L - lw:
                    w[x + y * L] = 1.0
    elif box_type == "circle_walls":
        h = L // 2
        for y in range(L):
            for x in range(L):
                dist = h - math.sqrt((h - x) * (h - x) + (h - y) * (h - y))
                w[x + y * L] = 1.0 if (dist <= 0 or dist < lw) else 0.0
    elif box_type == "triangle_walls":
        w[:] = 1.0
        side = int(L - lw * 2)
        height = int(math.floor(side * math.sqrt(3.0) / 2.0 + 0.5))   # C round(): half away from zero
        center = L // 2
        top = L // 2 - height // 2
        for row in range(height):
            half = int((side / 2.0) * (1.0 - row / height))
            y = top + row
            if 0 <= y < L:
                for x in range(center - half, min(center + half, L - 1) + 1):
                    if x >= 0:
                        w[x + y * L] = 0.0
    else:
        raise RCexception("box type '%s' is not a wall box of a square box (%s)" % (box_type, ", ".join(WALL_BOXES)))
    return w


def read_forces(text: str) -> List[Dict[str, object]]:
    """The ``{ key = value }`` blocks of an external_forces_file, in order
    (# comments dropped); every block must act on every cell (field = -1)."""
    text = re.sub(r"#[^\n]*", "", text)
    blocks = re.findall(r"\{([^{}]*)\}", text)
    out = []
    for b in blocks:
        kv = {}
        for line in b.splitlines():
            if "=" in line:
                k, v = line.split("=", 1)
                kv[k.strip()] = v.strip()
        t = kv.get("type")
        if t not in FORCE_KINDS:
            raise RCexception("Invalid force type `%s'" % t)
        if str(kv.get("field", "-1")).strip() != "-1":
            raise RCexception("external force `%s': only field = -1 (every cell) is supported by the HIP backend" % t)
        out.append({"type": t, "kind": FORCE_KINDS[t], "kappa_wall": float(kv.get("kappa_wall", 0.0)),
                    "lambda": float(kv.get("lambda", 2.0)), "shear_rate": float(kv.get("shear_rate", 0.0))})
    return out

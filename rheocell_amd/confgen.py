"""Initial-configuration generator for the benchmark workloads.

The reference generator only writes 30x30 step-disk patches
(src/Fields/BaseField.cpp:5, src/Managers/GeneratorManager.cpp:130-157), so
this module writes its own cells -- a tanh profile
phi = 0.5 (1 - tanh((r - R)/lambda)) on an L_sub x L_sub patch centred on a
jittered lattice point -- in the reference conf format
(src/Observables/Configuration.cpp:36-55).

Polarity: the conf stores (Q00, Q01) = (cos t, sin t) with t ~ U(-pi, pi];
the reference reader turns that into theta = 0.5 atan2(2 Q01, 2 Q00)
(src/Backends/SimBackend.cpp:283-286).
"""
from __future__ import annotations

import math
from typing import Dict, List, Optional, Sequence

import numpy as np

from .conf import CellConf, Conf


def tanh_cell(L: int, cx: float, cy: float, R: float, lam: float, side: int, theta: float,
              side_y: Optional[int] = None) -> CellConf:
    """One cell: patch of side x side_y sites whose centre is at (cx, cy)."""
    sy = side if side_y is None else side_y
    x0 = int(math.floor(cx)) - side // 2
    y0 = int(math.floor(cy)) - sy // 2
    qx = np.tile(np.arange(side), sy)
    qy = np.repeat(np.arange(sy), side)
    X = x0 + qx
    Y = y0 + qy
    r = np.sqrt((X - cx) ** 2 + (Y - cy) ** 2)
    phi = 0.5 * (1.0 - np.tanh((r - R) / lam))
    phi = np.array(["%.15g" % v for v in phi], dtype=np.float64)  # conf precision
    box = (X % L) + (Y % L) * L
    corner = (x0 % L) + (y0 % L) * L
    return CellConf(side, sy, (cx % L, cy % L), (0, 0), int(corner), (float(x0), float(y0)),
                    float("%.15g" % math.cos(theta)), float("%.15g" % math.sin(theta)),
                    box.astype(np.int64), phi)


def lattice_conf(L: int, nside: int, seed: int, R: float = 8.0, lam: float = 2.0, side: int = 41,
                 jitter: float = 2.0, radii: Optional[Sequence[float]] = None,
                 sides: Optional[Sequence[int]] = None, ncells: Optional[int] = None) -> Conf:
    """``nside`` x ``nside`` jittered lattice (spacing L/nside) in an L x L box.

    ``radii``/``sides`` give per-cell initial radii and patch sides (uneven
    patches); ``ncells`` keeps only the first cells of the lattice.
    """
    rng = np.random.default_rng(seed)
    h = L / nside
    cells: List[CellConf] = []
    n = nside * nside if ncells is None else ncells
    for i in range(n):
        a, b = i % nside, i // nside
        cx = (a + 0.5) * h + rng.uniform(-jitter, jitter)
        cy = (b + 0.5) * h + rng.uniform(-jitter, jitter)
        th = rng.uniform(-math.pi, math.pi)
        Ri = R if radii is None else radii[i]
        si = side if sides is None else sides[i]
        cells.append(tanh_cell(L, cx, cy, Ri, lam, si, th))
    return Conf(0, L, L, cells)


# Parameter sets of the benchmark workloads (reference input keys).
BASE_KEYS: Dict[str, object] = {
    "backend": "HIP", "sim_type": "FD_CPU", "interaction_type": "wetpolarmodel", "box": "square",
    "R": 8, "lambda": 2.0, "gamma": 0.008, "mu": 45, "kappa": 0.1, "omega": 0.0, "J0": 1, "dt": 0.1,
    "zetaQ_self": 0.1, "zetaQ_inter": math.sqrt(2e-3), "J_Q": 0, "anchoring": 0, "beta": 0,
    "passive_alpha": 1, "friction": 1, "friction_cell": 0, "CGtolerance": 1e-8, "wall_slip": 0.5,
    "equilibration_steps": 0, "restart_step_counter": 1, "time_scale": "linear", "T": 0.1,
    "topology": "topology.top", "conf_file": "start_conf.dat", "trajectory_file": "trajectory.dat",
}

WORKLOADS = {
    # name: (L, nside, steps, friction_cell)
    "dry140_64": (140, 8, 1000, 0),      # CPU-runnable dry case
    "dry512_1024": (512, 32, 2000, 0),   # dry, scaled up
    "wet256_400": (256, 20, 500, 1),     # wet, CG to 1e-8
    "wet512_1600": (512, 40, 500, 1),    # wet, the headline workload
    "wet1024_6400": (1024, 80, 500, 1),  # wet, solver vectors (172 MB each) larger than the 256 MB last-level cache
}


def replica_ensemble(nrep: int = 32, seed: int = 1, coverage: float = 0.9):
    """Independent dry replicas: box sides from {128,...,384}, cell radii
    U[6,10] with patch side 5R-1 rounded to odd (29..49), ~90% area
    coverage, own seed and zetaQ_self in [0.05, 0.2] per replica."""
    rng = np.random.default_rng(seed)
    out = []
    for r in range(nrep):
        L = int(rng.choice([128, 192, 256, 320, 384]))
        n_target = coverage * L * L / (math.pi * 8.0 ** 2)
        nside = max(1, int(round(math.sqrt(n_target))))
        radii = rng.uniform(6, 10, nside * nside)
        sides = [int(2 * round((5 * R - 1) / 2.0) + 1) if int(round(5 * R - 1)) % 2 == 0 else int(round(5 * R - 1))
                 for R in radii]
        sides = [min(49, max(29, s)) for s in sides]
        rs = int(rng.integers(1, 2 ** 31 - 1))
        conf = lattice_conf(L, nside, rs, radii=[float(x) for x in radii], sides=sides)
        keys = dict(BASE_KEYS)
        keys["seed"] = rs
        keys["zetaQ_self"] = float(rng.uniform(0.05, 0.2))
        keys["steps"] = 1000
        out.append((conf, keys))
    return out


def wet_replica_ensemble(nrep: int = 8, seed: int = 1, coverage: float = 0.9, tol: float = 1e-8):
    """Independent wet replicas (friction_cell = 1, CG to ``tol``): the boxes,
    uneven radii and patches, seeds and zetaQ_self of ``replica_ensemble``."""
    out = []
    for conf, keys in replica_ensemble(nrep, seed, coverage):
        keys = dict(keys)
        keys["friction_cell"] = 1.0
        keys["CGtolerance"] = tol
        keys["steps"] = 500
        out.append((conf, keys))
    return out


# The reference's own example model (examples/input: interaction_type =
# wetmodel and its parameters) on the 512^2 / 1600-cell lattice; and the dry
# polar model in a square wall box (SquareWalls, lambda_wall = 4) at 512^2 /
# 1024 cells
EXAMPLE_NEMATIC_KEYS: Dict[str, object] = {
    "interaction_type": "wetmodel", "gamma": 0.06, "mu": 20, "lambda": 2.0, "kappa": 0.5, "omega": 0,
    "wall_slip": 2, "passive_alpha": 1, "beta": 0, "friction": 1, "friction_cell": 1, "J_Q": 0.1, "anchoring": 0,
    "zetaQ_self": 0.3, "zetaQ_inter": 0, "CGtolerance": 1e-5, "dt": 0.1, "R": 8, "J0": 1,
}
EXTRA_WORKLOADS = {
    "nematic512_1600": (512, 40, 500, EXAMPLE_NEMATIC_KEYS),
    "walls512_1024": (512, 32, 2000, {"type": "square_walls", "lambda_wall": 4.0}),
}


def workload(name: str, seed: int = 1):
    if name in EXTRA_WORKLOADS:
        L, nside, steps, extra = EXTRA_WORKLOADS[name]
        conf = lattice_conf(L, nside, seed)
        keys = dict(BASE_KEYS)
        keys.update(extra)
        keys["steps"] = steps
        keys["seed"] = seed
        return conf, keys
    L, nside, steps, fc = WORKLOADS[name]
    conf = lattice_conf(L, nside, seed)
    keys = dict(BASE_KEYS)
    keys["friction_cell"] = fc
    keys["steps"] = steps
    keys["seed"] = seed
    return conf, keys

"""Long-run statistics of a trajectory: the cell-averaged area, deformation
and centre-of-mass speed, sampled every few steps after a transient, and
their time averages with a block-averaged standard error.

These are the quantities the reference's own post-processing averages over
long runs (scripts/calcAvgShape.py, scripts/calcAvgVelocity.py): per cell
the area sum phi^2, the magnitude of the shape tensor (S00, S01), and the
centre-of-mass displacement over the sampling interval (minimum image in the
periodic box) divided by the elapsed time.  Works on the library's
Simulation and the oracle's OracleSim alike.
"""
from __future__ import annotations

import math
from typing import Callable, List, Sequence, Tuple

import numpy as np

STATS = ("area", "deformation", "speed")


def _sample(cells: Sequence, prev: np.ndarray, L: int, elapsed: float) -> Tuple[np.ndarray, Tuple[float, float, float]]:
    com = np.array([[c.CoM[0], c.CoM[1]] for c in cells])
    if prev is None:
        return com, None
    d = com - prev
    d -= L * np.round(d / L)
    speed = np.hypot(d[:, 0], d[:, 1]) / elapsed
    area = np.array([c.area for c in cells])
    deformation = np.array([math.hypot(c.S00, c.S01) for c in cells])
    return com, (float(area.mean()), float(deformation.mean()), float(speed.mean()))


def series(step: Callable[[int], None], cells: Callable[[], Sequence], L: int, steps: int, every: int, skip: int,
           dt: float) -> np.ndarray:
    """Run ``steps`` steps in chunks of ``every``; one row (area, deformation,
    speed) per chunk that ends after ``skip`` steps."""
    rows: List[Tuple[float, float, float]] = []
    prev = None
    done = 0
    while done < steps:
        step(every)
        done += every
        com, row = _sample(cells(), prev, L, every * dt)
        if row is not None and done > skip:
            rows.append(row)
        prev = com
    return np.array(rows)


def sim_series(sim, L: int, steps: int, every: int, skip: int, dt: float) -> np.ndarray:
    return series(sim.step, lambda: sim.download()[0], L, steps, every, skip, dt)


def oracle_series(orc, ncells: int, L: int, steps: int, every: int, skip: int, dt: float) -> np.ndarray:
    return series(orc.step, lambda: [orc.cell(i) for i in range(ncells)], L, steps, every, skip, dt)


def mean_and_error(s: np.ndarray, blocks: int = 5) -> Tuple[np.ndarray, np.ndarray]:
    """Time average of each column and its standard error from ``blocks``
    consecutive block means (correlated samples inside a block)."""
    n = len(s) // blocks * blocks
    b = s[:n].reshape(blocks, -1, s.shape[1]).mean(axis=1)
    return s.mean(axis=0), b.std(axis=0, ddof=1) / math.sqrt(blocks)


def z_scores(a: np.ndarray, b: np.ndarray, blocks: int = 5) -> np.ndarray:
    """|mean_a - mean_b| in units of the combined standard error, per column."""
    ma, ea = mean_and_error(a, blocks)
    mb, eb = mean_and_error(b, blocks)
    return np.abs(ma - mb) / np.sqrt(ea ** 2 + eb ** 2)

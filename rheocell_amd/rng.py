"""The reference's random stream, restated for the host side.

* drand48 by its POSIX definition: x <- (0x5DEECE66D x + 0xB) mod 2^48,
  srand48(seed) sets x = seed << 16 | 0x330E (SimManager.cpp:86 seeds it).
* Utils::gaussian: Marsaglia's polar method with the cached second variate
  (src/Utilities/Utils.h:210-234).

Draw order of a reference run (one process = one stream):
  read_topology draws one drand48 per cell (WetPolarModel.cpp:72),
  read_next_configuration draws one per cell in MultiPhaseField::init
  (MultiPhaseField.cpp:79), then every predictor draws one gaussian per cell in
  cell order (WetPolarModel.cpp:365).
"""
from __future__ import annotations

import math

import numpy as np

_A = 0x5DEECE66D
_C = 0xB
_M = (1 << 48) - 1


class Drand48:
    def __init__(self, seed: int):
        self.x = ((seed & 0xFFFFFFFF) << 16) | 0x330E
        self.has_next = False
        self.next_g = -1e6

    def drand48(self) -> float:
        self.x = (_A * self.x + _C) & _M
        return self.x / 281474976710656.0

    def gaussian(self) -> float:
        if self.has_next:
            self.has_next = False
            return self.next_g
        w = 2.0
        while w >= 1.0:
            u = 2.0 * self.drand48() - 1.0
            v = 2.0 * self.drand48() - 1.0
            w = u * u + v * v
        w = math.sqrt((-2.0 * math.log(w)) / w)
        self.next_g = v * w
        self.has_next = True
        return u * w


class ReferenceStream:
    """The random stream of one box, positioned after initialisation."""

    def __init__(self, seed: int, ncells: int):
        self.rng = Drand48(seed)
        for _ in range(2 * ncells):
            self.rng.drand48()
        self.ncells = ncells

    def gaussians(self, nsteps: int) -> np.ndarray:
        g = self.rng.gaussian
        return np.array([g() for _ in range(nsteps * self.ncells)], dtype=np.float64).reshape(nsteps, self.ncells)

"""The reference's random stream, restated for the host side.

* drand48 by its POSIX definition: x <- (0x5DEECE66D x + 0xB) mod 2^48,
  srand48(seed) sets x = seed << 16 | 0x330E (SimManager.cpp:86 seeds it).
* Utils::gaussian: Marsaglia's polar method with the cached second variate
  (src/Utilities/Utils.h:210-234).

Draw order of a reference run (one process = one stream):
  read_topology draws one drand48 per cell (WetPolarModel.cpp:72; the nematic
  WetModel draws none, WetModel.cpp:67-81),
  read_next_configuration draws one per cell in MultiPhaseField::init
  (MultiPhaseField.cpp:79), then every predictor draws one gaussian per cell in
  cell order (WetPolarModel.cpp:365; the nematic model draws no noise).
"""
from __future__ import annotations

import ctypes
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


PI_F = float(np.float32(math.pi))   # src/defs.h:4 keeps PI a float
_libm = ctypes.CDLL("libm.so.6")
_libm.sincos.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
_libm.sincos.restype = None


def sincos(x: float):
    """glibc sincos: the reference build evaluates cos and sin of one argument
    with one such call."""
    s, c = ctypes.c_double(), ctypes.c_double()
    _libm.sincos(x, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def nematic_directors(seed: int, ncells: int):
    """nemQ of each cell as MultiPhaseField::init sets it (MultiPhaseField.cpp:79-95)
    from the cell's drand48 draw: theta = PI (1 - 2 u), Q = (cos 2 theta,
    sin 2 theta), nemQ = its unit director with nx >= 0.  The nematic model
    draws nothing before (WetModel.cpp:67-81), so cell i takes draw i."""
    r = Drand48(seed)
    out = []
    for _ in range(ncells):
        th = PI_F * (1 - 2 * r.drand48())
        q01, q00 = sincos(2 * th)
        mod = math.sqrt(q00 * q00 + q01 * q01)
        nx = math.sqrt((1 + q00 / mod) / 2)
        sgn = 1.0 if q01 > 0 else (-1.0 if q01 < 0 else 0.0)
        ny = sgn * math.sqrt((1 - q00 / mod) / 2)
        out.append((nx, ny))
    return out


class ReferenceStream:
    """The random stream of one box, positioned after initialisation."""

    def __init__(self, seed: int, ncells: int, nematic: bool = False):
        self.rng = Drand48(seed)
        for _ in range((1 if nematic else 2) * ncells):
            self.rng.drand48()
        self.ncells = ncells
        self.nematic = nematic

    def gaussians(self, nsteps: int) -> np.ndarray:
        if self.nematic:   # no noise: WetModel's predictor draws nothing
            return np.zeros((nsteps, self.ncells))
        g = self.rng.gaussian
        return np.array([g() for _ in range(nsteps * self.ncells)], dtype=np.float64).reshape(nsteps, self.ncells)

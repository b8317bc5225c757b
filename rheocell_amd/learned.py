"""Learned CG warm start: snapshot generation from the wet solver, GPU
training of the convolutional network, and attachment to a simulation.

Everything numeric runs in librheocell.so: the snapshots come from the
wet solve itself (rc_snapshot_box), the network's forward and backward passes
and Adam updates are the library's BF16-MFMA kernels (rc_net_*).  This module
only moves arrays and orders the calls.
"""
from __future__ import annotations

import ctypes as C
import time
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib, confgen
from .conf import RCexception
from .sim import Simulation

LAYERS, CH = 6, 32


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class ConvNet:
    """The 6-layer 3x3 network on the GPU (rc_net_*), batch N of H x W fields."""

    def __init__(self, N: int, H: int, W: int):
        self.L = _lib.load()
        self.N, self.H, self.W = N, H, W
        self.h = C.c_void_p()
        rc = self.L.rc_net_create(N, H, W, C.byref(self.h))
        if rc:
            raise RCexception("rc_net_create: %s" % (self.L.rc_net_last_error(self.h).decode() if self.h else rc))

    def _check(self, rc):
        if rc:
            raise RCexception("%s: %s" % (_lib.RC_ERR.get(rc, rc), self.L.rc_net_last_error(self.h).decode()))

    def close(self):
        if getattr(self, "h", None):
            self.L.rc_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, w: np.ndarray, b: np.ndarray):
        w = np.ascontiguousarray(w, dtype=np.float32)
        b = np.ascontiguousarray(b, dtype=np.float32)
        assert w.shape == (LAYERS, 9, CH, CH) and b.shape == (LAYERS, CH)
        self._check(self.L.rc_net_set_params(self.h, _fp(w), _fp(b)))

    def get_params(self) -> Tuple[np.ndarray, np.ndarray]:
        w = np.empty((LAYERS, 9, CH, CH), dtype=np.float32)
        b = np.empty((LAYERS, CH), dtype=np.float32)
        self._check(self.L.rc_net_get_params(self.h, _fp(w), _fp(b)))
        return w, b

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty(x.shape[:3] + (2,), dtype=np.float32)
        self._check(self.L.rc_net_forward(self.h, _fp(x), _fp(out)))
        return out

    def grads(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        gw = np.empty((LAYERS, 9, CH, CH), dtype=np.float32)
        gb = np.empty((LAYERS, CH), dtype=np.float32)
        loss = C.c_double()
        self._check(self.L.rc_net_grads(self.h, _fp(x), _fp(y), _fp(gw), _fp(gb), C.byref(loss)))
        return loss.value, gw, gb

    def train_step(self, x, y, lr: float) -> float:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        loss = C.c_double()
        self._check(self.L.rc_net_train_step(self.h, _fp(x), _fp(y), C.c_float(lr), C.byref(loss)))
        return loss.value


def init_params(seed: int = 0) -> Tuple[np.ndarray, np.ndarray]:
    """He-style initialisation; channels absent from the first/last layer are zero."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((LAYERS, 9, CH, CH)) * np.sqrt(2.0 / (9 * CH))
    w[0] = rng.standard_normal((9, CH, CH)) * np.sqrt(2.0 / (9 * 3))
    w[0, :, :, 3:] = 0
    w[LAYERS - 1, :, 2:, :] = 0
    w[LAYERS - 1] *= 0.1
    b = np.zeros((LAYERS, CH))
    return w.astype(np.float32), b.astype(np.float32)


def snapshot(sim: Simulation) -> np.ndarray:
    """(sum fx, sum fy, sum phi^2, sum vx, sum vy) box fields of the last solve, [L, L, 5]."""
    L = sim.confs[0].Lx
    out = np.empty((L * L * 5,), dtype=np.float32)
    sim._check(sim.L.rc_snapshot_box(sim.h, _fp(out)))
    return out.reshape(L, L, 5)


def record_run(L: int, nside: int, seed: int, steps: int, burn_in: int = 10, params=None,
               stages: int = 1) -> np.ndarray:
    """Training records of one wet run, [records, L, L, 5]: box sums of the
    residual of a start (x, y), sum phi^2, box sums of the correction from the
    start to the converged solution (x, y), both in units of the start's
    residual scale.  Without params every solve gives one record (start = the
    previous solution); with params = (w, b, scale) the network refines each
    start over `stages` stages and every stage's start gives a record."""
    conf = confgen.lattice_conf(L, nside, seed)
    keys = dict(confgen.BASE_KEYS, friction_cell=1, CGtolerance=1e-8, seed=seed)
    sim = Simulation([(conf, keys)])
    per_solve = 1
    if params is None:
        sim.set_warm_start(_lib.RC_WARM_PREVIOUS)
    else:
        attach(sim, *params, stages=stages)
        per_solve = stages
    if burn_in:
        sim.step(burn_in)
    cap = 2 * steps * per_solve
    sim._check(sim.L.rc_record_solves(sim.h, cap))
    sim.step(steps)
    n = C.c_int()
    buf = np.empty((cap, L, L, 5), dtype=np.float32)
    sim._check(sim.L.rc_get_recorded(sim.h, _fp(buf), C.byref(n)))
    sim.close()
    return buf[:n.value]


def generate_snapshots(L: int = 256, nside: int = 20, seeds: Sequence[int] = (11, 12), steps: int = 200,
                       burn_in: int = 10, log=None, params=None, stages: int = 1) -> List[np.ndarray]:
    """Wet-model runs (friction_cell = 1, CG to 1e-8); every solve is recorded."""
    runs = []
    for sd in seeds:
        t0 = time.time()
        runs.append(record_run(L, nside, sd, steps, burn_in, params=params, stages=stages))
        if log:
            log("[learned] seed %d: %d records in %.1f s" % (sd, len(runs[-1]), time.time() - t0))
    return runs


def pairs(runs: List[np.ndarray]) -> Tuple[np.ndarray, np.ndarray]:
    """Training pairs: (residual sums, sum phi^2) -> correction sums."""
    x = np.concatenate([r[..., 0:3] for r in runs])
    y = np.concatenate([r[..., 3:5] for r in runs])
    return x, y


def scales_of(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """[residual multiplier, sum phi^2 multiplier, output multiplier] for unit RMS."""
    return np.array([1.0 / np.sqrt(np.mean(x[..., 0:2] ** 2)), 1.0 / np.sqrt(np.mean(x[..., 2] ** 2)),
                     np.sqrt(np.mean(y ** 2))], dtype=np.float32)


def normalise(x: np.ndarray, y: np.ndarray, scale: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    xn = np.empty(x.shape, dtype=np.float32)
    xn[..., 0:2] = x[..., 0:2] * scale[0]
    xn[..., 2] = x[..., 2] * scale[1]
    return xn, (y / scale[2]).astype(np.float32)


def train(runs: List[np.ndarray], epochs: int = 20, batch: int = 8, lr: float = 2e-3, seed: int = 0, log=None,
          params: Optional[Tuple[np.ndarray, np.ndarray]] = None, scale: Optional[np.ndarray] = None):
    """Trains on the GPU (the library's forward/backward kernels and Adam);
    returns (w, b, scale, history of the mean epoch loss).  Fine-tuning passes
    the previous params and scale."""
    x, y = pairs(runs)
    if scale is None:
        scale = scales_of(x, y)
    x, y = normalise(x, y, scale)
    n, H, W = x.shape[:3]
    net = ConvNet(batch, H, W)
    w, b = init_params(seed) if params is None else params
    net.set_params(w, b)
    rng = np.random.default_rng(seed)
    hist = []
    nb = n // batch
    for ep in range(epochs):
        order = rng.permutation(n)
        tot = 0.0
        lr_ep = lr * (0.5 * (1 + np.cos(np.pi * ep / max(epochs, 1))))
        for k in range(nb):
            idx = np.sort(order[k * batch:(k + 1) * batch])
            tot += net.train_step(x[idx], y[idx], lr_ep)
        hist.append(tot / max(nb, 1))
        if log:
            log("[learned] epoch %d loss %.4e (lr %.2e)" % (ep, hist[-1], lr_ep))
    w, b = net.get_params()
    net.close()
    return w, b, scale, hist


def save(path: str, w, b, scale, stages: int = 1):
    np.savez(path, w=w, b=b, scale=scale, stages=np.int32(stages))


def load(path: str):
    d = np.load(path)
    return d["w"], d["b"], d["scale"]


def trained_stages(path: str) -> int:
    """Refinement stages the weights were trained for (1 when not stored)."""
    d = np.load(path)
    return int(d["stages"]) if "stages" in d.files else 1


def attach(sim: Simulation, w, b, scale, stages: int = 1) -> ConvNet:
    """Inference network on the simulation's box; enables RC_WARM_LEARNED with
    `stages` refinement stages per solve."""
    L = sim.confs[0].Lx
    net = ConvNet(1, L, L)
    net.set_params(w, b)
    sc = np.ascontiguousarray(scale, dtype=np.float32)
    sim._check(sim.L.rc_attach_net(sim.h, net.h, _fp(sc)))
    sim._check(sim.L.rc_set_learned_stages(sim.h, stages))
    sim.set_warm_start(_lib.RC_WARM_LEARNED)
    sim._net = net  # keep alive
    return net

"""Host-side mirror of the reference backend for Python callers.

``Simulation`` plays the part of ``SimBackend``/``FD_CPUBackend`` with the
``WetPolarModel`` interaction (src/Backends/SimBackend.cpp:171-314,
src/Backends/FD_CPUBackend.cpp:19-160) on top of the C ABI: it reads confs
with the reference semantics, owns the random streams and drives
``rc_step``.  All simulation arithmetic runs in librheocell.so on the GPU.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from .conf import CellConf, Conf, RCexception
from . import walls
from .rng import ReferenceStream, nematic_directors

# constructor / BaseInteraction defaults of each interaction
_KEY_DEFAULTS = {
    # WetPolarModel.cpp:4-21
    "wetpolarmodel": {"gamma": 0.06, "lambda": 2.5, "omega": 0.004, "mu": 20.0, "kappa": 0.1, "CGtolerance": 1e-4,
                      "wall_slip": 0.5, "passive_alpha": 1.0, "J0": 1.0, "anchoring": 0},
    # WetModel.cpp:4-20 (it has no beta)
    "wetmodel": {"gamma": 0.01, "lambda": 2.5, "omega": 0.004, "mu": 3.0, "kappa": 0.1, "CGtolerance": 1e-4,
                 "wall_slip": 0.5, "passive_alpha": 1.0, "J0": 1.0, "anchoring": 0, "beta": 0.0},
}
_REQUIRED = {"wetpolarmodel": ["R", "dt", "zetaQ_self", "zetaQ_inter", "J_Q", "friction", "friction_cell", "beta"],
             "wetmodel": ["R", "dt", "zetaQ_self", "zetaQ_inter", "J_Q", "friction", "friction_cell"]}


def interaction_of(keys: Dict[str, object]) -> str:
    it = str(keys.get("interaction_type", "wetpolarmodel")).lower()
    if it not in _KEY_DEFAULTS:
        raise RCexception("interaction_type '%s' is not supported by the HIP backend" % keys.get("interaction_type"))
    return it


def model_params(keys: Dict[str, object], cg_max_iter: int = 0) -> _lib.rc_model_params:
    """Input keys -> rc_model_params.  The *_active keys are mandatory here:
    the reference leaves those members uninitialised when they are absent
    (WetPolarModel.cpp:4-21 vs :37-46, WetModel.cpp:4-20 vs :27-45)."""
    it = interaction_of(keys)
    for k in _REQUIRED[it]:
        if k not in keys:
            raise RCexception("key '%s' is required (the reference leaves it uninitialised)" % k)
    g = lambda k: float(keys.get(k, _KEY_DEFAULTS[it].get(k, 0.0)))  # noqa: E731
    if str(keys.get("box", "square")).lower() != "square":
        raise RCexception("only box = square is supported (with type = none or %s)" % ", ".join(walls.WALL_BOXES))
    return _lib.rc_model_params(
        gamma=g("gamma"), lambda_=g("lambda"), omega=g("omega"), mu=g("mu"), kappa=g("kappa"),
        R=int(float(keys["R"])), anchoring=int(float(keys.get("anchoring", 0))),
        zetaQ_self=g("zetaQ_self"), zetaQ_inter=g("zetaQ_inter"), J_Q=g("J_Q"), friction=g("friction"),
        friction_cell=g("friction_cell"), CGtolerance=g("CGtolerance"), wall_slip=g("wall_slip"),
        passive_alpha=g("passive_alpha"), beta=g("beta") if it == "wetpolarmodel" else 0.0, J0=g("J0"), dt=g("dt"),
        cg_max_iter=cg_max_iter, interaction=_lib.RC_INTERACTION[it])


def cell_from_conf(c: CellConf) -> _lib.rc_cell:
    """read_next_configuration per-cell state (SimBackend.cpp:273-296)."""
    x = _lib.rc_cell()
    x.LsubX, x.LsubY = c.LsubX, c.LsubY
    x.offset[0], x.offset[1] = c.offset
    x.corner = x.corner_old = c.corner
    x.x_sub_left, x.y_sub_bottom = c.LsubX, c.LsubY
    x.CoM[0], x.CoM[1] = c.CoM
    x.CoM_old[0], x.CoM_old[1] = c.CoM
    x.unrap[0], x.unrap[1] = int(c.unrap[0]), int(c.unrap[1])  # set_positions takes ints
    x.Q00, x.Q01 = c.Q00, c.Q01
    x.thetaQ = 0.5 * math.atan2(2 * c.Q01, 2 * c.Q00)
    x.thetaQ_old = x.thetaQ
    area = 0.0
    sumF = 0.0
    for v in c.phi.tolist():  # sequential sums, as the reader does
        area += v * v
        sumF += v
    x.area, x.sumF = area, sumF
    return x


class Simulation:
    """One handle: one or more independent boxes advanced together."""

    def __init__(self, boxes: Sequence[Tuple[Conf, Dict[str, object]]], seeds: Optional[Sequence[int]] = None,
                 cg_max_iter: int = 0, init: bool = True, equilibrate: bool = True,
                 cg_form: Optional[int] = None):
        self.L = _lib.load()
        self._keep = []    # arrays rc_create reads
        self.confs = [b[0] for b in boxes]
        self.keys = [b[1] for b in boxes]
        descs = (_lib.rc_box_desc * len(boxes))()
        for i, (conf, keys) in enumerate(boxes):
            if conf.Lx != conf.Ly:
                raise RCexception("The box in the configuration file is not a square")
            descs[i].L = conf.Lx
            descs[i].ncells = len(conf.cells)
            descs[i].params = model_params(keys, cg_max_iter)
            self._set_walls_and_forces(descs[i], keys, conf.Lx)
        self.h = C.c_void_p()
        rc = self.L.rc_create(descs, len(boxes), C.byref(self.h))
        self._check(rc)
        if cg_form is not None:
            self.set_cg_form(cg_form)
        cells = [c for conf in self.confs for c in conf.cells]
        self.ncells = len(cells)
        self.counts = [len(conf.cells) for conf in self.confs]
        sd = [int(k.get("seed", 1)) for k in self.keys] if seeds is None else list(seeds)
        nematic = [interaction_of(k) == "wetmodel" for k in self.keys]
        # the nematic director of each cell, from its drand48 draw at start-up
        dirs = [d for s, n, nem in zip(sd, self.counts, nematic) for d in (nematic_directors(s, n) if nem else [(0., 0.)] * n)]
        arr = (_lib.rc_cell * self.ncells)()
        for i, c in enumerate(cells):
            arr[i] = cell_from_conf(c)
            arr[i].nemQ[0], arr[i].nemQ[1] = dirs[i]
            arr[i].nemQ_old[0], arr[i].nemQ_old[1] = dirs[i]
        phi = np.ascontiguousarray(np.concatenate([c.phi for c in cells]) if cells else np.zeros(1))
        self._check(self.L.rc_upload_cells(self.h, arr, phi.ctypes.data_as(C.POINTER(C.c_double))))
        self.streams = [ReferenceStream(s, n, nem) for s, n, nem in zip(sd, self.counts, nematic)]
        self.step_count = 0
        if init:
            self.init_forces()
            if equilibrate:
                self.apply_changes_after_equilibration()

    def _set_walls_and_forces(self, desc, keys, L):
        """The box's wall field (BoxFactory: box = square with type = a wall
        box, src/Boxes/BoxFactory.cpp:16-26) and the external forces of
        external_forces_file (ForceFactory::make_forces, ForceFactory.cpp:61-150)."""
        btype = str(keys.get("type", "none")).lower()
        if btype in walls.WALL_BOXES:
            if "lambda_wall" not in keys:
                raise RCexception("box type '%s' needs lambda_wall" % btype)
            w = np.ascontiguousarray(walls.wall_field(btype, L, float(keys["lambda_wall"])))
            self._keep.append(w)
            desc.walls = w.ctypes.data_as(C.POINTER(C.c_double))
        if str(keys.get("external_forces", "0")).lower() in ("1", "true", "yes"):
            if "external_forces_file" not in keys:
                raise RCexception("external_forces = 1 needs external_forces_file")
            with open(str(keys["external_forces_file"])) as f:
                forces = walls.read_forces(f.read())
            if len(forces) > _lib.RC_MAX_FORCES:
                raise RCexception("at most %d external forces" % _lib.RC_MAX_FORCES)
            desc.nforces = len(forces)
            for j, fo in enumerate(forces):
                desc.forces[j] = _lib.rc_ext_force(kind=fo["kind"], kappa_wall=fo["kappa_wall"], lambda_=fo["lambda"],
                                                   shear_rate=fo["shear_rate"])

    # --- errors -----------------------------------------------------------
    def _check(self, rc: int):
        if rc != _lib.RC_OK:
            msg = self.L.rc_last_error(self.h).decode() if self.h else ""
            raise RCexception("%s: %s" % (_lib.RC_ERR.get(rc, rc), msg))

    def close(self):
        if getattr(self, "h", None):
            self.L.rc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- protocol of the reference backend --------------------------------
    def init_forces(self):
        self._check(self.L.rc_init_forces(self.h))

    def apply_changes_after_equilibration(self):
        self._check(self.L.rc_apply_changes_after_equilibration(self.h))

    def box_field(self, which: int, size_grid: int, box: int = 0):
        """VelocityField (which 0) / ShapeField (which 1) box fields of one box:
        (fx, fy, coarse_x, coarse_y), each L x L indexed [y, x]
        (src/Observables/VelocityField.cpp:69-128, ShapeField.cpp:69-128)."""
        L = self.confs[box].Lx if 0 <= box < len(self.confs) else 1   # the library rejects a bad box
        out = [np.zeros(L * L) for _ in range(4)]
        ptrs = [o.ctypes.data_as(C.POINTER(C.c_double)) for o in out]
        self._check(self.L.rc_box_field(self.h, box, which, size_grid, *ptrs))
        return tuple(o.reshape(L, L) for o in out)

    def set_warm_start(self, mode: int):
        self._check(self.L.rc_set_warm_start(self.h, mode))

    def set_cg_form(self, form: int):
        """_lib.RC_CG_SITES (default) or _lib.RC_CG_ROWS: the wet solve's iterates
        as box fields or as rows (the same iteration, rounding apart)."""
        self._check(self.L.rc_set_cg_form(self.h, form))

    def gaussians(self, nsteps: int) -> np.ndarray:
        parts = [s.gaussians(nsteps) for s in self.streams]
        return np.ascontiguousarray(np.concatenate(parts, axis=1))

    def step(self, nsteps: int = 1, gauss: Optional[np.ndarray] = None):
        if gauss is None:
            gauss = self.gaussians(nsteps)
        gauss = np.ascontiguousarray(gauss, dtype=np.float64)
        self._check(self.L.rc_step(self.h, nsteps, gauss.ctypes.data_as(C.POINTER(C.c_double))))
        self.step_count += nsteps

    # --- sync / observables -----------------------------------------------
    def total_rows(self) -> int:
        n = C.c_longlong()
        self._check(self.L.rc_total_rows(self.h, C.byref(n)))
        return n.value

    def download(self) -> Tuple[List[_lib.rc_cell], np.ndarray]:
        arr = (_lib.rc_cell * self.ncells)()
        phi = np.empty(self.total_rows())
        self._check(self.L.rc_download_cells(self.h, arr, phi.ctypes.data_as(C.POINTER(C.c_double))))
        return list(arr), phi

    def rows(self, which: int) -> np.ndarray:
        out = np.empty(self.total_rows())
        self._check(self.L.rc_get_rows(self.h, which, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def site_array(self, which: int) -> np.ndarray:
        out = np.empty(self.total_rows())
        self._check(self.L.rc_get_site_array(self.h, which, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def set_stage_shape(self, threads: int = 0, rows: int = 0):
        """Force the field-update kernel's shape (threads per cell, rows per
        thread strip); 1024, 0 forces the persistent dry stage (one launch
        over every cell); 0, 0 lets the autotuner choose."""
        self._check(self.L.rc_set_stage_shape(self.h, threads, rows))

    def stage_groups(self) -> List[Dict]:
        """Field-update kernel groups: shape in use, cells, valid shapes and the
        tuner's time per valid shape (us)."""
        ng = C.c_int()
        shape = (C.c_int * (3 * 16))()
        valid = (C.c_int * (16 * 16))()
        us = (C.c_double * (16 * 8))()
        self._check(self.L.rc_stage_groups(self.h, 16, C.byref(ng), shape, valid, us))
        out = []
        for g in range(min(ng.value, 16)):
            v = [(valid[16 * g + 2 * i], valid[16 * g + 2 * i + 1]) for i in range(8) if valid[16 * g + 2 * i]]
            out.append({"threads": shape[3 * g], "rows": shape[3 * g + 1], "cells": shape[3 * g + 2], "valid": v,
                        "tune_us": {"%dx%d" % sv: round(us[8 * g + i], 2) for i, sv in enumerate(v)}})
        return out

    def stage_persistent(self) -> Dict:
        """The persistent form of the dry fused stage (one launch over every
        cell): whether the layout allows it, whether it is used, and the
        tuner's predictor times of it and of the groups (us)."""
        el, used, tp, tg = C.c_int(), C.c_int(), C.c_double(), C.c_double()
        self._check(self.L.rc_stage_persistent(self.h, C.byref(el), C.byref(used), C.byref(tp), C.byref(tg)))
        return {"eligible": bool(el.value), "used": bool(used.value), "tune_us_persistent": round(tp.value, 2),
                "tune_us_groups": round(tg.value, 2)}

    def set_graphs(self, on: bool = True):
        self._check(self.L.rc_set_graphs(self.h, 1 if on else 0))

    def graph_stats(self) -> Tuple[int, int]:
        """(graph replays, steps launched one by one)"""
        g, e = C.c_longlong(), C.c_longlong()
        self._check(self.L.rc_graph_stats(self.h, C.byref(g), C.byref(e)))
        return g.value, e.value

    def growth_stats(self) -> Tuple[int, float]:
        """(patch growths served, host milliseconds spent on them)"""
        n, ms = C.c_longlong(), C.c_double()
        self._check(self.L.rc_growth_stats(self.h, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def cg_stats(self) -> _lib.rc_cg_stats_t:
        st = _lib.rc_cg_stats_t()
        self._check(self.L.rc_cg_stats(self.h, C.byref(st)))
        return st

    def enable_timing(self, on: bool = True):
        self._check(self.L.rc_enable_timing(self.h, 1 if on else 0))

    def timing(self, kind: int) -> Tuple[float, int]:
        ms = C.c_double()
        n = C.c_longlong()
        self._check(self.L.rc_timing(self.h, kind, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def to_conf(self, box: int = 0) -> Conf:
        """Configuration observable of one box (Configuration.cpp:46-55)."""
        cells, phi = self.download()
        b0 = sum(self.counts[:box])
        out = []
        start = sum(c.LsubX * c.LsubY for c in cells[:b0])
        L = self.confs[box].Lx
        for c in cells[b0:b0 + self.counts[box]]:
            S = c.LsubX * c.LsubY
            q = np.arange(S)
            lx = (q % c.LsubX + c.offset[0]) % c.LsubX
            ly = (q // c.LsubX + c.offset[1]) % c.LsubY
            cx, cy = c.corner % L, c.corner // L
            k = (cx + lx) % L + ((cy + ly) % L) * L
            out.append(CellConf(c.LsubX, c.LsubY, (c.CoM[0], c.CoM[1]), (c.offset[0], c.offset[1]), c.corner,
                                (c.unrap[0], c.unrap[1]), c.Q00, c.Q01, k.astype(np.int64), phi[start:start + S].copy()))
            start += S
        return Conf(self.confs[box].step + self.step_count, L, L, out)

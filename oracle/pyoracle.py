"""ctypes binding of the C oracle (rc_oracle.c).  Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


class OrcParams(C.Structure):
    _fields_ = [
        ("gamma", C.c_double), ("lambda_", C.c_double), ("omega", C.c_double), ("mu", C.c_double),
        ("kappa", C.c_double), ("R", C.c_int), ("anchoring", C.c_int),
        ("zetaQ_self", C.c_double), ("zetaQ_inter", C.c_double), ("J_Q", C.c_double),
        ("friction", C.c_double), ("friction_cell", C.c_double), ("CGtolerance", C.c_double),
        ("wall_slip", C.c_double), ("passive_alpha", C.c_double), ("beta", C.c_double),
        ("J0", C.c_double), ("dt", C.c_double), ("cg_max_iter", C.c_int), ("model", C.c_int),
    ]


MODELS = {"wetpolarmodel": 0, "wetmodel": 1}


class OrcCellInfo(C.Structure):
    _fields_ = [
        ("LsubX", C.c_int), ("LsubY", C.c_int), ("subSize", C.c_int), ("offset", C.c_int * 2),
        ("corner", C.c_int), ("corner_old", C.c_int), ("x_sub_left", C.c_int), ("y_sub_bottom", C.c_int),
        ("start", C.c_int),
        ("CoM", C.c_double * 2), ("CoM_old", C.c_double * 2), ("unrap", C.c_double * 2),
        ("Q00", C.c_double), ("Q01", C.c_double), ("thetaQ", C.c_double), ("thetaQ_old", C.c_double),
        ("area", C.c_double), ("area_old", C.c_double), ("sumF", C.c_double), ("sumF_old", C.c_double),
        ("perimeter", C.c_double), ("S00", C.c_double), ("S01", C.c_double),
        ("total_force", C.c_double * 2), ("total_force_repulsion", C.c_double * 2),
        ("nemQ", C.c_double * 2), ("nemQ_old", C.c_double * 2),
    ]


class OrcCgStats(C.Structure):
    _fields_ = [("solves", C.c_longlong), ("iters_total", C.c_longlong), ("last_iters", C.c_int * 2),
                ("last_relres", C.c_double * 2), ("last_cold", C.c_int)]


class OrcRng(C.Structure):
    _fields_ = [("x", C.c_ulonglong), ("has_next", C.c_int), ("next", C.c_double)]


def lib_path(flavour: str = "strict") -> str:
    """liborc_<flavour>.so beside this file, or flavour itself when it is a
    path to a build of the oracle (the planted-mistake builds of the tests)."""
    if flavour.endswith(".so"):
        return flavour
    return os.path.join(_HERE, "liborc_%s.so" % flavour)


_LIBS: Dict[str, C.CDLL] = {}


def _lib(flavour: str = "strict") -> C.CDLL:
    if flavour in _LIBS:
        return _LIBS[flavour]
    path = lib_path(flavour)
    if not os.path.exists(path):
        raise RuntimeError("oracle library %s is not built (run `make -C oracle`)" % path)
    L = C.CDLL(path)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.c_int
    L.orc_create.restype = vp
    L.orc_create.argtypes = [ip, ip, C.POINTER(OrcParams), C.c_long]
    L.orc_destroy.argtypes = [vp]
    L.orc_last_error.restype = C.c_char_p
    L.orc_last_error.argtypes = [vp]
    L.orc_set_cell.argtypes = [vp, ip, ip, ip, C.c_double, C.c_double, ip, ip, ip, C.c_double, C.c_double,
                               C.c_double, C.c_double, dp]
    L.orc_init.argtypes = [vp]
    L.orc_apply_changes_after_equilibration.argtypes = [vp]
    L.orc_step.argtypes = [vp, ip]
    L.orc_num_cells.argtypes = [vp]
    L.orc_size_rows.argtypes = [vp]
    L.orc_get_cell.argtypes = [vp, ip, C.POINTER(OrcCellInfo)]
    L.orc_get_site_array.argtypes = [vp, ip, ip, dp]
    L.orc_get_rows.argtypes = [vp, ip, dp]
    L.orc_get_box.argtypes = [vp, ip, dp]
    L.orc_get_cg_stats.argtypes = [vp, C.POINTER(OrcCgStats)]
    L.orc_set_warm_mode.argtypes = [vp, ip]
    L.orc_rng_seed.argtypes = [C.POINTER(OrcRng), C.c_long]
    L.orc_rng_drand48.restype = C.c_double
    L.orc_rng_drand48.argtypes = [C.POINTER(OrcRng)]
    L.orc_rng_gaussian.restype = C.c_double
    L.orc_rng_gaussian.argtypes = [C.POINTER(OrcRng)]
    L.orc_trig_table.argtypes = [ip, ip, dp]
    L.orc_apply_A.argtypes = [vp, dp, dp]
    L.orc_assemble_dense.argtypes = [vp, dp]
    L.orc_test_nbrs.argtypes = [ip, ip, C.POINTER(C.c_int)]
    L.orc_test_set_positions.argtypes = [ip, C.POINTER(C.c_int), dp, dp, ip]
    L.orc_test_set_site.argtypes = [vp, ip, ip, dp]
    L.orc_test_set_rows.argtypes = [vp, ip, dp]
    L.orc_test_set_polarity.argtypes = [vp, ip, C.c_double, C.c_double, C.c_double, C.c_double]
    L.orc_test_set_director.argtypes = [vp, ip, C.c_double, C.c_double, C.c_double, C.c_double]
    L.orc_test_first_step.argtypes = [vp, ip]
    L.orc_set_walls.argtypes = [vp, dp]
    L.orc_add_force.argtypes = [vp, ip, C.c_double, C.c_double, C.c_double]
    L.orc_test_forces.argtypes = [vp]
    _LIBS[flavour] = L
    return L


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def params_from_keys(keys: Dict[str, object], cg_max_iter: int = 0) -> OrcParams:
    g = lambda k, d=None: float(keys[k]) if k in keys else d  # noqa: E731
    nem = str(keys.get("interaction_type", "wetpolarmodel")).lower() == "wetmodel"
    # constructor defaults: WetPolarModel.cpp:4-21, WetModel.cpp:4-20
    return OrcParams(gamma=g("gamma", 0.01 if nem else 0.06), lambda_=g("lambda", 2.5), omega=g("omega", 0.004),
                     mu=g("mu", 3.0 if nem else 20.0),
                     kappa=g("kappa", 0.1), R=int(float(keys["R"])), anchoring=int(float(keys.get("anchoring", 0))),
                     zetaQ_self=g("zetaQ_self"), zetaQ_inter=g("zetaQ_inter"), J_Q=g("J_Q"),
                     friction=g("friction"), friction_cell=g("friction_cell"), CGtolerance=g("CGtolerance", 1e-4),
                     wall_slip=g("wall_slip", 0.5), passive_alpha=g("passive_alpha", 1.0),
                     beta=g("beta", 0.0) if nem else g("beta"),
                     J0=g("J0", 1.0), dt=g("dt"), cg_max_iter=cg_max_iter,
                     model=MODELS[str(keys.get("interaction_type", "wetpolarmodel")).lower()])


class OracleSim:
    """One reference process: box, cells, RNG, interaction state."""

    def __init__(self, conf, keys: Dict[str, object], seed: Optional[int] = None, flavour: str = "strict",
                 cg_max_iter: int = 0, init: bool = True, equilibrate: bool = True, walls=None, forces=()):
        """walls: the box's getWalls field (L*L) or None; forces: dicts with
        kind, kappa_wall, lambda, shear_rate (rheocell_amd.walls.read_forces)."""
        self.L = _lib(flavour)
        self.N = len(conf.cells)
        self.prm = params_from_keys(keys, cg_max_iter)
        sd = int(keys.get("seed", 1)) if seed is None else seed
        self.h = self.L.orc_create(conf.Lx, self.N, C.byref(self.prm), sd)
        err = self.L.orc_last_error(self.h)
        if err:
            raise RuntimeError(err.decode())
        for i, c in enumerate(conf.cells):
            phi = np.ascontiguousarray(c.phi, dtype=np.float64)
            rc = self.L.orc_set_cell(self.h, i, c.LsubX, c.LsubY, c.CoM[0], c.CoM[1], c.offset[0], c.offset[1],
                                     c.corner, int(c.unrap[0]), int(c.unrap[1]), c.Q00, c.Q01, _dp(phi))
            self._check(rc)
        if walls is not None:
            self._walls = np.ascontiguousarray(walls, dtype=np.float64)
            self._check(self.L.orc_set_walls(self.h, _dp(self._walls)))
        for f in forces:
            self._check(self.L.orc_add_force(self.h, int(f["kind"]), float(f["kappa_wall"]), float(f["lambda"]),
                                             float(f["shear_rate"])))
        if init:
            self._check(self.L.orc_init(self.h))
            if equilibrate:
                self.L.orc_apply_changes_after_equilibration(self.h)

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self.L.orc_last_error(self.h).decode())

    def __del__(self):
        try:
            self.L.orc_destroy(self.h)
        except Exception:
            pass

    def step(self, n: int = 1):
        self._check(self.L.orc_step(self.h, n))

    def set_warm_mode(self, mode: int):
        self.L.orc_set_warm_mode(self.h, mode)

    def cell(self, i: int) -> OrcCellInfo:
        info = OrcCellInfo()
        self._check(self.L.orc_get_cell(self.h, i, C.byref(info)))
        return info

    def site(self, i: int, which: int = 0) -> np.ndarray:
        n = self.cell(i).subSize
        out = np.empty(n)
        self._check(self.L.orc_get_site_array(self.h, i, which, _dp(out)))
        return out

    def phi(self, i: int) -> np.ndarray:
        return self.site(i, 0)

    def rows(self, which: int) -> np.ndarray:
        out = np.empty(self.L.orc_size_rows(self.h))
        self._check(self.L.orc_get_rows(self.h, which, _dp(out)))
        return out

    def box(self, which: int, Lbox: int) -> np.ndarray:
        out = np.empty(Lbox * Lbox)
        self._check(self.L.orc_get_box(self.h, which, _dp(out)))
        return out

    def cg_stats(self) -> OrcCgStats:
        st = OrcCgStats()
        self.L.orc_get_cg_stats(self.h, C.byref(st))
        return st

    # --- hooks on hand-set state (closed-form pins) -------------------------
    def set_site(self, i: int, which: int, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        assert v.size == self.cell(i).subSize
        self._check(self.L.orc_test_set_site(self.h, i, which, _dp(v)))

    def set_rows(self, which: int, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        assert v.size == self.L.orc_size_rows(self.h)
        self._check(self.L.orc_test_set_rows(self.h, which, _dp(v)))

    def set_polarity(self, i: int, theta: float, theta_old: float, tr=(0.0, 0.0)):
        self._check(self.L.orc_test_set_polarity(self.h, i, theta, theta_old, tr[0], tr[1]))

    def set_director(self, i: int, n, n_old):
        self._check(self.L.orc_test_set_director(self.h, i, n[0], n[1], n_old[0], n_old[1]))

    def first_step(self, store: bool):
        self._check(self.L.orc_test_first_step(self.h, 1 if store else 0))

    def forces(self):
        self._check(self.L.orc_test_forces(self.h))

    def apply_changes_after_equilibration(self):
        self.L.orc_apply_changes_after_equilibration(self.h)

    def apply_A(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        self._check(self.L.orc_apply_A(self.h, _dp(x), _dp(y)))
        return y

    def assemble_dense(self) -> np.ndarray:
        n = self.L.orc_size_rows(self.h)
        A = np.empty((n, n))
        self._check(self.L.orc_assemble_dense(self.h, _dp(A)))
        return A


def rng_stream(seed: int, n: int, kind: str = "drand48", flavour: str = "strict") -> np.ndarray:
    L = _lib(flavour)
    r = OrcRng()
    L.orc_rng_seed(C.byref(r), seed)
    f = L.orc_rng_drand48 if kind == "drand48" else L.orc_rng_gaussian
    return np.array([f(C.byref(r)) for _ in range(n)])


def trig_table(L: int, which: int, flavour: str = "strict") -> np.ndarray:
    out = np.empty(L)
    _lib(flavour).orc_trig_table(L, which, _dp(out))
    return out


def test_nbrs(LX: int, LY: int, flavour: str = "strict") -> np.ndarray:
    out = np.empty(9 * LX * LY, dtype=np.int32)
    _lib(flavour).orc_test_nbrs(LX, LY, out.ctypes.data_as(C.POINTER(C.c_int)))
    return out


def test_set_positions(Lbox, LX, LY, offx, offy, corner, ux, uy, comx, comy, xsl, ysb, phi, flavour="strict"):
    """MultiPhaseField::set_positions(box) restated; returns the new state."""
    ints = np.array([LX, LY, offx, offy, corner, corner, xsl, ysb], dtype=np.int32)
    dbl = np.array([ux, uy, comx, comy, 0.0, 0.0])
    cap = 4 * max(LX * LY, 64 * 64)
    buf = np.zeros(cap)
    buf[:LX * LY] = phi
    rc = _lib(flavour).orc_test_set_positions(Lbox, ints.ctypes.data_as(C.POINTER(C.c_int)), _dp(dbl), _dp(buf), cap)
    if rc != 0:
        raise RuntimeError("set_positions failed")
    S = int(ints[0]) * int(ints[1])
    return {"LX": int(ints[0]), "LY": int(ints[1]), "offset": [int(ints[2]), int(ints[3])], "corner": int(ints[4]),
            "corner_old": int(ints[5]), "unrap": [dbl[0], dbl[1]], "CoM": [dbl[2], dbl[3]], "area": dbl[4],
            "phi": buf[:S].copy()}

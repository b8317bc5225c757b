"""Walls and external forces: the host's wall fields and the oracle's
wall-masked stencils against the reference's own wall boxes and
BaseInteraction code (golden vectors, tests/golden/make_golden.py), and the
external forces' free energy and velocity profiles against closed forms.
Each oracle pin is also run against oracle builds with one planted mistake.

  walls_stencils     BaseInteraction.h:196-263 in SquareWalls / CircleWalls / TriangleWalls boxes
  wall_free_energy   BaseField.h:54-66, ChannelWalls.cpp:26-31, ShearFlowChannel.cpp:33-36
  wall_forces_zero   WetPolarModel.cpp:329-339, WetModel.cpp:629-640 (no force on a wall site)
  shear_profile      ShearFlowChannel.cpp:41-45, FD_CPUBackend.cpp:62-72, BaseField.h:107-115
  pois_profile       PoiseuilleFlowChannel.cpp (velocity_profile)
  stretching         Stretching.cpp (velocity_profile of the cell index)
  nem_wall_density   WetModel.cpp:540-610 with BaseField.h:67-79, ShearFlowChannel.cpp:33-46
                     (the forces' free-energy density in the nematic stress)
"""
import json
import math
import os
import shutil

import numpy as np
import pytest

from oracle import OracleSim
from rheocell_amd import confgen
from rheocell_amd.conf import CellConf, Conf
from rheocell_amd.walls import read_forces, wall_field
from tests.test_oracle_nematic import _nkeys
from tests.test_oracle_pins import _cell, _keys, _logical, build_planted

HERE = os.path.dirname(os.path.abspath(__file__))
G = json.load(open(os.path.join(HERE, "golden", "ref_walls.json")))


@pytest.mark.parametrize("rec", G["walls"], ids=lambda r: "%s-%d" % (r["type"], r["L"]))
def test_wall_fields_match_the_reference_boxes(rec):
    assert np.array_equal(wall_field(rec["type"], rec["L"], rec["lambda_wall"]), np.array(rec["walls"]))


def test_read_forces_file():
    text = "# comment\n{\ntype = shear_flow_channel\nfield = -1\nkappa_wall = 0.5\nlambda = 3.0\nshear_rate = 0.01\n}\n" \
           "{\ntype = square_walls\nfield = -1\nkappa_wall = 0.2\n}\n"
    f = read_forces(text)
    assert [x["kind"] for x in f] == [1, 0]
    assert f[0]["kappa_wall"] == 0.5 and f[0]["lambda"] == 3.0 and f[0]["shear_rate"] == 0.01
    assert f[1]["lambda"] == 2.0


# --------------------------------------------------------------------------
def pin_walls_stencils(flavour="strict"):
    """The oracle's stencils in wall boxes equal the reference's bit for bit."""
    for rec in G["stencilw"]:
        Lb, LX, LY, ox, oy, corner = rec["args"]
        S = LX * LY
        c = CellConf(LX, LY, (LX / 2, LY / 2), (ox, oy), corner, (0.0, 0.0), 1.0, 0.0, np.zeros(S, dtype=np.int64),
                     np.array(rec["phi"]))
        w = wall_field(rec["type"], Lb, rec["lambda_wall"])
        s = OracleSim(Conf(0, Lb, Lb, [c]), _keys(wall_slip=rec["wall_slip"]), flavour=flavour, walls=w)
        assert np.array_equal(s.site(0, 2), np.array(rec["dx"]))
        assert np.array_equal(s.site(0, 3), np.array(rec["dy"]))
        assert np.array_equal(s.site(0, 4), np.array(rec["lap"]))


def _flat_in_walls(flavour, forces, c=0.4, **kw):
    """A flat patch reaching into the bottom wall of a square_walls box."""
    L, LX = 48, 15
    w = wall_field("square_walls", L, 4.0)
    keys = _keys(gamma=0.02, mu=30.0, **kw)
    s = OracleSim(Conf(0, L, L, [_cell(L, 16, 0, LX, LX, lambda x, y: c)]), keys, flavour=flavour, walls=w,
                  forces=forces)
    lx, ly = _logical(LX, LX)
    return s, keys, w[(16 + lx) + ly * L], lx, ly


def pin_wall_free_energy(flavour="strict"):
    """The walls' free energy 2 kappa_w / lambda_f phi w^2 (channel walls) plus
    4 kappa_w / lambda_f phi w^2 (a shear-flow force) is added to mu on wall
    sites only."""
    forces = [dict(kind=0, kappa_wall=0.3, **{"lambda": 2.0}, shear_rate=0.0),
              dict(kind=1, kappa_wall=0.5, **{"lambda": 3.0}, shear_rate=0.0)]
    c = 0.4
    s, keys, w, lx, ly = _flat_in_walls(flavour, forces, c)
    s0, _, _, _, _ = _flat_in_walls(flavour, [], c)
    mu, mu0 = s.site(0, 1), s0.site(0, 1)
    inner = (lx >= 1) & (lx <= 13) & (ly >= 1) & (ly <= 13)
    ext = 2.0 * 0.3 / 2.0 * c + 4.0 * 0.5 / 3.0 * c
    wall, opn = inner & (w == 1), inner & (w == 0)
    assert wall.sum() > 5 and opn.sum() > 5
    assert np.allclose(mu[wall] - mu0[wall], ext, rtol=1e-12, atol=0)
    assert np.array_equal(mu[opn], mu0[opn])


def pin_wall_forces_zero(flavour="strict"):
    """No force, hence no dry velocity, on a wall site; the open sites next to
    the walls feel the jump of mu."""
    forces = [dict(kind=0, kappa_wall=0.3, **{"lambda": 2.0}, shear_rate=0.0)]
    s, keys, w, lx, ly = _flat_in_walls(flavour, forces)
    vy = s.rows(1)
    assert np.all(vy[w == 1] == 0)
    assert np.abs(vy[w == 0]).max() > 1e-4


def _advected(flavour, kind, rate=0.02, lam=3.0, LX=15, y0=10, index_cells=1, equilibrate=True):
    """The predictor of a patch with mu = v = 0 and a unit x-gradient: the
    field moves by -dt V_ext(box site) (first_step's external advection)."""
    L = 48
    forces = [dict(kind=0, kappa_wall=0.1, **{"lambda": 2.0}, shear_rate=0.0),
              dict(kind=kind, kappa_wall=0.0, **{"lambda": lam}, shear_rate=rate)]
    cells = [_cell(L, 16, y0, LX, LX, lambda x, y: 0.5) for _ in range(index_cells)]
    s = OracleSim(Conf(0, L, L, cells), _keys(dt=0.1), flavour=flavour, forces=forces, equilibrate=equilibrate)
    n = LX * LX
    rows = index_cells * n
    s.set_rows(0, np.zeros(rows))
    s.set_rows(1, np.zeros(rows))
    for i in range(index_cells):
        s.set_site(i, 1, np.zeros(n))
        s.set_site(i, 2, np.ones(n))
        s.set_site(i, 3, np.zeros(n))
    s.first_step(True)
    _, ly = _logical(LX, LX)
    V = [-(s.phi(i) - 0.5) / 0.1 for i in range(index_cells)]
    return V, y0 + ly, L


def pin_shear_profile(flavour="strict"):
    """A linear shear between the walls: V_x(y) is linear in y with slope
    rate / (L - 2 (lambda + 1)), zero at y = lambda + 1/2 + h/2 (the channel's
    mid-line), and 0 for y <= lambda or y >= L - lambda; before equilibration
    ends the rate is 0."""
    rate, lam = 0.02, 3.0
    V0, _, _ = _advected(flavour, 1, rate, lam, LX=15, y0=0, equilibrate=False)
    assert np.all(V0[0] == 0)      # the shear rate is 0 until equilibration ends
    V, y, L = _advected(flavour, 1, rate, lam, LX=15, y0=0)
    V = V[0]
    h = L - 2 * (lam + 1)
    inside = (y > lam) & (y < L - lam)
    assert np.all(V[~inside] == 0)
    slope, icpt = np.polyfit(y[inside], V[inside], 1)
    assert abs(slope - rate / h) < 1e-12 and abs(-icpt / slope - ((lam + 1) - 0.5 + 0.5 * h)) < 1e-9
    assert np.abs(V[inside] - (slope * y[inside] + icpt)).max() < 1e-14


def pin_pois_profile(flavour="strict"):
    """A parabola 4 rate u (1 - u / h) / h of u = y - (lambda + 1),
    h = L - 2 (lambda + 1): zero at u = 0, peak rate at u = h / 2."""
    rate, lam = 0.02, 3.0
    V, y, L = _advected(flavour, 2, rate, lam, LX=15, y0=17)
    V = V[0]
    h = L - 2 * (lam + 1)
    u = y - (lam + 1)
    a, b, c = np.polyfit(u, V, 2)
    assert abs(a + 4 * rate / h ** 2) < 1e-12 and abs(b - 4 * rate / h) < 1e-12 and abs(c) < 1e-12


def pin_stretching(flavour="strict"):
    """Stretching pulls cells whose index is 0 mod 10 with -rate and 9 mod 10
    with +rate along x; the others are not advected."""
    rate = 0.02
    V, _, _ = _advected(flavour, 3, rate, LX=9, y0=17, index_cells=11)
    for i, v in enumerate(V):
        expect = -rate if i % 10 == 0 else (rate if i % 10 == 9 else 0.0)
        assert np.allclose(v, expect, rtol=1e-12, atol=1e-15), (i, v[:3], expect)


def pin_nem_wall_density(flavour="strict"):
    """The nematic model's stress is f - mu phi with the forces' density in f
    and their free energy in mu: on a flat patch c across the bottom wall of a
    square_walls box (gradients and Laplacians 0, repulsion 0 for one cell),
    the stress jumps by (fD - fF) c^2 from an open site to a wall site, where
    a ShearFlowChannel force has fF = 4 kw / lf and fD = 2 kw / lf.  An open
    site whose south neighbour is a wall feels f_y = -1/2 (s_N - s_S) =
    1/2 (fD - fF) c^2 (dry: v = f / xi, xi = 1); the open sites above it
    feel nothing."""
    L, LX, c, kw, lf = 48, 15, 0.4, 0.5, 3.0
    w = wall_field("square_walls", L, 4.0)
    forces = [dict(kind=1, kappa_wall=kw, **{"lambda": lf}, shear_rate=0.0)]
    keys = _nkeys(friction_cell=0, friction=1.0, passive_alpha=1.0)
    s = OracleSim(Conf(0, L, L, [_cell(L, 16, 0, LX, LX, lambda x, y: c)]), keys, flavour=flavour, walls=w,
                  forces=forces)
    lx, ly = _logical(LX, LX)
    vy = s.rows(1)
    inner = (lx >= 2) & (lx <= 12)
    expect = 0.5 * (2.0 * kw / lf - 4.0 * kw / lf) * c * c
    assert np.allclose(vy[inner & (ly == 4)], expect, rtol=1e-12, atol=0)
    assert np.all(vy[inner & (ly >= 5) & (ly <= 10)] == 0)
    assert np.all(vy[inner & (ly <= 3)] == 0)   # wall sites


PINS = {name[4:]: fn for name, fn in list(globals().items()) if name.startswith("pin_")}


@pytest.mark.parametrize("name", sorted(PINS))
def test_pin(name):
    PINS[name]()


MUTANTS = [
    ("open_inclusive", "return get_walls(s, k) < s->act.wall_slip;", "return get_walls(s, k) <= 1.0;", ["walls_stencils"]),
    ("derivx_one_side", "if(open_site(s, bn[5]) && open_site(s, bn[3]))", "if(open_site(s, bn[5]))", ["walls_stencils"]),
    ("lap_one_sided", "else if(E) return f[nb[7]] + f[nb[1]] + 2 * f[nb[5]] - 4 * f[q];",
     "else if(E) return f[nb[7]] + f[nb[1]] + f[nb[5]] - 4 * f[q];", ["walls_stencils"]),
    ("channel_coefficient", "case ORC_FORCE_CHANNEL_WALLS: return 2.0 * kappa_wall", "case ORC_FORCE_CHANNEL_WALLS: return 4.0 * kappa_wall",
     ["wall_free_energy"]),
    ("flow_coefficient", "case ORC_FORCE_POIS_FLOW: return 4.0 * kappa_wall", "case ORC_FORCE_POIS_FLOW: return 2.0 * kappa_wall",
     ["wall_free_energy"]),
    ("force_on_walls", "	if(open_site(s, p->map[q])) {\n		s->fx[row] = f_passive_x * s->act.passive_alpha + fQx",
     "	if(1) {\n		s->fx[row] = f_passive_x * s->act.passive_alpha + fQx", ["wall_forces_zero"]),
    ("shear_offset", "V[0] = ((y - (lambda + 1) + 0.5) -", "V[0] = ((y - (lambda + 1) - 0.5) -", ["shear_profile"]),
    ("pois_shape", "V[0] = 4 * (y - (lambda + 1)) * shear_rate * (1 - (y - (lambda + 1)) /",
     "V[0] = 4 * (y - (lambda + 1)) * shear_rate * (1 + (y - (lambda + 1)) /", ["pois_profile"]),
    ("stretching_side", "else if(x == 10 - 1) V[0] = shear_rate;", "else if(x == 10 - 1) V[0] = -shear_rate;", ["stretching"]),
    ("first_force_wins", "		V[0] = 0.;\n		V[1] = 0.;\n", "		if(f > 0) break;\n		V[0] = 0.;\n		V[1] = 0.;\n",
     ["shear_profile"]),
    ("rate_before_equilibration", "s->force[s->nforce].rate = 0.;", "s->force[s->nforce].rate = shear_rate;", ["shear_profile"]),
    ("flow_density", "return 2.0 * kappa_wall / lambda * phi * phi * walls * walls;",
     "return 4.0 * kappa_wall / lambda * phi * phi * walls * walls;", ["nem_wall_density"]),
]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc is needed to build the planted-mistake oracles")
@pytest.mark.parametrize("label,old,new,pins", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_planted_mistake_is_caught(tmp_path, label, old, new, pins):
    lib = build_planted(tmp_path, old, new, label)
    for name in pins:
        with pytest.raises(AssertionError):
            PINS[name](flavour=lib)

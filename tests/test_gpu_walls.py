"""Walls and external forces on the GPU against the oracle: cells pressed
against the walls of the reference's square wall boxes (square, circle,
triangle), the walls' free energy, and the shear, Poiseuille and stretching
advection; both interactions on the dry path, and the wet solve with an
external flow in a periodic box.

Tolerance: max|dphi| and max|dv| <= 1e-10 (FP64)."""
import numpy as np
import pytest

from oracle import OracleSim
from rheocell_amd import confgen
from rheocell_amd.sim import Simulation
from rheocell_amd.walls import read_forces, wall_field

pytestmark = pytest.mark.gpu

TOL = 1e-10

FORCES = {
    "channel+shear": "{\ntype = square_walls\nfield = -1\nkappa_wall = 0.2\nlambda = 2.0\n}\n"
                     "{\ntype = shear_flow_channel\nfield = -1\nkappa_wall = 0.1\nlambda = 3.0\nshear_rate = 0.05\n}\n",
    "pois": "{\ntype = channel_walls\nfield = -1\nkappa_wall = 0.3\n}\n"
            "{\ntype = pois_flow_channel\nfield = -1\nkappa_wall = 0.05\nlambda = 4.0\nshear_rate = 0.04\n}\n",
    "stretching": "{\ntype = triangle_walls\nfield = -1\nkappa_wall = 0.2\n}\n"
                  "{\ntype = stretching\nfield = -1\nshear_rate = 0.03\n}\n",
}


def _case(tmp_path, box_type, forces, L=96, nside=4, seed=41, **kw):
    ff = tmp_path / "external.conf"
    ff.write_text(FORCES[forces])
    keys = dict(confgen.BASE_KEYS, seed=seed, external_forces=1, external_forces_file=str(ff), **kw)
    walls = None
    if box_type != "none":
        keys.update(type=box_type, lambda_wall=4.0)
        walls = wall_field(box_type, L, 4.0)
    conf = confgen.lattice_conf(L, nside, seed=seed)
    sim = Simulation([(conf, keys)])
    orc = OracleSim(conf, keys, walls=walls, forces=read_forces(FORCES[forces]))
    return sim, orc


def _compare(sim, orc):
    cells, phi = sim.download()
    start, worst = 0, 0.0
    for i, c in enumerate(cells):
        oc = orc.cell(i)
        assert (c.LsubX, c.LsubY, c.offset[0], c.offset[1], c.corner) == \
            (oc.LsubX, oc.LsubY, oc.offset[0], oc.offset[1], oc.corner), i
        S = c.LsubX * c.LsubY
        worst = max(worst, float(np.abs(phi[start:start + S] - orc.phi(i)).max()))
        start += S
    dv = max(float(np.abs(sim.rows(0) - orc.rows(0)).max()), float(np.abs(sim.rows(1) - orc.rows(1)).max()))
    return worst, dv


@pytest.mark.parametrize("box_type,forces", [("square_walls", "channel+shear"), ("circle_walls", "pois"),
                                             ("triangle_walls", "stretching")])
@pytest.mark.parametrize("interaction", ["wetpolarmodel", "wetmodel"])
def test_walls_dry_match_oracle(tmp_path, box_type, forces, interaction):
    extra = dict(interaction_type="wetmodel", J_Q=0.1, zetaQ_self=0.3, zetaQ_inter=0.1) if interaction == "wetmodel" \
        else dict(omega=0.01)
    sim, orc = _case(tmp_path, box_type, forces, **extra)
    for n in (1, 14):
        sim.step(n)
        orc.step(n)
        w, dv = _compare(sim, orc)
        print(interaction, box_type, forces, "step", sim.step_count, "max|dphi| =", w, "max|dv| =", dv, flush=True)
        assert w <= TOL and dv <= TOL
    # cells reached the walls: some velocity rows sit on wall sites (zero force there)
    assert np.count_nonzero(sim.rows(0) == 0) > 0


def test_wet_solve_with_an_external_flow(tmp_path):
    sim, orc = _case(tmp_path, "none", "channel+shear", L=64, nside=3, seed=42, friction_cell=1, CGtolerance=1e-12)
    for n in (1, 4):
        sim.step(n)
        orc.step(n)
        w, dv = _compare(sim, orc)
        print("wet, shear flow, step", sim.step_count, "max|dphi| =", w, "max|dv| =", dv, flush=True)
        assert w <= TOL and dv <= TOL


def test_walls_with_the_wet_solve_are_refused(tmp_path):
    from rheocell_amd.conf import RCexception
    with pytest.raises(RCexception, match="walls with the wet solve"):
        _case(tmp_path, "square_walls", "pois", friction_cell=1)

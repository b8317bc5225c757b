"""The nematic wet model (interaction_type = wetmodel, WetModel.cpp) on the
GPU against the oracle: dry and wet, with adhesion and anchoring, across the
solver's 100-solve restart, and at the parameters of the reference's own
example input (examples/input: 100 cells, 140^2, friction_cell = 1).

Tolerance: max|dphi|, max|dv| and the director within 1e-10 (FP64; the wet
cases solve to 1e-12 so both solvers' answers are closer than that)."""
import numpy as np
import pytest

from oracle import OracleSim
from rheocell_amd import confgen
from rheocell_amd.sim import Simulation

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _keys(**kw):
    # examples/input of the reference (WetModel parameters)
    k = dict(confgen.BASE_KEYS, interaction_type="wetmodel", gamma=0.06, mu=20, **{"lambda": 2.0}, kappa=0.5,
             omega=0, wall_slip=2, passive_alpha=1, friction=1, friction_cell=0, J_Q=0.1, anchoring=0,
             zetaQ_self=0.3, zetaQ_inter=0.0, CGtolerance=1e-12)
    k.pop("beta", None)
    k.update(kw)
    return k


def _compare(sim, orc):
    cells, phi = sim.download()
    start, worst, worst_n = 0, 0.0, 0.0
    for i, c in enumerate(cells):
        oc = orc.cell(i)
        assert (c.LsubX, c.LsubY, c.offset[0], c.offset[1], c.corner) == \
            (oc.LsubX, oc.LsubY, oc.offset[0], oc.offset[1], oc.corner), i
        S = c.LsubX * c.LsubY
        worst = max(worst, float(np.abs(phi[start:start + S] - orc.phi(i)).max()))
        worst_n = max(worst_n, abs(c.nemQ[0] - oc.nemQ[0]), abs(c.nemQ[1] - oc.nemQ[1]),
                      abs(c.Q00 - oc.Q00), abs(c.Q01 - oc.Q01))
        start += S
    dv = max(float(np.abs(sim.rows(0) - orc.rows(0)).max()), float(np.abs(sim.rows(1) - orc.rows(1)).max()))
    return worst, dv, worst_n


@pytest.mark.parametrize("opts", [dict(), dict(omega=0.01, zetaQ_inter=0.2), dict(anchoring=1)],
                         ids=["dry", "adhesion_inter", "anchoring"])
def test_nematic_dry_matches_oracle(opts):
    conf = confgen.lattice_conf(96, 4, seed=31)
    keys = _keys(seed=31, **opts)
    sim = Simulation([(conf, keys)])
    orc = OracleSim(conf, keys)
    for n in (1, 19):
        sim.step(n)
        orc.step(n)
        w, dv, wn = _compare(sim, orc)
        print(opts, "step", sim.step_count, "max|dphi| =", w, "max|dv| =", dv, "director", wn, flush=True)
        assert w <= TOL and dv <= TOL and wn <= TOL


def test_nematic_wet_across_the_restart():
    conf = confgen.lattice_conf(64, 3, seed=32)
    keys = _keys(seed=32, friction_cell=1, zetaQ_inter=0.1)
    sim = Simulation([(conf, keys)])
    orc = OracleSim(conf, keys)
    for n in (1, 4, 50, 1):    # 112 force evaluations: the 101st wet solve starts cold
        sim.step(n)
        orc.step(n)
        w, dv, wn = _compare(sim, orc)
        print("wet nematic step", sim.step_count, "max|dphi| =", w, "max|dv| =", dv, "director", wn, "iters",
              list(sim.cg_stats().last_iters), list(orc.cg_stats().last_iters), flush=True)
        assert w <= TOL and dv <= TOL and wn <= TOL
    # the same solves with the same iteration counts: the cold restart fell on
    # the same solve (a tight tolerance hides it in the fields)
    assert sim.cg_stats().solves == orc.cg_stats().solves
    assert sim.cg_stats().iters_total == orc.cg_stats().iters_total


def test_nematic_reference_example_parameters():
    # examples/input: wetmodel, 100 cells in a 140^2 box, friction_cell = 1
    conf = confgen.lattice_conf(140, 10, seed=1)
    keys = _keys(seed=1, friction_cell=1)
    sim = Simulation([(conf, keys)])
    orc = OracleSim(conf, keys)
    sim.step(3)
    orc.step(3)
    w, dv, wn = _compare(sim, orc)
    print("examples/input parameters, 3 steps: max|dphi| =", w, "max|dv| =", dv, "director", wn, flush=True)
    assert w <= TOL and dv <= TOL and wn <= TOL

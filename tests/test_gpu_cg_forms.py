"""The two forms of the wet solve's Jacobi-PCG, rows and box-site fields, on
the GPU: the same iteration (Eigen's ConjugateGradient, WetPolarModel.cpp:200-233)
from the same start, differing in rounding only.

At the reference tolerance (1e-8) every solve takes the same number of
iterations in both forms, and the fields agree to within FP64 rounding of the
iteration: max|dv| <= 1e-11 |v| and max|dphi| <= 1e-12 after several steps,
including a cold restart; against the oracle the site form keeps the parity
bar (1e-10 on phi and v at a tight tolerance, where both solvers' answers are
closer than that)."""
import numpy as np
import pytest

from oracle import OracleSim
from rheocell_amd import _lib, confgen
from rheocell_amd.sim import Simulation

pytestmark = pytest.mark.gpu


def _fields(sim):
    _, phi = sim.download()
    return phi, sim.rows(0), sim.rows(1)


@pytest.mark.parametrize("workload,steps", [("wet256_400", 13), ("wet512_1600", 3)])
def test_site_and_row_forms_agree(workload, steps):
    conf, keys = confgen.workload(workload, seed=6)
    sims = {f: Simulation([(conf, keys)], cg_form=f) for f in (_lib.RC_CG_SITES, _lib.RC_CG_ROWS)}
    for step in range(steps):
        its = {}
        for f, s in sims.items():
            s.step(1)
            its[f] = list(s.cg_stats().last_iters)
        # step 11 holds the reference's cold restart (every 11th solve)
        print(workload, "step", step + 1, "iterations sites", its[_lib.RC_CG_SITES], "rows", its[_lib.RC_CG_ROWS],
              flush=True)
        assert its[_lib.RC_CG_SITES] == its[_lib.RC_CG_ROWS]
    (ps, vxs, vys), (pr, vxr, vyr) = _fields(sims[_lib.RC_CG_SITES]), _fields(sims[_lib.RC_CG_ROWS])
    vmax = float(max(np.abs(vxr).max(), np.abs(vyr).max()))
    dv = float(max(np.abs(vxs - vxr).max(), np.abs(vys - vyr).max()))
    dphi = float(np.abs(ps - pr).max())
    worst = max(s.cg_stats().worst_relres for s in sims.values())
    print(workload, "max|dphi| %.3g max|dv| %.3g (|v| %.3g), worst relres %.3g" % (dphi, dv, vmax, worst), flush=True)
    assert worst < 1e-8
    assert dv <= 1e-11 * vmax and dphi <= 1e-12


def test_site_form_matches_oracle_at_the_reference_tolerance():
    """Same iteration counts as the oracle's row-space CG at 1e-8, fields
    within the solver's tolerance of it."""
    conf, keys = confgen.workload("wet256_400", seed=8)
    sim = Simulation([(conf, keys)], cg_form=_lib.RC_CG_SITES)
    orc = OracleSim(conf, keys)
    for _ in range(3):
        sim.step(1)
        orc.step(1)
        si, oi = list(sim.cg_stats().last_iters), list(orc.cg_stats().last_iters)
        print("iterations library", si, "oracle", oi, flush=True)
        assert si == oi
    vmax = float(np.abs(orc.rows(0)).max())
    dv = float(max(np.abs(sim.rows(0) - orc.rows(0)).max(), np.abs(sim.rows(1) - orc.rows(1)).max()))
    print("max|dv| %.3g (|v| %.3g)" % (dv, vmax), flush=True)
    assert dv <= 1e-10 * vmax


def test_cg_form_argument():
    conf, keys = confgen.workload("wet256_400", seed=6)
    sim = Simulation([(conf, keys)], init=False)
    with pytest.raises(Exception, match="RC_ERR_ARG"):
        sim.set_cg_form(2)

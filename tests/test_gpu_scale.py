"""The benchmark workloads at their real sizes on the GPU against the oracle:
dry 512^2 with 1024 cells, the 32-replica ensemble (uneven boxes and
patches, ~11,700 cells in one handle), wet 256^2 with 400 cells at a tight
solver tolerance, a 100-step wet run, and wet 512^2 with 1600 cells driven by
the trained warm start.

Tolerance: max|dphi| and max|dv| <= 1e-10 against the oracle (FP64 path; the
wet cases solve to 1e-12 / 1e-13 so that both solvers' answers are closer
than that).  The learned case is compared with the reference warm start of
the library itself at the reference tolerance 1e-8."""
import os
import time

import numpy as np
import pytest

from oracle import OracleSim
from rheocell_amd import _lib, confgen, learned
from rheocell_amd.sim import Simulation

pytestmark = pytest.mark.gpu

TOL = 1e-10
WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "rheocell_amd", "weights",
                       "warmstart_wet256.npz")


def _compare(sim, orcs, check_v=True):
    """max |dphi| (and |dv|) of every cell of every box; geometry must agree."""
    cells, phi = sim.download()
    vx, vy = (sim.rows(0), sim.rows(1)) if check_v else (None, None)
    start, i, worst, worst_v = 0, 0, 0.0, 0.0
    for o in orcs:
        ov = (o.rows(0), o.rows(1)) if check_v else None
        r0 = start
        for j in range(o.N):
            c, oc = cells[i], o.cell(j)
            assert (c.LsubX, c.LsubY, c.offset[0], c.offset[1], c.corner) == \
                (oc.LsubX, oc.LsubY, oc.offset[0], oc.offset[1], oc.corner), (i, j)
            S = c.LsubX * c.LsubY
            worst = max(worst, float(np.abs(phi[start:start + S] - o.phi(j)).max()))
            start += S
            i += 1
        if check_v:
            worst_v = max(worst_v, float(np.abs(vx[r0:start] - ov[0]).max()), float(np.abs(vy[r0:start] - ov[1]).max()))
    return worst, worst_v


def test_dry512_1024_cells_matches_oracle():
    conf, keys = confgen.workload("dry512_1024", seed=2)
    sim = Simulation([(conf, keys)])
    orc = OracleSim(conf, keys)
    for n in (1, 4):
        sim.step(n)
        orc.step(n)
        w, wv = _compare(sim, [orc])
        print("dry512_1024 step", sim.step_count, "max|dphi| =", w, "max|dv| =", wv, flush=True)
        assert w <= TOL and wv <= TOL


def test_ensemble_of_32_replicas_matches_oracle():
    ens = confgen.replica_ensemble(32, seed=3)
    sim = Simulation(ens)
    groups = sim.stage_groups()
    print("ensemble: %d cells in %d boxes, stage groups %s" % (sim.ncells, len(ens),
          [(g["threads"], g["rows"], g["cells"]) for g in groups]), flush=True)
    assert sim.ncells > 10000 and len(groups) > 1
    orcs = [OracleSim(c, k) for c, k in ens]
    sim.step(3)
    for o in orcs:
        o.step(3)
    w, wv = _compare(sim, orcs)
    print("ensemble32 3 steps max|dphi| =", w, "max|dv| =", wv, flush=True)
    assert w <= TOL and wv <= TOL


def test_wet256_400_cells_matches_oracle():
    conf, keys = confgen.workload("wet256_400", seed=4)
    keys["CGtolerance"] = 1e-12
    sim = Simulation([(conf, keys)])
    orc = OracleSim(conf, keys)
    for _ in range(2):
        sim.step(1)
        orc.step(1)
        w, wv = _compare(sim, [orc])
        print("wet256_400 step", sim.step_count, "max|dphi| =", w, "max|dv| =", wv, "iters",
              list(sim.cg_stats().last_iters), list(orc.cg_stats().last_iters), flush=True)
        assert w <= TOL and wv <= TOL


def test_wet_100_steps_matches_oracle():
    conf = confgen.lattice_conf(64, 3, 7)
    keys = dict(confgen.BASE_KEYS, friction_cell=1, CGtolerance=1e-13, seed=7)
    sim = Simulation([(conf, keys)])
    orc = OracleSim(conf, keys)
    for n in (50, 50):
        sim.step(n)
        orc.step(n)
        w, wv = _compare(sim, [orc])
        print("wet64 step", sim.step_count, "max|dphi| =", w, "max|dv| =", wv, flush=True)
        assert w <= TOL and wv <= TOL


def test_wet512_1600_cells_learned_start():
    """Every solve reaches 1e-8; the fields agree with the reference warm
    start's to the solver tolerance; with the iteration over rows (Eigen's
    form) the learned start cuts CG time >= 2x."""
    conf, keys = confgen.workload("wet512_1600", seed=5)
    ref = Simulation([(conf, keys)], cg_form=_lib.RC_CG_ROWS)
    lrn = Simulation([(conf, keys)], cg_form=_lib.RC_CG_ROWS)
    w, b, scale = learned.load(WEIGHTS)
    learned.attach(lrn, w, b, scale, stages=8)
    ms = {"ref": 0.0, "lrn": 0.0}
    for step in range(24):
        for name, s in (("ref", ref), ("lrn", lrn)):
            t0 = s.cg_stats().total_solve_ms
            s.step(1)
            if step >= 4:        # past the cold start
                ms[name] += s.cg_stats().total_solve_ms - t0
    sr, sl = ref.cg_stats(), lrn.cg_stats()
    _, pr = ref.download()
    _, pl = lrn.download()
    dphi = float(np.abs(pr - pl).max())
    dv = float(max(np.abs(ref.rows(0) - lrn.rows(0)).max(), np.abs(ref.rows(1) - lrn.rows(1)).max()))
    vmax = float(np.abs(ref.rows(0)).max())
    print("wet512_1600: worst relres ref %.3g learned %.3g; CG ms/step ref %.3f learned %.3f (%.2fx); "
          "iterations %d vs %d; net kept %d/%d; max|dphi| %.3g max|dv| %.3g (|v| %.3g)"
          % (sr.worst_relres, sl.worst_relres, ms["ref"] / 20, ms["lrn"] / 20, ms["ref"] / ms["lrn"],
             sr.iters_total, sl.iters_total, sl.learned_used, sl.learned_tried, dphi, dv, vmax), flush=True)
    assert sr.worst_relres < 1e-8 and sl.worst_relres < 1e-8
    assert dphi < 5e-10 and dv < 2e-7 * vmax      # measured 2.6e-11 and 3.7e-8 |v|
    assert ms["ref"] >= 2.0 * ms["lrn"]

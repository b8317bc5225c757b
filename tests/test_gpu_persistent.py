"""The persistent form of the dry fused stage (one launch over every cell of
every box, k_stage_p) forced on, against the oracle: uneven patches in
several replicas (3 and 2 strip rows), 100 steps of the dry 140^2/64-cell
case, a patch growth, and HIP-graph replays against ordinary launches.

Tolerance: max|dphi| <= 1e-10 (FP64)."""
import numpy as np
import pytest

from oracle import OracleSim
from rheocell_amd import confgen
from rheocell_amd.conf import RCexception
from rheocell_amd.sim import Simulation
from tests.test_gpu_features import _growth_conf, _max_dphi

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _persistent(ens):
    sim = Simulation(ens)
    sim.set_stage_shape(1024, 0)
    info = sim.stage_persistent()
    assert info["eligible"] and info["used"], info
    return sim


def test_uneven_replicas():
    ens = []
    for r, (L, n) in enumerate([(96, 3), (128, 3), (160, 4)]):
        rng = np.random.default_rng(40 + r)
        sides = [int(s) for s in rng.choice([29, 37, 41, 45, 49], n * n)]
        radii = [float(x) for x in rng.uniform(6, 10, n * n)]
        conf = confgen.lattice_conf(L, n, seed=40 + r, radii=radii, sides=sides)
        ens.append((conf, dict(confgen.BASE_KEYS, seed=40 + r, zetaQ_self=0.05 + 0.05 * r)))
    sim = _persistent(ens)
    orcs = [OracleSim(c, k) for c, k in ens]
    for n in (1, 19):
        sim.step(n)
        for o in orcs:
            o.step(n)
        w = _max_dphi(sim, orcs)
        print("persistent, 3 replicas, step", sim.step_count, "max|dphi| =", w, flush=True)
        assert w <= TOL


def test_dry140_100_steps():
    conf, keys = confgen.workload("dry140_64", seed=3)
    sim = _persistent([(conf, keys)])
    orc = OracleSim(conf, keys)
    sim.step(100)
    orc.step(100)
    w = _max_dphi(sim, [orc])
    print("persistent, dry140_64, 100 steps: max|dphi| =", w, flush=True)
    assert w <= TOL


def test_growth():
    conf = _growth_conf()
    keys = dict(confgen.BASE_KEYS, seed=1)
    sim = _persistent([(conf, keys)])
    orc = OracleSim(conf, keys)
    sim.step(6)
    orc.step(6)
    cells, _ = sim.download()
    assert (cells[0].LsubX, cells[0].LsubY) == (orc.cell(0).LsubX, orc.cell(0).LsubY) != (41, 41)
    assert sim.growth_stats()[0] >= 1 and sim.stage_persistent()["used"]
    w = _max_dphi(sim, [orc])
    print("persistent, growth to %dx%d: max|dphi| = %.3e" % (cells[0].LsubX, cells[0].LsubY, w), flush=True)
    assert w <= TOL


def test_graph_replays_match_eager_steps():
    conf = confgen.lattice_conf(128, 4, seed=12)
    keys = dict(confgen.BASE_KEYS, seed=12)
    a = _persistent([(conf, keys)])
    b = _persistent([(conf, keys)])
    b.set_graphs(False)
    for chunk in (20, 33):
        a.step(chunk)
        b.step(chunk)
        assert np.array_equal(a.download()[1], b.download()[1])
    assert a.graph_stats()[0] > 0


def test_refused_where_it_does_not_apply():
    conf = confgen.lattice_conf(96, 3, seed=5)
    for extra in (dict(J_Q=0.3), dict(omega=0.01), dict(friction_cell=1)):
        sim = Simulation([(conf, dict(confgen.BASE_KEYS, seed=5, **extra))])
        assert not sim.stage_persistent()["eligible"]
        with pytest.raises(RCexception, match="persistent stage kernel"):
            sim.set_stage_shape(1024, 0)
        sim.close()

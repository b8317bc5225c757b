"""Several wet boxes in one handle: uneven boxes and patches, each box its own
reference solve (its own CG scalars, stop test and restart counter,
WetPolarModel.cpp:151-157, 200-233) in the same launches as the others.

Tolerance: max|dphi|, max|dv| <= 1e-10 against one oracle per box at a tight
solver tolerance (both solvers' answers closer than that); bit-identical to
one handle per box at the reference tolerance, through a cold restart."""
import numpy as np
import pytest

from oracle import OracleSim
from rheocell_amd import _lib, confgen
from rheocell_amd.sim import RCexception, Simulation

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _boxes(tol):
    """Three wet replicas: 64^2, 96^2, 128^2 with patch sides 29..49."""
    rng = np.random.default_rng(21)
    out = []
    for L, n, seed in ((64, 3, 11), (96, 4, 12), (128, 5, 13)):
        radii = rng.uniform(6, 10, n * n)
        sides = [min(49, max(29, int(round(5 * R - 1)) | 1)) for R in radii]
        conf = confgen.lattice_conf(L, n, seed, radii=[float(r) for r in radii], sides=sides)
        keys = dict(confgen.BASE_KEYS, friction_cell=1, CGtolerance=tol, seed=seed)
        out.append((conf, keys))
    return out


def test_wet_replicas_match_the_oracle():
    boxes = _boxes(1e-12)
    sim = Simulation(boxes)
    orcs = [OracleSim(c, k) for c, k in boxes]
    sim.step(3)
    for o in orcs:
        o.step(3)
    _, phi = sim.download()
    vx, vy = sim.rows(0), sim.rows(1)
    start, worst, worst_v = 0, 0.0, 0.0
    for o in orcs:
        n = sum(len(o.phi(j)) for j in range(o.N))
        ref = np.concatenate([o.phi(j) for j in range(o.N)])
        worst = max(worst, float(np.abs(phi[start:start + n] - ref).max()))
        worst_v = max(worst_v, float(np.abs(vx[start:start + n] - o.rows(0)).max()),
                      float(np.abs(vy[start:start + n] - o.rows(1)).max()))
        start += n
    print("3 wet replicas, 3 steps: max|dphi| %.3g max|dv| %.3g" % (worst, worst_v), flush=True)
    assert worst <= TOL and worst_v <= TOL


def test_wet_replicas_equal_one_handle_per_box():
    """At the reference tolerance, 12 steps (24 solves per box: the cold
    restart after 10 warm ones included): the batched handle's fields and
    iteration counts equal those of one handle per box, bit for bit."""
    boxes = _boxes(1e-8)
    sim = Simulation(boxes, init=False)
    singles = [Simulation([b], init=False) for b in boxes]
    for s in [sim] + singles:
        s.set_stage_shape(512, 8)      # the same field-update shape in every handle
        s.init_forces()
        s.apply_changes_after_equilibration()
    g = sim.gaussians(12)
    cols = np.cumsum([0] + [len(c.cells) for c, _ in boxes])
    sim.step(12, g)
    for i, s in enumerate(singles):
        s.step(12, np.ascontiguousarray(g[:, cols[i]:cols[i + 1]]))
    _, phi = sim.download()
    vx = sim.rows(0)
    start = 0
    its = 0
    for s in singles:
        _, p = s.download()
        n = len(p)
        assert np.array_equal(phi[start:start + n], p)
        assert np.array_equal(vx[start:start + n], s.rows(0))
        start += n
        its += s.cg_stats().iters_total
    st = sim.cg_stats()
    print("iterations: batched %d, one handle per box %d; worst relres %.3g" % (st.iters_total, its, st.worst_relres),
          flush=True)
    assert st.iters_total == its and st.worst_relres < 1e-8


def test_nematic_wet_replicas_match_the_oracle():
    """The reference's example model (interaction_type = wetmodel) in two
    boxes of one handle, each against its own oracle."""
    boxes = []
    for L, n, seed in ((64, 3, 31), (96, 4, 32)):
        conf = confgen.lattice_conf(L, n, seed)
        keys = dict(confgen.BASE_KEYS)
        keys.update(confgen.EXAMPLE_NEMATIC_KEYS)
        keys.update(CGtolerance=1e-12, seed=seed)
        boxes.append((conf, keys))
    sim = Simulation(boxes)
    orcs = [OracleSim(c, k) for c, k in boxes]
    sim.step(3)
    for o in orcs:
        o.step(3)
    _, phi = sim.download()
    vx = sim.rows(0)
    start, worst, worst_v = 0, 0.0, 0.0
    for o in orcs:
        ref = np.concatenate([o.phi(j) for j in range(o.N)])
        n = len(ref)
        worst = max(worst, float(np.abs(phi[start:start + n] - ref).max()))
        worst_v = max(worst_v, float(np.abs(vx[start:start + n] - o.rows(0)).max()))
        start += n
    print("2 nematic wet replicas, 3 steps: max|dphi| %.3g max|dvx| %.3g" % (worst, worst_v), flush=True)
    assert worst <= TOL and worst_v <= TOL


def test_row_form_and_learned_start_need_one_box():
    sim = Simulation(_boxes(1e-8), init=False)
    with pytest.raises(RCexception, match="RC_ERR_ARG"):
        sim.set_cg_form(_lib.RC_CG_ROWS)

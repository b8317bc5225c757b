"""GPU path vs the oracle across model options, batched uneven replicas,
invariants and the reference's error conditions."""
import math

import numpy as np
import pytest

from oracle import OracleSim
from rheocell_amd import confgen
from rheocell_amd.conf import CellConf, Conf, RCexception
from rheocell_amd.sim import Simulation

pytestmark = pytest.mark.gpu


def _max_dphi(sim, orcs):
    cells, phi = sim.download()
    start = 0
    worst = 0.0
    i = 0
    for o in orcs:
        for j in range(o.N):
            c = cells[i]
            S = c.LsubX * c.LsubY
            worst = max(worst, float(np.abs(phi[start:start + S] - o.phi(j)).max()))
            oc = o.cell(j)
            assert (c.offset[0], c.offset[1], c.corner) == (oc.offset[0], oc.offset[1], oc.corner)
            start += S
            i += 1
    return worst


@pytest.mark.parametrize("opts", [dict(J_Q=0.5), dict(omega=0.02), dict(beta=0.3), dict(anchoring=1),
                                  dict(J_Q=0.2, omega=0.01, beta=0.1, zetaQ_inter=0.3)],
                         ids=["JQ", "omega", "beta", "anchoring", "all"])
def test_dry_options(opts):
    conf = confgen.lattice_conf(96, 4, seed=11)
    keys = dict(confgen.BASE_KEYS, seed=11, **opts)
    sim = Simulation([(conf, keys)])
    orc = OracleSim(conf, keys)
    sim.step(20)
    orc.step(20)
    w = _max_dphi(sim, [orc])
    print(opts, "max|dphi| =", w)
    assert w <= 1e-10
    dv = np.abs(sim.rows(0) - orc.rows(0)).max()
    assert dv <= 1e-10


def test_uneven_replicas_one_launch():
    rng = np.random.default_rng(21)
    boxes, orcs = [], []
    for b, L in enumerate((96, 128, 64)):
        n = {96: 3, 128: 4, 64: 2}[L]
        sides = [int(s) for s in rng.choice([29, 33, 41, 49], n * n)]
        radii = [float(r) for r in rng.uniform(6, 10, n * n)]
        conf = confgen.lattice_conf(L, n, seed=100 + b, radii=radii, sides=sides)
        keys = dict(confgen.BASE_KEYS, seed=100 + b, zetaQ_self=float(rng.uniform(0.05, 0.2)))
        boxes.append((conf, keys))
        orcs.append(OracleSim(conf, keys))
    sim = Simulation(boxes)
    sim.step(15)
    for o in orcs:
        o.step(15)
    w = _max_dphi(sim, orcs)
    print("3 replicas, uneven patches: max|dphi| =", w)
    assert w <= 1e-10


def test_translation_invariance():
    conf = confgen.lattice_conf(128, 3, seed=8)
    sh = (37, 91)
    cells = []
    for c in conf.cells:
        cx, cy = c.corner % 128, c.corner // 128
        corner = (cx + sh[0]) % 128 + ((cy + sh[1]) % 128) * 128
        cells.append(CellConf(c.LsubX, c.LsubY, ((c.CoM[0] + sh[0]) % 128, (c.CoM[1] + sh[1]) % 128), c.offset,
                              corner, (c.unrap[0] + sh[0], c.unrap[1] + sh[1]), c.Q00, c.Q01, c.box_index, c.phi))
    keys = dict(confgen.BASE_KEYS, seed=8)
    a = Simulation([(conf, keys)])
    b = Simulation([(Conf(0, 128, 128, cells), keys)])
    a.step(12)
    b.step(12)
    assert np.array_equal(a.download()[1], b.download()[1])


def test_site_arrays_dry_and_wet():
    for fc in (0, 1):
        conf = confgen.lattice_conf(64, 3, seed=13)
        keys = dict(confgen.BASE_KEYS, seed=13, friction_cell=fc, CGtolerance=1e-13)
        sim = Simulation([(conf, keys)])
        orc = OracleSim(conf, keys)
        sim.step(3)
        orc.step(3)
        for which in (1, 2, 3, 4):
            ref = np.concatenate([orc.site(i, which) for i in range(orc.N)])
            got = sim.site_array(which)
            assert np.abs(got - ref).max() <= 1e-10, (fc, which)


def test_equilibration_then_active():
    conf = confgen.lattice_conf(64, 3, seed=17)
    keys = dict(confgen.BASE_KEYS, seed=17, zetaQ_self=0.3, J_Q=0.1)
    sim = Simulation([(conf, keys)], equilibrate=False)
    orc = OracleSim(conf, keys, equilibrate=False)
    sim.step(4)
    orc.step(4)
    sim.apply_changes_after_equilibration()
    orc.L.orc_apply_changes_after_equilibration(orc.h)
    sim.step(6)
    orc.step(6)
    assert _max_dphi(sim, [orc]) <= 1e-10


def test_wet_reference_tolerance():
    conf = confgen.lattice_conf(64, 3, seed=19)
    keys = dict(confgen.BASE_KEYS, seed=19, friction_cell=1, CGtolerance=1e-8)
    sim = Simulation([(conf, keys)])
    orc = OracleSim(conf, keys)
    for _ in range(12):  # crosses the cold restart after 10 warm solves
        sim.step(1)
        orc.step(1)
        gi, oi = list(sim.cg_stats().last_iters), list(orc.cg_stats().last_iters)
        assert all(abs(a - b) <= 2 for a, b in zip(gi, oi)), (gi, oi)
        assert max(sim.cg_stats().last_relres) < 1e-8
    w = _max_dphi(sim, [orc])
    print("wet @1e-8, 12 steps: max|dphi| =", w)
    assert w < 1e-7


def test_nan_field_halts():
    # one NaN in cell 4's field: the first step names the first cell whose field
    # it reached (through the box sum of phi^2 it reaches the cells overlapping
    # that site in the same launch) and halts before recentering
    conf = confgen.lattice_conf(64, 3, seed=7)
    conf.cells[4].phi = conf.cells[4].phi.copy()
    conf.cells[4].phi[100] = float("nan")
    keys = dict(confgen.BASE_KEYS, seed=7)
    with pytest.raises(RCexception, match=r"RC_ERR_NAN: the field of cell \d+ became NaN"):
        sim = Simulation([(conf, keys)])
        sim.step(2)
    # a clean configuration still runs on a new handle
    sim = Simulation([(confgen.lattice_conf(64, 3, seed=7), keys)])
    sim.step(1)
    sim.close()


def test_nan_in_wet_solve_halts():
    # a NaN cell friction: the first wet solve (the corrector's force
    # evaluation) has r.z = NaN at its first residual and stops there instead
    # of iterating to its cap (the cap is also set low, so the test cannot run
    # long either way)
    conf = confgen.lattice_conf(64, 3, seed=7)
    keys = dict(confgen.BASE_KEYS, seed=7, friction_cell=float("nan"), CGtolerance=1e-8)
    with pytest.raises(RCexception, match="RC_ERR_NAN: the wet solve's right-hand side or residual became NaN"):
        sim = Simulation([(conf, keys)], cg_max_iter=200)
        sim.step(1)


def test_api_argument_and_order_errors():
    # the C ABI's own checks: a bad box, a call out of order, a patch that does
    # not fit its box, an unknown warm-start mode
    import ctypes as C
    import numpy as np
    from rheocell_amd import _lib
    from rheocell_amd.sim import cell_from_conf, model_params
    L = _lib.load()
    conf = confgen.lattice_conf(64, 2, seed=3)
    keys = dict(confgen.BASE_KEYS, seed=3)
    desc = (_lib.rc_box_desc * 1)()
    desc[0].L, desc[0].ncells, desc[0].params = 0, len(conf.cells), model_params(keys)
    h = C.c_void_p()
    assert L.rc_create(desc, 1, C.byref(h)) == 1          # RC_ERR_ARG: side 0
    desc[0].L = 64
    assert L.rc_create(desc, 1, C.byref(h)) == 0
    try:
        assert L.rc_step(h, 1, None) == 6                  # RC_ERR_STATE: nothing uploaded or initialised
        assert L.rc_last_error(h) == b"rc_step before rc_init_forces"
        assert L.rc_init_forces(h) == 6
        assert L.rc_last_error(h) == b"rc_init_forces before rc_upload_cells"
        cells = (_lib.rc_cell * len(conf.cells))()
        for i, c in enumerate(conf.cells):
            cells[i] = cell_from_conf(c)
        cells[1].corner = 64 * 64 + 5                      # corner outside the box
        phi = np.ascontiguousarray(np.concatenate([c.phi for c in conf.cells]))
        ptr = phi.ctypes.data_as(C.POINTER(C.c_double))
        assert L.rc_upload_cells(h, cells, ptr) == 1       # RC_ERR_ARG
        assert b"corner" in L.rc_last_error(h)
        cells[1] = cell_from_conf(conf.cells[1])
        assert L.rc_upload_cells(h, cells, ptr) == 0
        assert L.rc_set_warm_start(h, 7) == 1              # RC_ERR_ARG: unknown mode
        assert L.rc_step(h, -1, None) in (1, 6)
        assert L.rc_init_forces(h) == 0
        assert L.rc_step(h, -1, None) == 1                 # RC_ERR_ARG: negative step count
        assert L.rc_step(h, 1, None) == 0
    finally:
        L.rc_destroy(h)


def test_overlap_limit_raises():
    # 21 patches covering the same sites: the reference throws
    L = 64
    cells = [confgen.tanh_cell(L, 32.0 + 0.01 * i, 32.0, 8.0, 2.0, 41, 0.1 * i) for i in range(21)]
    conf = Conf(0, L, L, cells)
    keys = dict(confgen.BASE_KEYS, seed=1)
    with pytest.raises(RuntimeError, match="Too many field patches overlap"):
        OracleSim(conf, keys)
    with pytest.raises(RCexception, match="RC_ERR_OVERLAP"):
        sim = Simulation([(conf, keys)])
        sim.step(1)


def _growth_conf(L=96):
    # a cell whose interface comes closer than 4 sites to the patch's left edge
    c = confgen.tanh_cell(L, 40.0, 48.0, 8.0, 2.0, 41, 0.2)
    phi = np.roll(c.phi.reshape(41, 41), -9, axis=1)
    phi[:, -9:] = 0.0
    cell = CellConf(41, 41, c.CoM, (0, 0), (c.corner + 9) % (L * L), (c.unrap[0] + 9, c.unrap[1]), c.Q00, c.Q01,
                    c.box_index, phi.ravel().copy())
    other = confgen.tanh_cell(L, 80.0, 20.0, 8.0, 2.0, 41, -1.0)
    return Conf(0, L, L, [cell, other])


@pytest.mark.parametrize("fc", [0, 1], ids=["dry", "wet"])
def test_patch_growth_parity(fc):
    conf = _growth_conf()
    keys = dict(confgen.BASE_KEYS, seed=1, friction_cell=fc, CGtolerance=1e-13, J_Q=0.3)
    orc = OracleSim(conf, keys)
    sim = Simulation([(conf, keys)])
    sim.step(1)
    orc.step(1)
    cells, _ = sim.download()
    assert (cells[0].LsubX, cells[0].LsubY) == (orc.cell(0).LsubX, orc.cell(0).LsubY) != (41, 41)
    sim.step(5)
    orc.step(5)
    w = _max_dphi(sim, [orc])
    print("growth (%s): patch %dx%d, max|dphi| = %.3e" % ("wet" if fc else "dry", cells[0].LsubX, cells[0].LsubY, w))
    assert w <= 1e-10
    assert np.abs(sim.rows(0) - orc.rows(0)).max() <= 1e-10


def test_wet_solve_is_deterministic():
    conf = confgen.lattice_conf(96, 5, seed=23)
    keys = dict(confgen.BASE_KEYS, seed=23, friction_cell=1, CGtolerance=1e-8)
    a = Simulation([(conf, keys)])
    b = Simulation([(conf, keys)])
    a.step(4)
    b.step(4)
    assert np.array_equal(a.download()[1], b.download()[1])
    assert np.array_equal(a.rows(0), b.rows(0))
    assert a.cg_stats().iters_total == b.cg_stats().iters_total


@pytest.mark.parametrize("fc", [0, 1], ids=["dry", "wet"])
def test_every_stage_shape_matches_oracle(fc):
    # uneven patches 29..49 (several stage groups): every (threads per cell,
    # strip rows) shape of the field-update kernel, forced in turn on the groups
    # it fits, against the oracle
    rng = np.random.default_rng(5)
    n = 3
    sides = [int(s) for s in rng.choice([29, 41, 49], n * n)]
    radii = [float(r) for r in rng.uniform(6, 10, n * n)]
    conf = confgen.lattice_conf(96, n, seed=31, radii=radii, sides=sides)
    keys = dict(confgen.BASE_KEYS, seed=31, friction_cell=fc, CGtolerance=1e-13)
    orc = OracleSim(conf, keys)
    orc.step(6)
    probe = Simulation([(conf, keys)])
    groups = probe.stage_groups()
    probe.close()
    print("groups", groups)
    assert len(groups) >= 2 and sum(g["cells"] for g in groups) == n * n
    shapes = sorted({v for g in groups for v in g["valid"]})
    assert len(shapes) >= 3 and all(t in (256, 512) for t, _ in shapes)
    for t, r in shapes:
        sim = Simulation([(conf, keys)])
        sim.set_stage_shape(t, r)
        sim.step(6)
        for g in sim.stage_groups():
            if (t, r) in g["valid"]:
                assert (g["threads"], g["rows"]) == (t, r)
        w = _max_dphi(sim, [orc])
        print("shape", t, "x", r, "max|dphi| =", w)
        assert w <= 1e-10
        sim.close()


def test_stage_autotuner_picks_a_timed_shape():
    conf = confgen.lattice_conf(128, 4, seed=3)
    sim = Simulation([(conf, dict(confgen.BASE_KEYS, seed=3))])
    sim.step(1)
    (g,) = sim.stage_groups()
    print(g)
    assert (g["threads"], g["rows"]) in g["valid"] and g["cells"] == 16
    times = g["tune_us"]
    assert all(v > 0 for v in times.values())
    assert times["%dx%d" % (g["threads"], g["rows"])] == min(times.values())
    with pytest.raises(RCexception):
        sim.set_stage_shape(256, 4)   # 4 rows per strip do not cover a 41-row patch with 256 threads


@pytest.mark.parametrize("grow", [False, True], ids=["steady", "growth"])
def test_graph_replays_match_eager_steps(grow):
    # dry steps replayed from the 16-step HIP graph against the same steps
    # launched one by one: identical fields after every chunk (several replays
    # with host synchronisation between them); with growth the chunk falls
    # back to ordinary launches after the growth is served
    if grow:
        rng = np.random.default_rng(21)
        sides = [int(s) for s in rng.choice([29, 33, 41, 49], 9)]
        radii = [float(r) for r in rng.uniform(6, 10, 9)]
        conf = confgen.lattice_conf(96, 3, seed=100, radii=radii, sides=sides)
        keys = dict(confgen.BASE_KEYS, seed=100)
    else:
        conf = confgen.lattice_conf(128, 4, seed=12)
        keys = dict(confgen.BASE_KEYS, seed=12)
    a = Simulation([(conf, keys)])
    b = Simulation([(conf, keys)])
    b.set_graphs(False)
    for chunk in (20, 33, 16):
        a.step(chunk)
        b.step(chunk)
        ca, pa = a.download()
        cb, pb = b.download()
        assert pa.shape == pb.shape and np.array_equal(pa, pb)
        assert [(c.corner, c.LsubX, c.offset[0], c.offset[1]) for c in ca] == \
               [(c.corner, c.LsubX, c.offset[0], c.offset[1]) for c in cb]
    ga, ea = a.graph_stats()
    gb, eb = b.graph_stats()
    print("graph replays", ga, "eager steps", ea, "| reference run eager steps", eb)
    # steps queued after a growth request halt and are launched again once it is served
    assert gb == 0 and (eb == 69 if not grow else eb >= 69)
    assert ga >= 1 and ga * 16 + ea >= 69


def test_graph_replays_resume_after_a_growth():
    # a patch grows in the first step; with 79 steps left the re-laid-out steps
    # are captured again and replayed: identical fields to ordinary launches
    conf = _growth_conf()
    keys = dict(confgen.BASE_KEYS, seed=1)
    a = Simulation([(conf, keys)])
    b = Simulation([(conf, keys)])
    b.set_graphs(False)
    a.step(80)
    b.step(80)
    ca, pa = a.download()
    cb, pb = b.download()
    assert pa.shape == pb.shape and np.array_equal(pa, pb)
    assert [(c.corner, c.LsubX, c.LsubY) for c in ca] == [(c.corner, c.LsubX, c.LsubY) for c in cb]
    (ga, ea), (n, _) = a.graph_stats(), a.growth_stats()
    print("growths", n, "graph replays", ga, "eager steps", ea)
    assert n >= 1 and ga >= 3


def test_tile_list_overflow_raises():
    # 136 patches of 3 x 3 sites (three covering each site, far below the
    # overlap limit of 20) all intersect the first 32 x 8 box tile: more than
    # its list holds (128), which the gather reports as RC_ERR_TILE
    from rheocell_amd.conf import CellConf, Conf
    L = 64
    cells = []
    for x0 in range(-2, 32):
        for y0 in (-2, 1, 4, 7):
            corner = (x0 % L) + (y0 % L) * L
            cells.append(CellConf(3, 3, ((x0 + 1.5) % L, (y0 + 1.5) % L), (0, 0), corner, (float(x0), float(y0)),
                                  1.0, 0.0, np.zeros(9, dtype=np.int64), np.full(9, 0.5)))
    assert len(cells) == 136
    sim = Simulation([(Conf(0, L, L, cells), dict(confgen.BASE_KEYS, seed=2))])
    with pytest.raises(RCexception, match="RC_ERR_TILE"):
        sim.step(1)
    sim.close()

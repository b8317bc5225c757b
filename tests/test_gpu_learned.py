"""The learned warm start: BF16-MFMA conv kernels against the CPU oracle,
GPU training, and the CG driven by the trained network."""
import numpy as np
import pytest

from oracle import conv_ref
from oracle import OracleSim
from rheocell_amd import _lib, confgen, learned
from rheocell_amd.sim import Simulation

pytestmark = pytest.mark.gpu


def _data(seed=0, N=2, H=32, W=48):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((N, H, W, 3)).astype(np.float32)
    y = rng.standard_normal((N, H, W, 2)).astype(np.float32)
    return x, y


def test_forward_matches_oracle():
    x, _ = _data()
    w, b = conv_ref.init_params(5)
    b = (b + 0.05 * conv_ref.masks()[1]).astype(np.float32)
    net = learned.ConvNet(2, 32, 48)
    net.set_params(w, b)
    got = net.forward(x)
    _, ref = conv_ref.forward(x.astype(np.float64), w.astype(np.float64), b.astype(np.float64), bf=True)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("forward rel err", err)
    assert err < 8e-3          # measured 1.6e-3 (BF16 operands, FP32 accumulation)


def test_gradients_match_oracle():
    x, y = _data(1)
    w, b = conv_ref.init_params(6)
    net = learned.ConvNet(2, 32, 48)
    net.set_params(w, b)
    loss, gw, gb = net.grads(x, y)
    rl, rw, rb = conv_ref.loss_and_grads(x.astype(np.float64), y.astype(np.float64), w.astype(np.float64),
                                         b.astype(np.float64), bf=True)
    print("loss", loss, rl)
    assert abs(loss - rl) <= 1e-4 * rl      # measured 2.4e-6
    for l in range(6):
        e = np.linalg.norm(gw[l] - rw[l]) / np.linalg.norm(rw[l])
        eb = np.linalg.norm(gb[l] - rb[l]) / max(np.linalg.norm(rb[l]), 1e-30)
        print("layer", l, "rel err w", e, "b", eb)
        assert e < 3e-3 and eb < 3e-3     # measured <= 5.8e-4 (layer 0)
    mw, mb = conv_ref.masks()
    assert np.all(gw[mw == 0] == 0) and np.all(gb[mb == 0] == 0)


def test_training_reduces_loss():
    # learn a fixed smooth linear map of the inputs
    rng = np.random.default_rng(2)
    x = rng.standard_normal((8, 32, 32, 3)).astype(np.float32)
    xs = x + np.roll(x, 1, axis=1) + np.roll(x, -1, axis=2)
    y = np.stack([xs[..., 0] - 0.5 * xs[..., 2], 0.3 * xs[..., 1]], axis=-1).astype(np.float32)
    net = learned.ConvNet(8, 32, 32)
    net.set_params(*learned.init_params(0))
    losses = [net.train_step(x, y, 3e-3) for _ in range(60)]
    print("loss", losses[0], "->", losses[-1])
    assert losses[-1] < 0.2 * losses[0]


def test_learned_warm_start_end_to_end():
    # recorded solves of small wet runs -> brief training -> CG started from the net
    L = 64
    runs = learned.generate_snapshots(L=L, nside=4, seeds=(3, 4), steps=40, burn_in=5)
    assert all(r.shape == (80, L, L, 5) for r in runs)
    # each record: residual sums, sum phi^2, correction sums; sum phi^2 > 0 inside cells
    assert all(np.isfinite(r).all() and r[..., 2].max() > 0.5 for r in runs)
    w, b, scale, hist = learned.train(runs, epochs=20, batch=8, lr=3e-3)
    print("training loss", hist[0], "->", hist[-1])
    assert hist[-1] < hist[0]
    conf = confgen.lattice_conf(L, 4, 9)
    tight = dict(confgen.BASE_KEYS, friction_cell=1, CGtolerance=1e-12, seed=9)
    a = Simulation([(conf, tight)])
    a.set_warm_start(_lib.RC_WARM_PREVIOUS)
    bsim = Simulation([(conf, tight)])
    learned.attach(bsim, w, b, scale, stages=3)
    for _ in range(8):
        a.step(1)
        bsim.step(1)
        assert max(bsim.cg_stats().last_relres) < 1e-12
    sa, sb = a.cg_stats(), bsim.cg_stats()
    print("iterations previous-guess", sa.iters_total, "learned", sb.iters_total,
          "net step taken", sb.learned_used, "/", sb.learned_tried)
    assert sb.learned_tried == 16
    # both runs solve tightly: the fields agree to the solver tolerance
    _, pa = a.download()
    _, pb = bsim.download()
    assert np.abs(pa - pb).max() < 1e-9
    assert np.abs(a.rows(0) - bsim.rows(0)).max() < 1e-9


def test_stage_records_and_limits():
    # with the network attached, every stage's start is one training record
    L = 64
    w, b = learned.init_params(1)
    scale = np.array([1.0, 1.0, 0.01], dtype=np.float32)
    recs = learned.record_run(L, 4, 5, steps=3, burn_in=2, params=(w, b, scale), stages=3)
    assert recs.shape == (2 * 3 * 3, L, L, 5) and np.isfinite(recs).all()
    # the scaled residual sums of a start have unit-order RMS
    rms = np.sqrt(np.mean(recs[..., 0:2] ** 2, axis=(1, 2)))
    assert np.all(rms > 0.1) and np.all(rms < 10)
    conf = confgen.lattice_conf(L, 4, 5)
    sim = Simulation([(conf, dict(confgen.BASE_KEYS, friction_cell=1, seed=5))])
    with pytest.raises(Exception):
        learned.attach(sim, w, b, scale, stages=_lib_max_stages() + 1)


def _lib_max_stages():
    return 8


def test_trained_weights_cut_cg_iterations():
    # the committed network (trained on 256^2 / 400-cell runs, seeds 101..) on an
    # unseen seed of the same workload: 8 stages per solve, CG to 1e-8
    import os
    path = os.path.join(os.path.dirname(learned.__file__), "weights", "warmstart_wet256.npz")
    w, b, scale = learned.load(path)
    stages = learned.trained_stages(path)
    assert stages == 8
    conf = confgen.lattice_conf(256, 20, 77)
    keys = dict(confgen.BASE_KEYS, friction_cell=1, CGtolerance=1e-8, seed=77)
    it = {}
    for mode in ("previous", "learned"):
        sim = Simulation([(conf, keys)])
        g = sim.gaussians(8)
        if mode == "learned":
            learned.attach(sim, w, b, scale, stages=stages)
        else:
            sim.set_warm_start(_lib.RC_WARM_PREVIOUS)
        sim.step(3, g[:3])
        s0 = sim.cg_stats()
        sim.step(5, g[3:])
        s1 = sim.cg_stats()
        it[mode] = (s1.iters_total - s0.iters_total) / (s1.solves - s0.solves)
        assert max(s1.last_relres) < 1e-8
        sim.close()
    print("CG iterations per solve: previous", it["previous"], "learned", it["learned"])
    # each network stage costs about 2.3 CG iterations of device time; 8 stages
    # must leave fewer than previous - 8 * 2.3 iterations to pay off twice over
    assert it["learned"] < 0.25 * it["previous"]


def test_untrained_network_falls_back_safely():
    # a random network: the A-norm steps are small or useless, the fallback keeps
    # the previous solution where the residual got larger, and CG still converges
    # to the same solution as without the network
    L = 64
    conf = confgen.lattice_conf(L, 4, 21)
    tight = dict(confgen.BASE_KEYS, friction_cell=1, CGtolerance=1e-12, seed=21)
    a = Simulation([(conf, tight)])
    a.set_warm_start(_lib.RC_WARM_PREVIOUS)
    bsim = Simulation([(conf, tight)])
    w, b = learned.init_params(3)
    learned.attach(bsim, w, b, np.array([1.0, 1.0, 1.0], dtype=np.float32), stages=2)
    for _ in range(4):
        a.step(1)
        bsim.step(1)
        assert max(bsim.cg_stats().last_relres) < 1e-12
    st = bsim.cg_stats()
    print("kept", st.learned_used, "of", st.learned_tried)
    assert st.learned_tried == 8 and 0 <= st.learned_used <= 8
    _, pa = a.download()
    _, pb = bsim.download()
    assert np.abs(pa - pb).max() < 1e-9

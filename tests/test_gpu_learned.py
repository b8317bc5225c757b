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
    assert err < 2e-2


def test_gradients_match_oracle():
    x, y = _data(1)
    w, b = conv_ref.init_params(6)
    net = learned.ConvNet(2, 32, 48)
    net.set_params(w, b)
    loss, gw, gb = net.grads(x, y)
    rl, rw, rb = conv_ref.loss_and_grads(x.astype(np.float64), y.astype(np.float64), w.astype(np.float64),
                                         b.astype(np.float64), bf=True)
    print("loss", loss, rl)
    assert abs(loss - rl) <= 1e-3 * rl
    for l in range(6):
        e = np.linalg.norm(gw[l] - rw[l]) / np.linalg.norm(rw[l])
        eb = np.linalg.norm(gb[l] - rb[l]) / max(np.linalg.norm(rb[l]), 1e-30)
        print("layer", l, "rel err w", e, "b", eb)
        assert e < 3e-2 and eb < 3e-2
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
    learned.attach(bsim, w, b, scale)
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

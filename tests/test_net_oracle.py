"""The network's CPU oracle (oracle/conv_ref.py) pinned by finite
differences in float64 and by the circular-convolution adjoint identity."""
import numpy as np

from oracle import conv_ref


def _small(seed=0, H=6, W=5):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((2, H, W, 3))
    y = rng.standard_normal((2, H, W, 2))
    w, b = conv_ref.init_params(seed)
    b = b + 0.1 * rng.standard_normal(b.shape) * conv_ref.masks()[1]
    return x, y, w.astype(np.float64), b.astype(np.float64)


def test_conv_adjoint_identity():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((2, 7, 6, 4))
    g = rng.standard_normal((2, 7, 6, 5))
    w = rng.standard_normal((9, 5, 4))
    lhs = np.sum(conv_ref.conv(a, w) * g)
    assert np.isclose(lhs, np.sum(a * conv_ref.conv_t(g, w)), rtol=1e-12)
    assert np.isclose(lhs, np.sum(w * conv_ref.conv_w(a, g)), rtol=1e-12)


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(4)
    a = rng.standard_normal((1, 5, 4, 2))
    w = rng.standard_normal((9, 3, 2))
    out = conv_ref.conv(a, w)
    for (yy, xx, co) in [(0, 0, 0), (4, 3, 2), (2, 1, 1)]:
        s = 0.0
        for t in range(9):
            dy, dx = t // 3 - 1, t % 3 - 1
            s += a[0, (yy + dy) % 5, (xx + dx) % 4] @ w[t, co]
        assert np.isclose(out[0, yy, xx, co], s)


def test_gradients_by_finite_differences():
    x, y, w, b = _small()
    loss, gw, gb = conv_ref.loss_and_grads(x, y, w, b, bf=False)
    rng = np.random.default_rng(1)
    mw, mb = conv_ref.masks()
    idx = [tuple(int(rng.integers(s)) for s in w.shape) for _ in range(40)]
    idx = [i for i in idx if mw[i] > 0][:20]
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (conv_ref.loss_and_grads(x, y, wp, b, bf=False)[0] - conv_ref.loss_and_grads(x, y, wm, b, bf=False)[0]) / (2 * h)
        assert abs(fd - gw[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, gw[i])
    for l in (0, 5):
        for c in (0, 1):
            bp, bm = b.copy(), b.copy()
            bp[l, c] += h
            bm[l, c] -= h
            fd = (conv_ref.loss_and_grads(x, y, w, bp, bf=False)[0] - conv_ref.loss_and_grads(x, y, w, bm, bf=False)[0]) / (2 * h)
            assert abs(fd - gb[l, c]) <= 1e-6 * max(1.0, abs(fd))


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.005859375, -3.0e-3, 65504.0])
    r = conv_ref.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # halfway rounds to even
    assert r[2] == 1.0078125
    assert np.all(np.abs(r - v) <= np.abs(v) * 2 ** -8)

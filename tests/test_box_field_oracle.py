"""Pins oracle/box_field.py (the VelocityField / ShapeField box fields) by facts
that do not depend on its own formulas: hand-placed patches, the identity and
delta cases of the coarse average, totals."""
import numpy as np

from oracle.box_field import box_fields, coarse, site_box_index


def test_hand_placed_patch_with_offsets():
    # 3 x 3 patch, corner at box (3, 1) in a 5 x 5 box, ring offsets (1, 2):
    # storage site q = (qx, qy) is the logical site ((qx + 1) % 3, (qy + 2) % 3),
    # placed at x = (3 + lx) % 5, y = 1 + ly
    k = site_box_index(5, 3, 3, corner=1 * 5 + 3, offx=1, offy=2)
    by_hand = []
    for qy in range(3):
        for qx in range(3):
            lx, ly = (qx + 1) % 3, (qy + 2) % 3
            by_hand.append(((1 + ly) % 5) * 5 + (3 + lx) % 5)
    assert k.tolist() == by_hand
    # q = 0 -> logical (1, 2) -> box (4, 3); q = 2 -> logical (0, 2) -> box (3, 3) ... wraps in x
    assert k[0] == 3 * 5 + 4 and k[2] == 3 * 5 + 3 and k[1] == 3 * 5 + 0


def test_fields_weights_and_totals():
    rng = np.random.default_rng(4)
    L = 7
    geom = [(3, 3, 0, 0, 0), (4, 2, 5 * 7 + 5, 1, 1), (3, 4, 2 * 7 + 6, 2, 3)]
    phis = [rng.uniform(0, 1, lx * ly) for lx, ly, *_ in geom]
    rows = [rng.normal(size=(p.size, 2)) for p in phis]
    fx, fy = box_fields(L, geom, phis, rows)
    assert np.isclose(fx.sum(), sum((r[:, 0] * p).sum() for r, p in zip(rows, phis)), rtol=1e-13)
    assert np.isclose(fy.sum(), sum((r[:, 1] * p).sum() for r, p in zip(rows, phis)), rtol=1e-13)
    # support: exactly the union of the patches' sites
    cover = np.zeros(L * L, bool)
    for g in geom:
        cover[site_box_index(L, *g)] = True
    assert (fx.ravel() != 0).tolist() == cover.tolist()
    # a per-field pair weights every site of that field alike
    sx, sy = box_fields(L, geom[:1], phis[:1], [(2.0, -3.0)])
    k = site_box_index(L, *geom[0])
    assert np.array_equal(sx.ravel()[k], 2.0 * phis[0]) and np.array_equal(sy.ravel()[k], -3.0 * phis[0])
    # two fields on one site add in field order (sum of the two products)
    one = [(1, 1, 10, 0, 0), (1, 1, 10, 0, 0)]
    f2, _ = box_fields(L, one, [np.array([0.5]), np.array([0.25])], [(1.0, 0.0), (3.0, 0.0)])
    assert f2.ravel()[10] == 0.5 * 1.0 + 0.25 * 3.0


def test_coarse_identity_delta_constant():
    rng = np.random.default_rng(2)
    f = rng.normal(size=(6, 6))
    assert np.array_equal(coarse(f, 1), f)
    for g in (2, 3, 4):
        d = np.zeros((9, 9))
        d[4, 2] = 1.0
        c = coarse(d, g)
        h = g // 2
        # site (i, j) reads (i + k - h, j + l - h): the delta at (4, 2) reaches
        # rows 4 + h - (g - 1) .. 4 + h and columns 2 + h - (g - 1) .. 2 + h
        block = np.zeros((9, 9), bool)
        for k in range(g):
            for l in range(g):
                block[(4 - k + h) % 9, (2 - l + h) % 9] = True
        assert np.array_equal(c != 0, block)
        assert np.all(c[block] == 1.0 / (g * g))
        assert np.allclose(coarse(np.full((5, 5), 2.5), g), 2.5, rtol=0, atol=1e-15)

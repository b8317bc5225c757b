"""The oracle (and the product's host-side restatements) against golden
vectors produced by the reference's own MultiPhaseField / SquareBox / Utils
code compiled from source (tests/golden/make_golden.py)."""
import json
import os

import numpy as np
import pytest

from oracle import pyoracle
from rheocell_amd import rng as prng

HERE = os.path.dirname(os.path.abspath(__file__))
G = json.load(open(os.path.join(HERE, "golden", "ref_fields.json")))


@pytest.mark.parametrize("rec", G["trig"], ids=lambda r: "L%d" % r["L"])
def test_trig_tables_exact(rec):
    L = rec["L"]
    assert np.array_equal(pyoracle.trig_table(L, 0), np.array(rec["cos"]))
    assert np.array_equal(pyoracle.trig_table(L, 1), np.array(rec["sin"]))


def test_trig_tables_release_build_bound():
    # the reference's -O3 -ffast-math build evaluates the table with
    # vectorised libm: the IEEE table is within 5e-7 of it
    worst = 0.0
    for rec in G["trig_fast"]:
        worst = max(worst, np.abs(pyoracle.trig_table(rec["L"], 0) - np.array(rec["cos"])).max(),
                    np.abs(pyoracle.trig_table(rec["L"], 1) - np.array(rec["sin"])).max())
    assert worst < 5e-7


@pytest.mark.parametrize("rec", G["rng"], ids=lambda r: "seed%d" % r["seed"])
def test_drand48(rec):
    n = len(rec["values"])
    assert np.array_equal(pyoracle.rng_stream(rec["seed"], n, "drand48"), np.array(rec["values"]))
    r = prng.Drand48(rec["seed"])
    assert [r.drand48() for _ in range(n)] == rec["values"]


@pytest.mark.parametrize("rec", G["gauss"], ids=lambda r: "seed%d" % r["seed"])
def test_gaussian(rec):
    n = len(rec["values"])
    assert np.array_equal(pyoracle.rng_stream(rec["seed"], n, "gaussian"), np.array(rec["values"]))
    r = prng.Drand48(rec["seed"])
    assert [r.gaussian() for _ in range(n)] == rec["values"]


@pytest.mark.parametrize("rec", G["nbrsub"], ids=lambda r: "%dx%d" % (r["LX"], r["LY"]))
def test_patch_neighbours(rec):
    assert np.array_equal(pyoracle.test_nbrs(rec["LX"], rec["LY"]), np.array(rec["nbr"], dtype=np.int32))


@pytest.mark.parametrize("rec", G["subindex"], ids=lambda r: "-".join(map(str, r["args"])))
def test_ring_buffer_map(rec):
    Lb, LX, LY, ox, oy, corner = rec["args"]
    from rheocell_amd.conf import CellConf, Conf
    from oracle import OracleSim
    from rheocell_amd import confgen
    S = LX * LY
    c = CellConf(LX, LY, (LX / 2, LY / 2), (ox, oy), corner, (0.0, 0.0), 1.0, 0.0, np.zeros(S, dtype=np.int64),
                 np.zeros(S))
    sim = OracleSim(Conf(0, Lb, Lb, [c]), dict(confgen.BASE_KEYS), init=False)
    assert np.array_equal(sim.site(0, 8).astype(np.int64), np.array(rec["k"]))
    # the product's host conf writer computes the same box indices
    q = np.arange(S)
    lx = (q % LX + ox) % LX
    ly = (q // LX + oy) % LY
    k = (corner % Lb + lx) % Lb + ((corner // Lb + ly) % Lb) * Lb
    assert np.array_equal(k, np.array(rec["k"]))


@pytest.mark.parametrize("idx", range(len(G["setpos"])))
def test_set_positions(idx):
    rec = G["setpos"][idx]
    i = rec["input"]
    out = pyoracle.test_set_positions(i["Lbox"], i["LX"], i["LY"], i["offset"][0], i["offset"][1], i["corner"],
                                      i["unrap"][0], i["unrap"][1], i["CoM"][0], i["CoM"][1], i["xsl"], i["ysb"],
                                      np.array(i["phi"]))
    assert (out["LX"], out["LY"]) == (rec["LX"], rec["LY"])
    assert out["offset"] == rec["offset"]
    assert out["corner"] == rec["corner"]
    assert out["corner_old"] == rec["corner_old"]
    assert out["unrap"] == rec["unrap"]
    assert out["CoM"] == rec["CoM"]
    assert out["area"] == rec["area"]
    assert np.array_equal(out["phi"], np.array(rec["phi"]))

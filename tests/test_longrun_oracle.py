"""The long-run statistics criterion (tests/longrun.py) on the oracle alone:
two seeds of the same dynamics agree within Z_MAX combined standard errors,
and planted mistakes in the oracle's parameters (active force 30% too
strong, the repulsion term dropped) are caught by the centre-of-mass speed.
The GPU tests in test_gpu_statistics.py apply the same criterion to the
library against the oracle."""
import numpy as np
import pytest

from oracle import OracleSim
from rheocell_amd import confgen
from tests.longrun import STATS, mean_and_error, oracle_series, z_scores

Z_MAX = 5.0
STEPS, EVERY, SKIP, DT = 600, 20, 100, 0.1


def _run(seed, **change):
    conf, keys = confgen.workload("dry140_64", seed=seed)
    keys.update(change)
    orc = OracleSim(conf, keys)
    return oracle_series(orc, len(conf.cells), conf.Lx, STEPS, EVERY, SKIP, DT)


def test_block_error_on_known_series():
    # independent draws: the block error is sigma / sqrt(n) within sampling
    # error; a mean shifted by 10 of those errors gives z near 10; a series
    # against itself gives z = 0
    rng = np.random.default_rng(0)
    n, sigma = 4000, 2.0
    a = rng.normal(0.0, sigma, size=(n, 3))
    m, e = mean_and_error(a, blocks=20)
    assert np.allclose(e, sigma / np.sqrt(n), rtol=0.5)
    assert np.all(np.abs(m) < 5 * sigma / np.sqrt(n))
    b = rng.normal(0.0, sigma, size=(n, 3)) + 10 * np.sqrt(2) * sigma / np.sqrt(n)
    z = z_scores(a, b, blocks=20)
    assert np.all((z > 5) & (z < 15)), z
    assert np.all(z_scores(a, a) == 0)


@pytest.fixture(scope="module")
def base():
    return _run(3)


def test_independent_seeds_agree(base):
    z = z_scores(base, _run(4))
    assert (z < Z_MAX).all(), dict(zip(STATS, z))


@pytest.mark.parametrize("change", [dict(zetaQ_self=0.13), dict(kappa=0.0)], ids=["active_x1.3", "no_repulsion"])
def test_planted_mistake_is_caught(base, change):
    z = z_scores(base, _run(3, **change))
    assert z[STATS.index("speed")] > Z_MAX, dict(zip(STATS, z))

"""Reconstruction-quality metrics: the oracle's numpy restatement pinned
against the reference's own MetricsFramework / fast_rmsd (fixture
tests/golden/metrics_ref.json, written by tests/golden/make_anneal_fixtures.py
from the reference compiled in place) and against closed forms; the
product's C++ metrics (libpmmc_hip.so, host code) against both.

Tolerances: Pearson / SCC / distance correlation are FP64 sums rounded to
float, so two implementations may differ by the order of summation: 2e-6
absolute.  RMSD: 1e-5 relative (+1e-6 absolute), the float rounding of an
FP64 eigenvalue / singular-value computation."""
import json
import os

import numpy as np
import pytest
from scipy.spatial.transform import Rotation
from scipy.stats import pearsonr

from oracle import metrics as OM
from pmmc_amd import metrics as PM

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "metrics_ref.json")))["metrics"]


def _rmsd_close(a, b):
    return abs(a - b) <= 1e-5 * abs(b) + 1e-6


@pytest.mark.parametrize("impl", [OM, PM], ids=["oracle", "product"])
@pytest.mark.parametrize("k", range(len(CASES)))
def test_metrics_match_reference_fixture(impl, k):
    c = CASES[k]
    a = np.array(c["a"], np.float32).reshape(-1, 3)
    b = np.array(c["b"], np.float32).reshape(-1, 3)
    ha, hb = OM.distance_map(a), OM.distance_map(b)
    assert abs(impl.pearson(ha, hb) - c["map_pearson"]) <= 2e-6
    assert abs(impl.distance_correlation(a, b, c["step"]) - c["distance_correlation"]) <= 2e-6
    assert abs(impl.scc(ha, hb, c["h"]) - c["scc"]) <= 2e-6
    if np.isnan(c["rmsd"]):
        # the reference's fast_rmsd solves its cubic without guards and returns
        # NaN for coplanar sets (2 and 3 beads); here the least RMSD is
        # returned, checked against scipy's optimal proper rotation
        assert c["n"] <= 3
        _, rssd = Rotation.align_vectors(a.astype(np.float64) - a.mean(0),
                                         b.astype(np.float64) - b.mean(0))
        assert _rmsd_close(impl.rmsd(a, b), rssd / np.sqrt(c["n"]))
    else:
        assert _rmsd_close(impl.rmsd(a, b), c["rmsd"])


@pytest.mark.parametrize("impl", [OM, PM], ids=["oracle", "product"])
def test_pearson_closed_forms(impl):
    rng = np.random.default_rng(1)
    a = rng.normal(size=500).astype(np.float32)
    assert abs(impl.pearson(a, 2 * a + 1) - 1.0) <= 1e-6
    assert abs(impl.pearson(a, -3 * a) + 1.0) <= 1e-6
    assert impl.pearson(a, np.full_like(a, 2.0)) == 0.0  # degenerate denominator
    assert impl.pearson(a[:0], a[:0]) == 0.0
    b = rng.normal(size=500).astype(np.float32)
    assert abs(impl.pearson(a, b) - pearsonr(a.astype(np.float64), b.astype(np.float64))[0]) <= 2e-6


@pytest.mark.parametrize("impl", [OM, PM], ids=["oracle", "product"])
def test_rmsd_closed_forms(impl):
    rng = np.random.default_rng(2)
    x = (rng.normal(size=(50, 3)) * 5).astype(np.float32)
    R = Rotation.random(random_state=3).as_matrix()
    y = (x.astype(np.float64) @ R.T + np.array([3.0, -1.0, 7.0])).astype(np.float32)
    assert impl.rmsd(x, y) <= 1e-4                  # rigid motion: zero
    assert impl.rmsd(x, x) <= 1e-5
    # a mirror image is not superposable: compare with scipy's optimal proper
    # rotation on the centred sets
    m = x * np.array([1, 1, -1], np.float32)
    xc = x.astype(np.float64) - x.mean(0)
    mc = m.astype(np.float64) - m.mean(0)
    rot, rssd = Rotation.align_vectors(xc, mc)
    want = rssd / np.sqrt(len(x))
    assert abs(impl.rmsd(x, m) - want) <= 1e-5 * want + 1e-6
    # pure translation plus noise of known size on one bead
    z = x.copy()
    z[0] += np.array([0.0, 0.0, 0.001], np.float32)
    assert impl.rmsd(x, z) <= 0.001


@pytest.mark.parametrize("impl", [OM, PM], ids=["oracle", "product"])
def test_distance_correlation_and_scc_closed_forms(impl):
    rng = np.random.default_rng(4)
    x = rng.normal(size=(80, 3)).astype(np.float32)
    assert abs(impl.distance_correlation(x, x, 1) - 1.0) <= 1e-6
    assert abs(impl.distance_correlation(x, 4 * x, 1) - 1.0) <= 1e-6   # scale invariant
    h = OM.distance_map(x)
    assert abs(impl.scc(h, h, 5) - 1.0) <= 1e-6
    assert abs(impl.scc(h, 2 * h + 3, 5) - 1.0) <= 1e-6   # per diagonal affine invariant
    assert impl.scc(h[:1, :1], h[:1, :1], 5) == 0.0


def test_distance_correlation_subsampling_rule():
    """step <= 0 picks BenchmarkRunner's n/500 subsampling for n > 500."""
    rng = np.random.default_rng(5)
    x = rng.normal(size=(1100, 3)).astype(np.float32)
    y = (x + 0.3 * rng.normal(size=x.shape)).astype(np.float32)
    for impl in (OM, PM):
        assert impl.distance_correlation(x, y, 0) == impl.distance_correlation(x, y, 2)
    assert abs(PM.distance_correlation(x, y, 0) - OM.distance_correlation(x, y, 0)) <= 2e-6


@pytest.mark.parametrize("seed", range(6))
def test_product_metrics_match_oracle(seed):
    rng = np.random.default_rng(10 + seed)
    n = int(rng.integers(3, 300))
    a = (rng.normal(size=(n, 3)) * rng.uniform(0.1, 50)).astype(np.float32)
    b = (a + rng.normal(size=(n, 3)) * rng.uniform(0.01, 5)).astype(np.float32)
    ha, hb = PM.distance_map(a), PM.distance_map(b)
    assert np.array_equal(ha, OM.distance_map(a))
    assert abs(PM.pearson(ha, hb) - OM.pearson(ha, hb)) <= 2e-6
    assert abs(PM.distance_correlation(a, b) - OM.distance_correlation(a, b)) <= 2e-6
    assert abs(PM.scc(ha, hb) - OM.scc(ha, hb)) <= 2e-6
    assert _rmsd_close(PM.rmsd(a, b), OM.rmsd(a, b))

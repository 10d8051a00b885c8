"""The oracle's stray-lipid cleanup (delete_lipid, src/cleanup.h:29-91):
pinned by the reference's own delete_lipid (golden vectors) and by hand-made
cells whose medians and thresholds are known exactly."""
import os

import numpy as np
import pytest

from oracle import Oracle, REF_SO, Reference

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def test_cleanup_matches_reference_golden(orc):
    g = np.load(os.path.join(HERE, "golden", "stray.npz"))
    x = g["x"].astype(np.float64)
    dele = orc.stray_delete(x, np.zeros(len(x), np.int32), g["aff"], g["centroids"], float(g["tol"]))
    assert 0 < dele.sum() < len(x)
    assert np.array_equal(dele, ~g["keep"])


def _cell(center, d2s):
    # points at squared distances d2s from center (along x, exact in FP64 for these values)
    return [(center[0] + np.sqrt(v), center[1], center[2]) for v in d2s]


def test_cleanup_known_answers(orc):
    pts, aff = [], []
    # cell 0, odd size 11: sorted d2 [1,1,2,2,3,3,4,4,18.7,18.75,100] -> dr2[5] = 3, threshold 3 * 2.5^2 = 18.75:
    # 18.7 stays, 18.75 (not below) and 100 go
    c0 = (0.0, 0.0, 0.0)
    d0 = [1, 1, 2, 2, 3, 3, 4, 4, 18.7, 18.75, 100]
    # cell 1, even size 6: dr2[3] = 20 -> threshold 125: 100 stays (a lower median, 10, would remove it)
    c1 = (50.0, 0.0, 0.0)
    d1 = [1, 2, 10, 20, 30, 100]
    # cell 2: one lipid; threshold = its own d2 * 6.25 > d2, so it stays
    c2 = (100.0, 0.0, 0.0)
    d2 = [49]
    for k, (c, ds) in enumerate(((c0, d0), (c1, d1), (c2, d2))):
        pts += _cell(c, ds)
        aff += [k] * len(ds)
    cen = np.array([c0, c1, c2], np.float32)
    # a protein far from its cell and a lipid without a cell are never removed
    pts += [(0.0, 0.0, 90.0), (0.0, 0.0, -90.0)]
    aff += [0, -1]
    typ = np.zeros(len(pts), np.int32)
    typ[-2] = 2
    x = np.array(pts, np.float64)
    dele = orc.stray_delete(x, typ, np.array(aff, np.int32), cen, 2.5)
    want = np.zeros(len(pts), bool)
    want[9] = want[10] = True                     # 18.75 and 100 of cell 0
    assert np.array_equal(dele, want)
    # the threshold scales with tol^2, not tol: at tol = 6 (36 x 3 = 108) cell 0's 100 stays
    dele6 = orc.stray_delete(x, typ, np.array(aff, np.int32), cen, 6.0)
    assert not dele6.any()


@pytest.mark.skipif(not os.path.exists(REF_SO), reason="reference build absent (oracle/_ref)")
def test_cleanup_matches_reference_random(orc):
    rng = np.random.default_rng(3)
    from _systems import patch
    x = patch(24, 24, seed=5, proteins=False)["x"]
    lift = rng.choice(len(x), 15, replace=False)
    x[lift, 2] += rng.uniform(0.5, 5.0, len(lift))
    xs = x.astype(np.float32)
    cen = xs[5::14].copy()
    cen[:, 2] = 0.0
    aff, keep = Reference().stray(xs, cen, 2.5)
    dele = orc.stray_delete(xs.astype(np.float64), np.zeros(len(xs), np.int32), aff, cen, 2.5)
    assert np.array_equal(dele, ~keep)

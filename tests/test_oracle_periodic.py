"""Periodic boxes in the oracle (the reference has none; the build defines them).

Pinned by properties that do not restate the formula: the periodic cell list
equals periodic brute force; a rigid translation (then re-wrapping) leaves the
pair set and every force unchanged; a pair across the boundary interacts
exactly like the same pair unwrapped in an open box; a box much larger than
the system changes nothing; wrapping lands in [lo, lo + L) and is idempotent.
"""
import numpy as np
import pytest

from _systems import periodic_patch

pytestmark = []


@pytest.fixture
def pbox(oracle):
    yield oracle
    oracle.set_box(None)


def id_set(s, pairs):
    ids = s["id"].astype(np.int64)
    p = np.sort(np.stack([ids[pairs[:, 0]], ids[pairs[:, 1]]], 1), axis=1)
    return set(map(tuple, p.tolist()))


def test_wrap_range_and_idempotent(pbox):
    s = periodic_patch(12, 10, seed=1)
    lo, hi, per = s["box"]
    pbox.set_box(s["box"])
    w = pbox.wrap(s["x"])
    for d in (0, 1):
        assert np.all(w[:, d] >= lo[d]) and np.all(w[:, d] < hi[d])
        k = np.round((s["x"][:, d] - w[:, d]) / (hi[d] - lo[d]))
        assert np.allclose(s["x"][:, d] - w[:, d], k * (hi[d] - lo[d]), atol=1e-9)
    assert np.array_equal(w[:, 2], s["x"][:, 2])
    assert np.array_equal(pbox.wrap(w), w)


@pytest.mark.parametrize("seed", [0, 3])
def test_periodic_cell_list_equals_brute(pbox, seed):
    s = periodic_patch(14, 12, seed=seed)
    pbox.set_box(s["box"])
    x = pbox.wrap(s["x"])
    a = pbox.pairs_brute(x, s["type"])
    b = pbox.pairs_cell(x, s["type"])
    assert len(a) > 0 and np.array_equal(a, b)
    # the periodic sheet has more pairs than the same sheet with open edges
    pbox.set_box(None)
    assert len(pbox.pairs_brute(x, s["type"])) < len(a)


def test_periodic_translation_invariance(pbox):
    s = periodic_patch(14, 12, seed=2)
    lo, hi, _ = s["box"]
    pbox.set_box(s["box"])
    x0 = pbox.wrap(s["x"])
    shift = np.array([0.37 * (hi[0] - lo[0]), 0.81 * (hi[1] - lo[1]), 0.0])
    x1 = pbox.wrap(x0 + shift)
    p0 = pbox.pairs_cell(x0, s["type"])
    p1 = pbox.pairs_cell(x1, s["type"])
    assert id_set(s, p0) == id_set(s, p1)
    f0, t0, e0, _ = pbox.forces(x0, s["n"], s["type"], bonds=s["bonds"])
    f1, t1, e1, _ = pbox.forces(x1, s["n"], s["type"], bonds=s["bonds"])
    scale = np.abs(f0).max()
    assert np.abs(f0 - f1).max() < 1e-9 * scale
    assert np.abs(t0 - t1).max() < 1e-9 * max(1.0, np.abs(t0).max())
    assert abs(e0.sum() - e1.sum()) < 1e-9 * abs(e0.sum())


def test_pair_across_boundary_equals_open_pair(pbox):
    """A lipid pair split by the x boundary, and a bond split by the y boundary."""
    lo = np.array([0.0, 0.0, -50.0]); hi = np.array([20.0, 20.0, 50.0])
    n = np.array([[0.0, 0.0, 1.0], [0.1, 0.0, 0.995]])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    xa = np.array([[19.6, 5.0, 0.0], [0.5, 5.2, 0.1]])        # 0.9 apart through x = 20 == 0
    xo = np.array([[19.6, 5.0, 0.0], [20.5, 5.2, 0.1]])
    t = np.array([0, 0], np.int32)
    pbox.set_box((lo, hi, [1, 1, 0]))
    fa, ta, ea, _ = pbox.forces(xa, n, t, pairs=np.array([[0, 1]]))
    pbox.set_box(None)
    fo, to, eo, _ = pbox.forces(xo, n, t, pairs=np.array([[0, 1]]))
    assert np.abs(fo).max() > 0
    assert np.allclose(fa, fo, rtol=1e-12, atol=1e-14) and np.allclose(ta, to, rtol=1e-12, atol=1e-14)
    xb = np.array([[3.0, 19.9, 0.0], [3.0, 1.0, -2.0]])
    xbo = np.array([[3.0, 19.9, 0.0], [3.0, 21.0, -2.0]])
    bonds = (np.array([0], np.int32), np.array([0]), np.array([1]))
    pbox.set_box((lo, hi, [1, 1, 0]))
    fb, _, _ = pbox.bonded(xb, bonds)
    pbox.set_box(None)
    fbo, _, _ = pbox.bonded(xbo, bonds)
    assert np.abs(fbo).max() > 0 and np.allclose(fb, fbo, rtol=1e-12, atol=1e-14)


def test_large_periodic_box_changes_nothing(pbox):
    s = periodic_patch(10, 8, seed=4)
    x = s["x"]
    ref = pbox.forces(x, s["n"], s["type"], bonds=s["bonds"])
    pbox.set_box((np.array([-500.0, -500.0, -500.0]), np.array([500.0, 500.0, 500.0]), [1, 1, 1]))
    got = pbox.forces(x, s["n"], s["type"], bonds=s["bonds"])
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)

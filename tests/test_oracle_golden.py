"""The oracle against the OpenRBC reference's own outputs.

tests/golden/*.npz hold inputs and the reference's outputs (FP32, fast-math,
as shipped), written by tests/golden/make_golden.py from the reference
compiled in place.  The FP64 oracle must reproduce them to FP32 accuracy.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name))


def combined(z):
    """Reference containers -> one particle array (lipids first) + bonds by index."""
    nl, np_ = len(z["lx"]), len(z["px"])
    x = np.vstack([z["lx"], z["px"]]).astype(np.float64)
    n = np.vstack([z["ln"], z["pn"]]).astype(np.float64)
    t = np.concatenate([np.zeros(nl, np.int32), z["ptype"]])
    tag2idx = {int(tg): nl + k for k, tg in enumerate(z["ptag"])}
    bi = np.array([tag2idx[int(v)] for v in z["bi"]], np.int64)
    bj = np.array([tag2idx[int(v)] for v in z["bj"]], np.int64)
    return x, n, t, (z["bt"], bi, bj), nl


def test_force_field_tables(oracle):
    """Folded 6x6 tables == the reference's per-kind tables (FP32 rounding)."""
    g = load("ff_tables.npz")
    f = oracle.ff()
    # rep/att are constexpr-evaluated in FP32 by the reference (pow by 8 chained float products)
    rel = lambda a, b: abs(a - b) <= 2e-6 * max(abs(a), abs(b), 1e-30)
    for p in range(1, 6):        # lipid-protein uses the protein-type row (src/forcefield_canonical.h:116-122)
        assert rel(f["cut2_poly"][0, p], g["cutsqlp"][p])
        assert rel(f["rep"][0, p], g["replp"][p]) and rel(f["att"][0, p], g["attlp"][p])
        assert rel(f["alpha"][0, p], g["alphalp"][p])
        assert rel(f["cut2_lj"][p, 0], g["lj_cutsq"][p])          # lj_cutsq[type], row 0
        assert rel(f["lj1"][p, 0], g["lj_lj1"][p]) and rel(f["lj2"][p, 0], g["lj_lj2"][p])
    for a in range(1, 6):
        for b in range(1, 6):
            k = a + 6 * b           # type12 = type1 + type2 * n_type
            assert rel(f["cut2_poly"][a, b], g["cutsqpp"][k])
            assert rel(f["rep"][a, b], g["reppp"][k])
            assert rel(f["cut2_lj"][a, b], g["lj_cutsq"][k])
            assert rel(f["lj1"][a, b], g["lj_lj1"][k]) and rel(f["lj2"][a, b], g["lj_lj2"][k])
    assert np.allclose(f["mass"], g["mass"]) and np.allclose(f["radius"], g["radius"], rtol=1e-7)
    assert np.allclose(f["bond_r0"], g["r0"], rtol=1e-7) and np.allclose(f["bond_K"], g["K"])


@pytest.mark.parametrize("seed", [3, 8])
def test_forces_match_reference(oracle, seed):
    z = load(f"forces_patch{seed}.npz")
    x, n, t, bonds, nl = combined(z)
    f, tq, _, _ = oracle.forces(x, n, t, bonds=bonds)
    F = np.vstack([z["lf"], z["pf"]]); T = np.vstack([z["lt"], z["pt"]])
    scale_f = np.sqrt(np.mean(np.sum(f * f, 1)))
    scale_t = np.sqrt(np.mean(np.sum(tq * tq, 1)))
    assert np.abs(f - F).max() / scale_f < 2e-5
    assert np.abs(tq - T).max() / scale_t < 2e-5


def _state(z, pre):
    keys = ("x", "v", "n", "o", "f", "t")
    s = {k: np.vstack([z[f"{pre}l{k}"], z[f"{pre}p{k}"]]).astype(np.float64) for k in keys}
    return s


def test_langevin_kT0_matches_reference(oracle):
    z = load("langevin_T0.npz")
    s = _state(z, "")
    nl = len(z["lx"])
    s["type"] = np.concatenate([np.zeros(nl, np.int32), z["ptype"]]).astype(np.int32)
    s["id"] = np.arange(1, len(s["type"]) + 1, dtype=np.uint32)
    oracle.langevin(s, 0.0, 0.01, 0.01, 1, 0)
    o = _state(z, "out_")
    for k in ("x", "v", "n", "o"):
        assert np.abs(s[k] - o[k]).max() < 5e-6 * max(1.0, np.abs(o[k]).max()), k
    assert np.abs(s["f"]).max() == 0 and np.abs(s["t"]).max() == 0


@pytest.mark.parametrize("which", [0, 1])
def test_nose_hoover_matches_reference(oracle, which):
    z = load(f"nh{which}.npz")
    s = _state(z, "")
    nl = len(z["lx"])
    s["type"] = np.concatenate([np.zeros(nl, np.int32), z["ptype"]]).astype(np.int32)
    zeta, Q, kT = float(z["zeta_in"]), float(z["Q"]), float(z["kT"])
    n = len(s["type"])
    if which == 0:
        ke = oracle.nh_first(s, 0.01, zeta)
    else:
        ke = oracle.nh_second(s, 0.01, zeta)
    znew = oracle.nh_zeta(zeta, 0.01, Q, ke, n, kT)
    o = _state(z, "out_")
    for k in ("x", "v", "n", "o", "t"):
        assert np.abs(s[k] - o[k]).max() < 5e-6 * max(1.0, np.abs(o[k]).max()), k
    assert abs(znew - float(z["zeta_out"])) < 1e-5 * max(1.0, abs(znew))


def test_temperature_matches_reference(oracle):
    z = load("temperature.npz")
    v = np.vstack([z["lv"], z["pv"]]).astype(np.float64)
    t = np.concatenate([np.zeros(len(z["lv"]), np.int32), z["ptype"]])
    assert abs(oracle.temperature(v, t) - float(z["T"])) < 1e-6 * float(z["T"])


def test_morton_matches_reference(oracle):
    z = load("morton.npz")
    assert np.array_equal(oracle.morton_ref(z["pts"]), z["keys"])
    # reorder_morton: the reference's output is the input sorted by key
    sub = z["pts"][:500]
    keys = oracle.morton_ref(sub)
    got_keys = oracle.morton_ref(z["reordered"])
    assert np.all(np.diff(got_keys.astype(np.int64)) >= 0)
    assert np.array_equal(np.sort(keys), got_keys)
    assert sorted(map(tuple, sub.tolist())) == sorted(map(tuple, z["reordered"].tolist()))


def test_voronoi_queries_match_reference(oracle):
    z = load("voronoi.npz")
    cen, q = z["centroids"], z["queries"]
    aff = oracle.partition(q.astype(np.float64), cen)
    assert np.array_equal(aff, z["partition"])
    assert np.array_equal(aff, z["nearest"])
    off = 0
    for k, c in enumerate(z["within_counts"]):
        exp = z["within_flat"][off:off + c]
        off += c
        d = (cen - q[k]).astype(np.float32) ** 2
        got = np.where((d[:, 0] + d[:, 1]) + d[:, 2] < np.float32(81.0))[0]
        assert np.array_equal(np.sort(exp), got)

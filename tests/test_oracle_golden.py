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


# ---- WCA and protein-protein pairs (the contact patch) ------------------------------------
def _kinds_of(oracle, x, t, pairs):
    """Pair kind of each interacting pair: (min type, max type, 'poly' | 'lj')."""
    ff = oracle.ff()
    a, b = pairs[:, 0], pairs[:, 1]
    r2 = np.sum((x[a] - x[b]) ** 2, 1)
    poly = r2 < ff["cut2_poly"][t[a], t[b]]
    return [(int(min(p, q)), int(max(p, q)), "poly" if pl else "lj") for p, q, pl in zip(t[a], t[b], poly)]


def test_contact_forces_match_reference(oracle):
    """compute_pairwise_fused + compute_bonded of the reference on a patch that
    holds every WCA pair (lipid-actin, lipid-spectrin, band-3-actin,
    band-3-spectrin) and every protein-protein polynomial pair among types
    1-3: the oracle reproduces it to FP32 accuracy, and it would not without
    its WCA or protein-protein terms (src/pairwise_kernel_fused.h:63-97)."""
    z = load("forces_contacts.npz")
    x, n, t, bonds, nl = combined(z)
    pairs = oracle.pairs_cell(x, t)
    kinds = _kinds_of(oracle, x, t, pairs)
    need = [(0, 4, "lj"), (0, 5, "lj"), (1, 4, "lj"), (1, 5, "lj"), (2, 4, "lj"), (2, 5, "lj"),
            (1, 1, "poly"), (1, 2, "poly"), (1, 3, "poly"), (2, 2, "poly"), (2, 3, "poly"), (3, 3, "poly"),
            (0, 1, "poly"), (0, 2, "poly"), (0, 3, "poly")]
    for k in need:
        assert kinds.count(k) > 0, k
    f, tq, _, _ = oracle.forces(x, n, t, pairs=pairs, bonds=bonds)
    F = np.vstack([z["lf"], z["pf"]]); T = np.vstack([z["lt"], z["pt"]])
    scale_f = np.sqrt(np.mean(np.sum(f * f, 1)))
    scale_t = np.sqrt(np.mean(np.sum(tq * tq, 1)))
    assert np.abs(f - F).max() / scale_f < 2e-5
    assert np.abs(tq - T).max() / scale_t < 2e-5
    # the comparison is sensitive to each of the two terms
    for drop in ("lj", "pp"):
        keep = np.array([not ((drop == "lj" and k[2] == "lj") or (drop == "pp" and k[2] == "poly" and k[0] > 0))
                         for k in kinds])
        g, _, _, _ = oracle.forces(x, n, t, pairs=pairs[keep], bonds=bonds)
        assert np.abs(g - F).max() / scale_f > 100 * 2e-5, drop


# ---- the Voronoi cell stencil over many cells ----------------------------------------------
def test_voronoi_stencil_forces_match_reference(oracle):
    """The reference's cell-stencil sweep over 22 Voronoi cells with given
    centroids (src/compute_pairwise_fused.h:253-316): the oracle's stencil
    pair set (orc_pairs_voronoi_fp64: centroid distance < 9 / 8 / 6 for
    protein-protein / protein-lipid / lipid-lipid) gives the reference's
    forces; the cells are large enough that every radius drops pairs within
    the cutoff, and restoring the dropped pairs of any one kind breaks the
    agreement (so a swapped radius would be caught)."""
    z = load("voronoi_forces.npz")
    x, n, t, bonds, nl = combined(z)
    aff = np.concatenate([z["laff"], z["paff"]])
    cen = z["centroids"]
    vor = oracle.pairs_voronoi(x, t, aff, cen)
    full = oracle.pairs_cell(x, t)
    f, tq, _, _ = oracle.forces(x, n, t, pairs=vor, bonds=bonds)
    F = np.vstack([z["lf"], z["pf"]]); T = np.vstack([z["lt"], z["pt"]])
    sf = np.sqrt(np.mean(np.sum(f * f, 1))); st = np.sqrt(np.mean(np.sum(tq * tq, 1)))
    assert np.abs(f - F).max() / sf < 2e-5
    assert np.abs(tq - T).max() / st < 2e-5
    vs = set(map(tuple, np.sort(vor, 1).tolist()))
    dropped = np.array([p for p in map(tuple, np.sort(full, 1).tolist()) if p not in vs], np.int64)
    kind = np.where((t[dropped[:, 0]] == 0) & (t[dropped[:, 1]] == 0), 0,
                    np.where((t[dropped[:, 0]] == 0) | (t[dropped[:, 1]] == 0), 1, 2))
    for k in (0, 1, 2):                     # lipid-lipid, protein-lipid, protein-protein
        assert (kind == k).sum() > 0, k
        g, _, _, _ = oracle.forces(x, n, t, pairs=np.vstack([vor, dropped[kind == k]]), bonds=bonds)
        assert np.abs(g - F).max() / sf > 50 * 2e-5, k


# ---- the minimiser (src/openrbc.cpp:110-132), from the reference program's own frames -------
def test_minimize_move_free_matches_reference(oracle):
    """Caps of 1e9: every particle moves by x += f dt / m with the default dt
    (proteins of all four kinds and masses); spectrin directors, which the
    initialiser sets to un-normalised positions, come out normalised."""
    z = load("minimize.npz")
    x0, n0 = z["free_x0"].astype(np.float64), z["free_n0"].astype(np.float64)
    f0, ty = z["free_f0"].astype(np.float64), z["free_type"]
    st = dict(x=x0.copy(), n=n0.copy(), f=f0, t=np.zeros_like(f0), type=ty.astype(np.int32))
    oracle.minimize_move(st, 1e9, 1e9, 0.01)
    assert set(np.unique(ty).tolist()) == {2, 3, 4, 5}
    assert np.abs(st["x"] - z["free_x1"]).max() < 1e-5 * max(1.0, np.abs(x0).max()) / 100
    sp = ty == 5
    assert np.abs(st["n"][sp] - z["free_n1"][sp]).max() < 1e-6
    # the step is not trivially zero
    assert np.abs(z["free_x1"] - z["free_x0"]).max() > 0.05


def test_minimize_move_capped_matches_reference(oracle):
    """Default caps (0.05 / 0.05): every particle's step is capped either by
    its displacement or by its director rotation, as the reference's rule
    min(dr_opt / |f/m|, dn_opt / |n x t x n|) makes it.  The oracle's move,
    given the reference's force and the director force of its own exact pair
    set (for the particles whose force the reference's stencil reproduces),
    gives the reference's positions and directors."""
    z = load("minimize.npz")
    x0, n0 = z["capped_x0"].astype(np.float64), z["capped_n0"].astype(np.float64)
    x1, n1 = z["capped_x1"].astype(np.float64), z["capped_n1"].astype(np.float64)
    f0, ty, tag = z["capped_f0"].astype(np.float64), z["capped_type"], z["capped_tag"]
    m = oracle.ff()["mass"][ty]
    unit = np.abs(np.linalg.norm(n0, axis=1) - 1) < 1e-5
    dx = np.linalg.norm(x1 - x0, axis=1)
    w = n1 / np.sum(n1 * n0, 1)[:, None] - n0          # director update before normalisation
    dw = np.linalg.norm(w, axis=1)
    tcap = np.abs(dx / 0.05 - 1) < 1e-4
    rcap = np.abs(dw / 0.05 - 1) < 1e-4
    assert (tcap | rcap)[unit].all() and tcap.sum() > 0 and rcap.sum() > 0
    assert (dw[tcap & unit] <= 0.05 * (1 + 1e-4)).all() and (dx[rcap] <= 0.05 * (1 + 1e-4)).all()
    # director force from the oracle: bonds of the reference initialiser by tag
    import openrbc_amd.systems as systems
    import openrbc_amd as orbc
    s = orbc.init_rbc(z["vert"], z["edge"], z["face"], rho=1.05, seed=1)
    tag2i = {int(q): i for i, q in enumerate(tag) if q >= 0}
    bi = np.array([tag2i[int(q)] for q in s["bond_i"]]); bj = np.array([tag2i[int(q)] for q in s["bond_j"]])
    fo, to, _, _ = oracle.forces(x0, n0, ty.astype(np.int32), bonds=(s["bond_type"], bi, bj))
    good = unit & (np.linalg.norm(fo - f0, axis=1) / np.maximum(np.linalg.norm(f0, axis=1), 1.0) < 1e-4)
    assert good.sum() > 0.8 * len(ty)
    st = dict(x=x0.copy(), n=n0.copy(), f=f0, t=np.cross(n0, to), type=ty.astype(np.int32))
    oracle.minimize_move(st, 0.05, 0.05, 0.01)
    assert np.abs(st["x"] - x1)[good].max() < 1e-5
    assert np.abs(st["n"] - n1)[good].max() < 2e-6
    # a reversed director update would move every director the other way
    st2 = dict(x=x0.copy(), n=n0.copy(), f=f0, t=-np.cross(n0, to), type=ty.astype(np.int32))
    oracle.minimize_move(st2, 0.05, 0.05, 0.01)
    assert np.median(np.linalg.norm(st2["n"] - n1, axis=1)[good]) > 100 * 2e-6
    _ = systems

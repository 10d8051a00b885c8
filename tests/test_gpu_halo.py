"""Block halos of the fused steps (Langevin, Nose-Hoover second half).

With ``block_halo`` on, each block of 256 particles stages its own data (FP32
mirrors, or FP64 positions and directors) and that of its outside neighbours
in LDS and walks its list as 16-bit local indices; a block whose outside
neighbours do not fit gathers from the 32-bit list.  The pair terms and their
order are the same either way, so a run with halos must be bit-identical to a
run without them, through neighbour rebuilds, with and without blocks that
overflow.
"""
import numpy as np
import pytest

import openrbc_amd as orbc

pytestmark = pytest.mark.gpu


def _rbc(level=2, seed=1):
    """Icosphere cell through the initialiser: lipids, the four protein types
    and the spectrin network with its bonds."""
    v, e, f = orbc.icosphere(level)
    s = orbc.init_rbc(v, e, f, seed=seed)
    rng = np.random.default_rng(seed)
    x = np.vstack([s["lipid_x"], s["protein_x"]])
    n = np.vstack([s["lipid_n"], s["protein_n"]])
    t = np.concatenate([np.zeros(len(s["lipid_x"]), np.int32), s["protein_type"]])
    ids = np.concatenate([s["lipid_id"], s["protein_id"]])
    idx = {int(k): i for i, k in enumerate(ids)}
    bi = np.array([idx[int(k)] for k in s["bond_i"]]); bj = np.array([idx[int(k)] for k in s["bond_j"]])
    return dict(x=x, n=n, type=t, id=ids, v=rng.normal(0, 0.3, x.shape), o=np.zeros_like(x),
                bonds=(s["bond_type"], bi, bj))


def _with_blob(s, m=16, spacing=1.0, offset=(400.0, 0.0, 0.0), seed=3):
    """Append an m^3 cubic lattice of lipids (about 140 list entries each, so
    a block's outside neighbours overflow the halo) far from the patch."""
    rng = np.random.default_rng(seed)
    g = np.arange(m) * spacing
    p = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) + np.array(offset)
    p += rng.normal(0, 0.02, p.shape)
    nb = len(p)
    nrm = rng.normal(0, 1, (nb, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    out = dict(s)
    out["x"] = np.vstack([s["x"], p])
    out["n"] = np.vstack([s["n"], nrm])
    out["v"] = np.vstack([s["v"], rng.normal(0, 0.1, (nb, 3))])
    out["o"] = np.vstack([s["o"], rng.normal(0, 0.05, (nb, 3))])
    out["type"] = np.concatenate([s["type"], np.zeros(nb, np.int32)])
    out["id"] = np.concatenate([s["id"], np.arange(nb, dtype=np.uint32) + s["id"].max() + 1])
    return out


def _run(s, halo, nsteps, minimize=0, cells=False, mode=None, precision=orbc.MIXED, halo_capacity=0, **kw):
    # block_halo 2: halos at every rebuild; 0: never; 1: while lists last 4+ steps
    sim = orbc.Simulation(precision=precision, block_halo=2 * halo if mode is None else mode, **kw)
    if halo_capacity:
        sim.set_tuning(**{**sim.tuning(), "halo_capacity": halo_capacity})
    t = s["type"]
    lip = np.where(t == 0)[0]
    prot = np.where(t != 0)[0]
    sim.set_state(orbc.LIPID, s["x"][lip], s["n"][lip], s["id"][lip], v=s["v"][lip], o=s["o"][lip])
    if len(prot):
        sim.set_state(orbc.PROTEIN, s["x"][prot], s["n"][prot], s["id"][prot], v=s["v"][prot], o=s["o"][prot],
                      type=t[prot])
    bt, bi, bj = s["bonds"]
    if len(bt):
        sim.set_bonds(bt, s["id"][bi], s["id"][bj])
    if minimize:
        sim.minimize(minimize)
    if cells:
        sim.voronoi_init(0, 4)
    sim.run(nsteps)
    st = {sp: sim.get_state(sp) for sp in (orbc.LIPID, orbc.PROTEIN) if (len(lip) if sp == orbc.LIPID else len(prot))}
    obs = sim.observables()
    sim.close()
    return st, obs


def _same(a, b):
    for sp in a:
        for k in a[sp]:
            assert np.array_equal(a[sp][k], b[sp][k]), (sp, k)


@pytest.fixture(scope="module")
def rbc():
    return _rbc(2)


@pytest.mark.parametrize("precision", [orbc.MIXED, orbc.FP64])
@pytest.mark.parametrize("skin,thermostat", [(0.05, orbc.RUN_LANGEVIN), (0.4, orbc.RUN_LANGEVIN),
                                              (0.05, orbc.RUN_NOSE_HOOVER)])
def test_block_halo_run_bit_identical(rbc, skin, thermostat, precision):
    s = rbc
    print(f"halo run: icosphere-2 cell, n={len(s['type'])}, skin {skin}, thermostat {thermostat}, "
          f"precision {precision}", flush=True)
    kw = dict(minimize=30, seed=5, skin=skin, thermostat=thermostat, precision=precision)
    on, obs_on = _run(s, 1, 40, **kw)
    off, obs_off = _run(s, 0, 40, **kw)
    _same(on, off)
    assert obs_on["rebuilds"] == obs_off["rebuilds"] and obs_on["rebuilds"] > 1
    assert obs_on["halo_blocks"] == (len(s["type"]) + 255) // 256
    assert obs_on["halo_overflow"] < obs_on["halo_blocks"]
    assert obs_off["halo_blocks"] == 0


@pytest.mark.parametrize("precision", [orbc.MIXED, orbc.FP64])
@pytest.mark.parametrize("thermostat", [orbc.RUN_LANGEVIN, orbc.RUN_NOSE_HOOVER])
def test_block_halo_overflow_blocks_gather_globally(rbc, thermostat, precision):
    s = _with_blob(rbc)
    print(f"halo overflow run: n={len(s['type'])}, thermostat {thermostat}, precision {precision}", flush=True)
    kw = dict(minimize=20, seed=4, skin=0.2, thermostat=thermostat, precision=precision)
    on, obs_on = _run(s, 1, 6, **kw)
    off, _ = _run(s, 0, 6, **kw)
    _same(on, off)
    assert 0 < obs_on["halo_overflow"] < obs_on["halo_blocks"]


@pytest.mark.parametrize("precision", [orbc.MIXED, orbc.FP64])
@pytest.mark.parametrize("thermostat", [orbc.RUN_LANGEVIN, orbc.RUN_NOSE_HOOVER])
def test_wide_halo_kernels_bit_identical(rbc, thermostat, precision):
    """The wide halo kernels (1024 local indices, the autotuner's halo
    capacity knob) against gathers, with blocks that overflow even them."""
    s = _with_blob(rbc)
    print(f"wide halo run: n={len(s['type'])}, thermostat {thermostat}, precision {precision}", flush=True)
    kw = dict(minimize=20, seed=4, skin=0.2, thermostat=thermostat, precision=precision)
    wide, obs_w = _run(s, 1, 6, halo_capacity=1024, **kw)
    off, _ = _run(s, 0, 6, **kw)
    _, obs_d = _run(s, 1, 6, **kw)
    _same(wide, off)
    assert 0 < obs_w["halo_overflow"] <= obs_d["halo_overflow"] < obs_w["halo_blocks"]


@pytest.mark.parametrize("halo", [0, 1])
def test_fused_nose_hoover_matches_unfused(rbc, halo):
    """The Nose-Hoover second half fused with the forces (no Voronoi cells in
    the run) against the separate force and integrate kernels (taken while
    cells are kept for the stray-lipid cleanup, which finds nothing to remove
    here): the same arithmetic, so the same trajectory bit for bit."""
    s = rbc
    kw = dict(minimize=30, seed=5, skin=0.05, thermostat=orbc.RUN_NOSE_HOOVER)
    fused, obs_f = _run(s, halo, 30, **kw)
    split, obs_s = _run(s, halo, 30, cells=True, **kw)
    assert obs_s["lipids_deleted"] == 0
    assert obs_f["rebuilds"] == obs_s["rebuilds"] and obs_f["rebuilds"] > 1
    _same(fused, split)


def test_block_halo_automatic_mode_skips_short_lived_lists(rbc):
    """Lists that last a step or two (tiny skin): the automatic mode stops
    building halos after the first rebuilds; the trajectory is unchanged."""
    s = rbc
    auto, obs_a = _run(s, 0, 30, minimize=30, seed=5, skin=0.01, mode=1)
    off, obs_o = _run(s, 0, 30, minimize=30, seed=5, skin=0.01, mode=0)
    _same(auto, off)
    assert obs_a["rebuilds"] > 10
    assert obs_a["halo_blocks"] == 0


@pytest.mark.parametrize("precision", [orbc.MIXED, orbc.FP64])
@pytest.mark.parametrize("thermostat", [orbc.RUN_LANGEVIN, orbc.RUN_NOSE_HOOVER])
def test_block_halo_periodic_box_bit_identical(precision, thermostat):
    """Minimum images on the periodic axes: the halo walks take the same
    anchor-difference wrap (mixed) and FP64 minimum image as the gather walks."""
    from _systems import periodic_patch
    s = periodic_patch(48, 48, seed=3)
    lo, hi, per = s["box"]
    print(f"halo periodic run: n={len(s['type'])}, precision {precision}, thermostat {thermostat}", flush=True)
    kw = dict(minimize=20, seed=6, skin=0.1, thermostat=thermostat, precision=precision,
              box_lo=list(lo), box_hi=list(hi), periodic=[int(p) for p in per])
    on, obs_on = _run(s, 1, 30, **kw)
    off, _ = _run(s, 0, 30, **kw)
    _same(on, off)
    assert obs_on["halo_blocks"] == (len(s["type"]) + 255) // 256 and obs_on["rebuilds"] > 1

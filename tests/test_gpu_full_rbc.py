"""The full-size red blood cell (OpenRBC's example-large mesh populated by the
init_rbc rules: 2,639,690 lipids + 565,816 proteins, 584,066 bonds) on the
GPU: the FP64 drop-in and the mixed-precision path against the oracle on the
minimised configuration, and the main loop with the volume constraint in
both precisions.

The oracle's cell list runs on the host (about half a minute for the 3.2M
particles); the comparison state is the FP64-minimised one, as in the
benchmark.
"""
import numpy as np
import pytest

import openrbc_amd as orbc

pytestmark = pytest.mark.gpu

EPS_MIXED = 1e-5
EPS_FP64 = 1e-11


@pytest.fixture(scope="module")
def rbc(oracle):
    v, e, f = orbc.load_example_large()
    s = orbc.init_rbc(v, e, f, rho=1.05, seed=7)
    nl, npr = len(s["lipid_x"]), len(s["protein_x"])
    assert (nl, npr, len(s["bond_type"])) == (2639690, 565816, 584066)
    sim = orbc.Simulation(precision=orbc.FP64, seed=7, skin=0.6)
    orbc.system_to_simulation(s, sim)
    sim.minimize(60)
    L, P = sim.get_state(orbc.LIPID), sim.get_state(orbc.PROTEIN)
    sim.close()
    x = np.vstack([L["x"], P["x"]])
    n = np.vstack([L["n"], P["n"]])
    t = np.concatenate([np.zeros(nl, np.int32), s["protein_type"]]).astype(np.int32)
    ids = np.concatenate([s["lipid_id"], s["protein_id"]]).astype(np.int64)
    order = np.argsort(ids)
    to_idx = lambda q: order[np.searchsorted(ids[order], q)]
    bonds = (s["bond_type"], to_idx(s["bond_i"]), to_idx(s["bond_j"]))
    pairs = oracle.pairs_cell(x, t)
    f64, t64, _, _ = oracle.forces(x, n, t, pairs=pairs, bonds=bonds)
    mins = dict(lx=L["x"], ln=L["n"], px=P["x"], pn=P["n"])
    return dict(sys=s, x=x, n=n, t=t, ids=ids, to_idx=to_idx, bonds=bonds, pairs=pairs, f=f64, tq=t64, mins=mins)


def _sim(r, precision, **kw):
    s, m = r["sys"], r["mins"]
    sim = orbc.Simulation(precision=precision, seed=7, skin=0.6, **kw)
    sim.set_state(orbc.LIPID, m["lx"], m["ln"], s["lipid_id"])
    sim.set_state(orbc.PROTEIN, m["px"], m["pn"], s["protein_id"], type=s["protein_type"])
    sim.set_bonds(s["bond_type"], s["bond_i"], s["bond_j"])
    return sim


def _forces(sim, r):
    nl = len(r["sys"]["lipid_x"])
    sim.compute_pairwise(zero_first=True)
    sim.compute_bonded()
    lf, lt = sim.get_forces(orbc.LIPID)
    pf, pt = sim.get_forces(orbc.PROTEIN)
    return np.vstack([lf, pf]), np.vstack([lt, pt]), nl


def _sorted_id_pairs(r, pairs_idx):
    p = np.sort(r["ids"][pairs_idx], axis=1)
    key = p[:, 0] * (1 << 32) + p[:, 1]
    return np.sort(key)


def test_full_rbc_fp64_vs_oracle(rbc):
    """C2, the FP64 drop-in: the pair set equals the oracle's cell list bit
    for bit, and forces and director forces agree to 1e-11."""
    sim = _sim(rbc, orbc.FP64)
    got = sim.pairs()
    gk = np.sort(got[:, 0].astype(np.int64) * (1 << 32) + got[:, 1])
    assert np.array_equal(gk, _sorted_id_pairs(rbc, rbc["pairs"]))
    f, t, _ = _forces(sim, rbc)
    rms_f = np.sqrt(np.mean(np.sum(rbc["f"] ** 2, 1)))
    rms_t = np.sqrt(np.mean(np.sum(rbc["tq"] ** 2, 1)))
    assert np.abs(f - rbc["f"]).max() / rms_f < EPS_FP64
    assert np.abs(t - rbc["tq"]).max() / rms_t < EPS_FP64


def test_full_rbc_mixed_vs_oracle(rbc, oracle):
    """C3, the precision form: eps_F <= 1e-5 for forces and director forces.
    Pairs that FP32 selects differently sit within FP32 rounding of a cutoff
    (the WCA cut 1.1225 sigma, where the force jumps); they are checked to be
    that, and the oracle is evaluated on the mixed path's own pair set."""
    sim = _sim(rbc, orbc.MIXED)
    got = sim.pairs()
    gk = np.sort(got[:, 0].astype(np.int64) * (1 << 32) + got[:, 1])
    ek = _sorted_id_pairs(rbc, rbc["pairs"])
    flips = np.setxor1d(gk, ek)
    assert len(flips) <= 20
    if len(flips):
        a, b = rbc["to_idx"](flips >> 32), rbc["to_idx"](flips & 0xffffffff)
        rr = np.linalg.norm(rbc["x"][a] - rbc["x"][b], axis=1)
        ff = oracle.ff()
        ta, tb = rbc["t"][a], rbc["t"][b]
        cuts = np.stack([np.sqrt(ff["cut2_poly"][ta, tb]), np.sqrt(ff["cut2_lj"][ta, tb])], 1)
        assert np.all(np.min(np.abs(rr[:, None] / np.where(cuts > 0, cuts, 1e9) - 1), 1) < 1e-5)
        pairs = np.stack([rbc["to_idx"](gk >> 32), rbc["to_idx"](gk & 0xffffffff)], 1)
        fo, to, _, _ = oracle.forces(rbc["x"], rbc["n"], rbc["t"], pairs=pairs, bonds=rbc["bonds"])
    else:
        fo, to = rbc["f"], rbc["tq"]
    f, t, _ = _forces(sim, rbc)
    assert np.abs(f - fo).max() / np.sqrt(np.mean(np.sum(fo ** 2, 1))) < EPS_MIXED
    assert np.abs(t - to).max() / np.sqrt(np.mean(np.sum(to ** 2, 1))) < EPS_MIXED


def test_full_rbc_loop_with_volume_constraint(rbc):
    """The main loop of C2 and C3 (Langevin, skin rebuilds, Voronoi cells
    updated every 2 steps, the volume constraint at the initial enclosed
    volume) for 40 steps in both precisions from the same state with the same
    noise: both stay finite near kT's relaxation path, the mixed trajectory
    follows the FP64 one, and the enclosed volume stays at its target."""
    out = {}
    for prec in (orbc.FP64, orbc.MIXED):
        sim = _sim(rbc, prec, volume_target=-1.0, volume_strength=0.05)
        rng = np.random.default_rng(7)
        from openrbc_amd.systems import maxwell
        s = rbc["sys"]
        lv = maxwell(np.zeros(len(s["lipid_x"]), np.int32), 0.22, rng)
        pv = maxwell(s["protein_type"], 0.22, rng)
        sim.set_state(orbc.LIPID, rbc["mins"]["lx"], rbc["mins"]["ln"], s["lipid_id"], v=lv)
        sim.set_state(orbc.PROTEIN, rbc["mins"]["px"], rbc["mins"]["pn"], s["protein_id"], v=pv,
                      type=s["protein_type"])
        sim.voronoi_init(0, 8)
        V0 = sim.constrain_volume(1.0, 0.0)
        o0 = sim.observables()
        sim.run(40)
        o1 = sim.observables()
        st = sim.get_state(orbc.LIPID)
        assert np.isfinite(st["x"]).all() and np.isfinite(st["v"]).all()
        assert o1["step"] == 40 and 0.05 < o1["temperature"] < 0.3
        V1 = sim.constrain_volume(1.0, 0.0)
        assert abs(V1 / V0 - 1) < 1e-3
        out[prec] = (st["x"], (o1["kinetic"] + o1["potential"]) / o1["n"], (o0["kinetic"] + o0["potential"]) / o0["n"])
    x64, e64, e0 = out[orbc.FP64]
    x32, e32, _ = out[orbc.MIXED]
    assert np.abs(x32 - x64).max() < 1e-4
    assert abs(e32 - e64) < 1e-5 * abs(e64)
    assert abs(e64 - e0) < 0.05 * abs(e0)

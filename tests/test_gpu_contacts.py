"""GPU parity on the contact patch: a membrane whose pairs include every WCA
Lennard-Jones kind (lipid-actin, lipid-spectrin, band-3-actin,
band-3-spectrin) and every protein-protein polynomial kind among types 1-3
(src/pairwise_kernel_fused.h:63-97), next to the anisotropic lipid terms and
the bonds.  The oracle (pinned on this very patch by the reference's
forces_contacts golden) is the checker.

Bars: FP64 forces within 1e-11 of the oracle (relative to the rms force),
mixed precision within eps_F = max |dF| / rms |F| <= 1e-5.  For whole steps
the velocity and angular-velocity errors are converted back into force
errors (m dv / dt and do / dt), so the fused step kernels are held to the
same eps_F per force evaluation.
"""
import numpy as np
import pytest

import openrbc_amd as orbc
from _systems import contact_patch, pair_kinds

pytestmark = pytest.mark.gpu

EPS_MIXED = 1e-5
EPS_FP64 = 1e-11
DT = 0.01
MASS = np.array([1, 4, 4, 1, 10, 10], np.float64)


@pytest.fixture(scope="module", params=[0, 1])
def cps(request):
    return contact_patch(seed=request.param)


def make_sim(s, precision, **kw):
    sim = orbc.Simulation(precision=precision, **kw)
    t = s["type"]
    lip = np.where(t == 0)[0]
    prot = np.where(t != 0)[0]
    sim.set_state(orbc.LIPID, s["x"][lip], s["n"][lip], s["id"][lip], v=s["v"][lip], o=s["o"][lip])
    sim.set_state(orbc.PROTEIN, s["x"][prot], s["n"][prot], s["id"][prot], v=s["v"][prot], o=s["o"][prot],
                  type=t[prot])
    bt, bi, bj = s["bonds"]
    sim.set_bonds(bt, s["id"][bi], s["id"][bj])
    return sim, lip, prot


def gather_forces(sim, lip, prot, energy=False):
    n = len(lip) + len(prot)
    lf = sim.get_forces(orbc.LIPID, energy)
    pf = sim.get_forces(orbc.PROTEIN, energy)
    out = []
    for k in range(len(lf)):
        a = np.zeros((n, 3) if lf[k].ndim == 2 else (n,))
        a[lip] = lf[k]
        a[prot] = pf[k]
        out.append(a)
    return out


def gather_state(sim, lip, prot):
    n = len(lip) + len(prot)
    ls, ps = sim.get_state(orbc.LIPID), sim.get_state(orbc.PROTEIN)
    res = {}
    for k in ls:
        a = np.zeros((n, 3))
        a[lip] = ls[k]
        a[prot] = ps[k]
        res[k] = a
    return res


def id_pairs(s, pairs):
    ids = s["id"].astype(np.int64)
    p = np.sort(np.stack([ids[pairs[:, 0]], ids[pairs[:, 1]]], 1), axis=1)
    return p[np.lexsort((p[:, 1], p[:, 0]))]


def eps(a, b):
    return np.abs(a - b).max() / np.sqrt(np.mean(np.sum(b * b, 1)))


def test_contact_patch_has_every_pair_kind(cps):
    k = pair_kinds(cps["x"], cps["type"])
    for key in [(0, 4, "lj"), (0, 5, "lj"), (1, 4, "lj"), (1, 5, "lj"), (2, 4, "lj"), (2, 5, "lj"),
                (1, 1, "poly"), (1, 2, "poly"), (1, 3, "poly"), (2, 2, "poly"), (2, 3, "poly"), (3, 3, "poly")]:
        assert k.get(key, 0) > 0, key


@pytest.mark.parametrize("precision", [orbc.FP64, orbc.MIXED])
def test_contact_pair_set_bit_exact(cps, precision, oracle):
    s = cps
    sim, _, _ = make_sim(s, precision)
    got = sim.pairs()
    if precision == orbc.FP64:
        exp = id_pairs(s, oracle.pairs_brute(s["x"], s["type"]))
    else:
        exp = id_pairs(s, oracle.pairs_mixed(s["x"], s["type"], sim.observables()["anchor_h"]))
    assert len(got) == len(exp) > 0 and np.array_equal(got, exp)


@pytest.mark.parametrize("precision,tol", [(orbc.FP64, EPS_FP64), (orbc.MIXED, EPS_MIXED)])
def test_contact_forces_vs_oracle(cps, precision, tol, oracle):
    s = cps
    sim, lip, prot = make_sim(s, precision)
    sim.compute_pairwise(zero_first=True, energy=True)
    sim.compute_bonded(energy=True)
    f, t, e, w = gather_forces(sim, lip, prot, energy=True)
    of, ot, oe, ow = oracle.forces(s["x"], s["n"], s["type"], bonds=s["bonds"])
    assert eps(f, of) < tol and eps(t, ot) < tol
    # the WCA / protein-protein rows alone (proteins) against the same scale
    pr = s["type"] != 0
    assert np.abs(f[pr] - of[pr]).max() / np.sqrt(np.mean(np.sum(of * of, 1))) < tol
    assert np.abs(e - oe).max() / np.sqrt(np.mean(oe ** 2)) < 10 * tol
    assert np.abs(w - ow).max() / np.sqrt(np.mean(ow ** 2)) < 10 * tol


def _oracle_langevin(oracle, s, nsteps, seed, kT):
    st = {k: s[k].copy() for k in ("x", "v", "n", "o")}
    st.update(type=s["type"].astype(np.int32), id=s["id"].astype(np.uint32))
    F = []
    for step in range(nsteps):
        f, t, _, _ = oracle.forces(st["x"], st["n"], st["type"], bonds=s["bonds"])
        F.append((f, t))
        st["f"], st["t"] = f.copy(), t.copy()
        oracle.langevin(st, kT, 0.01, DT, seed, step)
    return st, F


@pytest.mark.parametrize("precision", [orbc.FP64, orbc.MIXED])
@pytest.mark.parametrize("block_halo", [0, 2])
@pytest.mark.parametrize("nsteps", [1, 2])
def test_contact_fused_langevin_vs_oracle(cps, precision, block_halo, nsteps, oracle):
    """orbc_run: the fused step kernel (block-halo walk with block_halo=2,
    32-bit gather walk with 0) against the oracle stepping verlet_langevin
    with its own forces and the same Philox noise."""
    s = cps
    seed, kT = 77, 0.22
    sim, lip, prot = make_sim(s, precision, seed=seed, kBT=kT, block_halo=block_halo, skin=0.4)
    sim.run(nsteps)
    obs = sim.observables()
    assert obs["step"] == nsteps
    assert (obs["halo_blocks"] > 0) == (block_halo == 2)
    got = gather_state(sim, lip, prot)
    st, F = _oracle_langevin(oracle, s, nsteps, seed, kT)
    m = MASS[s["type"]][:, None]
    rms_f = np.sqrt(np.mean(np.sum(F[-1][0] ** 2, 1)))
    rms_t = np.sqrt(np.mean(np.sum(F[-1][1] ** 2, 1)))
    dF = np.abs(got["v"] - st["v"]).max(1, keepdims=True) * m / DT
    dT = np.abs(got["o"] - st["o"]).max(1) / DT
    if precision == orbc.FP64:
        for k in ("x", "v", "n", "o"):
            assert np.abs(got[k] - st[k]).max() < 1e-10 * max(1.0, np.abs(st[k]).max()), k
    else:
        assert dF.max() / rms_f < nsteps * EPS_MIXED
        assert dT.max() / rms_t < nsteps * EPS_MIXED
        assert np.abs(got["x"] - st["x"]).max() < nsteps * DT * DT * EPS_MIXED * rms_f
        assert np.abs(got["n"] - st["n"]).max() < 1e-6


@pytest.mark.parametrize("precision", [orbc.FP64, orbc.MIXED])
@pytest.mark.parametrize("block_halo", [0, 2])
def test_contact_nose_hoover_vs_oracle(cps, precision, block_halo, oracle):
    """orbc_run with the Nose-Hoover thermostat: the fused second half
    (k_step_nh_second[_halo], forces and post_toque_final_update in one pass)
    and the first half against the oracle's fused NH halves
    (src/integrate_nh.h:178-273), two steps."""
    s = cps
    n = len(s["type"])
    Q, kT, zeta = 0.01 * n, 0.22, 0.0
    sim, lip, prot = make_sim(s, precision, thermostat=orbc.RUN_NOSE_HOOVER, Q=Q, kBT=kT, block_halo=block_halo,
                              skin=0.4)
    nsteps = 2
    sim.run(nsteps)
    got = gather_state(sim, lip, prot)
    st = {k: s[k].copy() for k in ("x", "v", "n", "o")}
    st.update(type=s["type"].astype(np.int32))
    F = []

    def forces():
        f, t, _, _ = oracle.forces(st["x"], st["n"], st["type"], bonds=s["bonds"])
        F.append(f)
        st["f"], st["t"] = f, t

    forces()
    ke = oracle.nh_second(st, DT, zeta)
    zeta = oracle.nh_zeta(zeta, DT, Q, ke, n, kT)
    for _ in range(nsteps):
        ke = oracle.nh_first(st, DT, zeta)
        zeta = oracle.nh_zeta(zeta, DT, Q, ke, n, kT)
        forces()
        ke = oracle.nh_second(st, DT, zeta)
        zeta = oracle.nh_zeta(zeta, DT, Q, ke, n, kT)
    if precision == orbc.FP64:
        for k in ("x", "v", "n", "o"):
            assert np.abs(got[k] - st[k]).max() < 1e-10 * max(1.0, np.abs(st[k]).max()), k
        assert abs(sim.observables()["zeta"] - zeta) < 1e-10 * max(1.0, abs(zeta))
    else:
        m = MASS[s["type"]][:, None]
        rms_f = np.sqrt(np.mean(np.sum(F[-1] ** 2, 1)))
        nev = nsteps + 1                          # force evaluations behind the final velocities
        dF = np.abs(got["v"] - st["v"]).max(1, keepdims=True) * m / DT
        assert dF.max() / rms_f < nev * EPS_MIXED
        assert abs(sim.observables()["zeta"] - zeta) < 1e-6 * max(1.0, abs(zeta))


def _oracle_minimize(oracle, s, nsteps, dr=0.05, dn=0.05):
    st = {k: s[k].copy() for k in ("x", "n")}
    st["type"] = s["type"].astype(np.int32)
    mass = MASS[s["type"]]
    margin = np.inf
    for _ in range(nsteps):
        f, t, _, _ = oracle.forces(st["x"], st["n"], st["type"], bonds=s["bonds"])
        tp = np.cross(st["n"], t)                                   # post_torque
        a = np.linalg.norm(f, axis=1) / mass
        b = np.linalg.norm(np.cross(tp, st["n"]), axis=1)
        # distance of every particle from the cap thresholds (the step jumps there)
        margin = min(margin, np.min(np.abs(a / dr - 1)), np.min(np.abs(b / dn - 1)))
        st["f"], st["t"] = f, tp
        oracle.minimize_move(st, dr, dn, DT)
    return st, margin


@pytest.mark.parametrize("precision", [orbc.FP64, orbc.MIXED])
def test_contact_minimize_vs_oracle(cps, precision, oracle):
    """orbc_minimize (src/openrbc.cpp:88-146: forces, post_torque, capped
    steepest-descent move, bounce_back) against the oracle's move
    (orc_minimize_move, pinned by the reference program's frames).  FP64:
    three steps to rounding.  Mixed: one step; a capped move is 0.05 along
    f / |f| (or rotates the director by 0.05), so an FP32 force error of
    eps_F rms |F| turns it by at most that over |f| > 0.05 m (|n x t x n| >
    0.05): below 0.05 * 5e-3 for this patch's rms force and director force."""
    s = cps
    nsteps = 3 if precision == orbc.FP64 else 1
    sim, lip, prot = make_sim(s, precision)
    sim.minimize(nsteps)
    got = gather_state(sim, lip, prot)
    st, margin = _oracle_minimize(oracle, s, nsteps)
    assert margin > 1e-3        # no particle sits on a cap threshold, where FP32 could pick the other branch
    moved = np.abs(st["x"] - s["x"]).max()
    assert moved > 0.04
    if precision == orbc.FP64:
        assert np.abs(got["x"] - st["x"]).max() < 1e-10 * max(1.0, np.abs(st["x"]).max())
        assert np.abs(got["n"] - st["n"]).max() < 1e-10
    else:
        f0, t0, _, _ = oracle.forces(s["x"], s["n"], s["type"], bonds=s["bonds"])
        rel_f = EPS_MIXED * np.sqrt(np.mean(np.sum(f0 * f0, 1))) / 0.05
        rel_t = EPS_MIXED * np.sqrt(np.mean(np.sum(t0 * t0, 1))) / 0.05
        bound = 0.05 * 2 * max(rel_f, rel_t)
        assert bound < 1e-3
        assert np.abs(got["x"] - st["x"]).max() < bound
        assert np.abs(got["n"] - st["n"]).max() < bound


@pytest.mark.parametrize("precision,tol", [(orbc.FP64, EPS_FP64), (orbc.MIXED, EPS_MIXED)])
def test_reupload_permuted_same_count(cps, precision, tol, oracle):
    """An adapter that re-uploads its containers after the reference has
    permuted them (reorder by cell, src/reorder.h:73-149) keeps the count but
    changes which tag and protein type sits at each position: the device
    must follow the new tags (bonds by tag, noise keys) and types
    (force-field rows)."""
    s = cps
    sim, lip, prot = make_sim(s, precision)
    sim.compute_pairwise(zero_first=True)           # the store is live
    rng = np.random.default_rng(9)
    pl, pp = rng.permutation(len(lip)), rng.permutation(len(prot))
    L, P = lip[pl], prot[pp]
    sim.set_state(orbc.LIPID, s["x"][L], s["n"][L], s["id"][L], v=s["v"][L], o=s["o"][L])
    sim.set_state(orbc.PROTEIN, s["x"][P], s["n"][P], s["id"][P], v=s["v"][P], o=s["o"][P], type=s["type"][P])
    sim.compute_pairwise(zero_first=True)
    sim.compute_bonded()
    f, t = gather_forces(sim, L, P)
    of, ot, _, _ = oracle.forces(s["x"], s["n"], s["type"], bonds=s["bonds"])
    assert eps(f, of) < tol and eps(t, ot) < tol


def test_nose_hoover_mass_fixed_at_first_use():
    """cfg.Q = 0 resolves to 0.01 N at the first Nose-Hoover step and keeps
    that value after the stray-lipid cleanup shrinks N (the reference sets Q
    once, src/runtime_parameter.h:76): the run equals one with Q = 0.01 N0
    given explicitly."""
    sphere = orbc.random_sphere(r=12.0, seed=2)
    rng = np.random.default_rng(102)
    idx = rng.choice(len(sphere["lipid_x"]), 12, replace=False)
    sphere["lipid_x"][idx] *= 16.0 / 12.0
    n0 = len(sphere["lipid_x"])
    res = []
    for Q in (0.0, 0.01 * n0):
        sim = orbc.Simulation(precision=orbc.FP64, thermostat=orbc.RUN_NOSE_HOOVER, Q=Q, skin=0.3, freq_cleanup=0)
        orbc.system_to_simulation(sphere, sim, kT=0.22, seed=4)
        sim.run(5)
        sim.voronoi_init(0, 4)
        assert sim.delete_stray_lipids(2.5) > 0
        sim.run(5)
        res.append((sim.get_state(orbc.LIPID)["v"], sim.observables()["zeta"]))
    assert np.array_equal(res[0][0], res[1][0]) and res[0][1] == res[1][1]


@pytest.mark.parametrize("precision,tol", [(orbc.FP64, EPS_FP64), (orbc.MIXED, EPS_MIXED)])
def test_reupload_other_system_into_live_context(precision, tol, oracle):
    """A live context that receives a different system (other counts) drops
    its device store and builds a new one: cell ranges, lists, block halos
    and bond lists are allocated for the new system, not taken from the freed
    one.  Three systems in a row, the last the same as the first; forces
    against the oracle after each."""
    sim = orbc.Simulation(precision=precision, block_halo=2)
    for seed in (0, 1, 0):
        s = contact_patch(seed=seed)
        t = s["type"]
        lip = np.where(t == 0)[0]
        prot = np.where(t != 0)[0]
        sim.set_state(orbc.LIPID, s["x"][lip], s["n"][lip], s["id"][lip], v=s["v"][lip], o=s["o"][lip])
        sim.set_state(orbc.PROTEIN, s["x"][prot], s["n"][prot], s["id"][prot], v=s["v"][prot], o=s["o"][prot],
                      type=t[prot])
        bt, bi, bj = s["bonds"]
        sim.set_bonds(bt, s["id"][bi], s["id"][bj])
        sim.compute_pairwise(zero_first=True)
        sim.compute_bonded()
        f, tq = gather_forces(sim, lip, prot)
        of, ot, _, _ = oracle.forces(s["x"], s["n"], s["type"], bonds=s["bonds"])
        assert eps(f, of) < tol and eps(tq, ot) < tol
        sim.run(3)
        assert np.isfinite(gather_state(sim, lip, prot)["x"]).all()

"""GPU parity: liborbc_hip.so (through the C ABI) against the oracle.

Bars: pair sets bit-exact (same selection arithmetic as the device: FP64, or
FP32 anchor-relative for the mixed path); FP64 forces within 1e-11 of the
oracle (relative to the rms force); mixed forces within eps_F <= 1e-5
(max |dF| / rms |F|, the metric of the force-error target); integrators within
FP64 rounding.
"""
import numpy as np
import pytest

import openrbc_amd as orbc
from _systems import patch

pytestmark = pytest.mark.gpu

EPS_MIXED = 1e-5      # eps_F target for FP32 forces (BASELINE force-error definition)
EPS_FP64 = 1e-11


def make_sim(s, precision, **kw):
    sim = orbc.Simulation(precision=precision, **kw)
    t = s["type"]
    lip = np.where(t == 0)[0]
    prot = np.where(t != 0)[0]
    sim.set_state(orbc.LIPID, s["x"][lip], s["n"][lip], s["id"][lip], v=s["v"][lip], o=s["o"][lip])
    sim.set_state(orbc.PROTEIN, s["x"][prot], s["n"][prot], s["id"][prot], v=s["v"][prot], o=s["o"][prot],
                  type=t[prot])
    bt, bi, bj = s["bonds"]
    if len(bt):
        sim.set_bonds(bt, s["id"][bi], s["id"][bj])
    return sim, lip, prot


def gather(sim, lip, prot, what="forces", energy=False):
    n = len(lip) + len(prot)
    if what == "forces":
        lf = sim.get_forces(orbc.LIPID, energy)
        pf = sim.get_forces(orbc.PROTEIN, energy)
        out = []
        for k in range(len(lf)):
            if lf[k] is None:
                out.append(None)
                continue
            shape = (n, 3) if lf[k].ndim == 2 else (n,)
            a = np.zeros(shape)
            a[lip] = lf[k]
            a[prot] = pf[k]
            out.append(a)
        return out
    ls = sim.get_state(orbc.LIPID)
    ps = sim.get_state(orbc.PROTEIN)
    res = {}
    for k in ls:
        a = np.zeros((n, 3))
        a[lip] = ls[k]
        a[prot] = ps[k]
        res[k] = a
    return res


def id_pairs(s, pairs):
    ids = s["id"].astype(np.int64)
    p = np.stack([ids[pairs[:, 0]], ids[pairs[:, 1]]], 1)
    p = np.sort(p, axis=1)
    return p[np.lexsort((p[:, 1], p[:, 0]))]


def eps_f(a, b):
    return np.abs(a - b).max() / np.sqrt(np.mean(np.sum(b * b, 1)))


@pytest.fixture(scope="module", params=[0, 5])
def system(request):
    return patch(16, 16, seed=request.param)


@pytest.mark.parametrize("precision", [orbc.FP64, orbc.MIXED])
def test_pair_set_bit_exact(system, precision, oracle):
    s = system
    sim, lip, prot = make_sim(s, precision)
    got = sim.pairs()
    if precision == orbc.FP64:
        exp = id_pairs(s, oracle.pairs_brute(s["x"], s["type"]))
    else:
        h = sim.observables()["anchor_h"]
        exp = id_pairs(s, oracle.pairs_mixed(s["x"], s["type"], h))
    assert len(got) == len(exp) > 0
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("precision,tol", [(orbc.FP64, EPS_FP64), (orbc.MIXED, EPS_MIXED)])
def test_forces_vs_oracle(system, precision, tol, oracle):
    s = system
    sim, lip, prot = make_sim(s, precision)
    sim.compute_pairwise(zero_first=True, energy=True)
    sim.compute_bonded(energy=True)
    f, t, e, w = gather(sim, lip, prot, energy=True)
    of, ot, oe, ow = oracle.forces(s["x"], s["n"], s["type"], bonds=s["bonds"])
    assert eps_f(f, of) < tol
    assert eps_f(t, ot) < tol
    scale = np.sqrt(np.mean(oe ** 2))
    assert np.abs(e - oe).max() / scale < 10 * tol
    assert np.abs(w - ow).max() / np.sqrt(np.mean(ow ** 2)) < 10 * tol


def test_bonded_only(system, oracle):
    s = system
    sim, lip, prot = make_sim(s, orbc.FP64)
    sim.compute_bonded(zero_first=True)
    f, t = gather(sim, lip, prot)
    of, _, _ = oracle.bonded(s["x"], s["bonds"])
    assert np.abs(f - of).max() <= 1e-12 * max(1.0, np.abs(of).max())
    assert np.abs(t).max() == 0.0


@pytest.mark.parametrize("precision", [orbc.FP64, orbc.MIXED])
@pytest.mark.parametrize("kT", [0.0, 0.22])
def test_langevin_step_vs_oracle(system, precision, kT, oracle):
    """integrate(verlet_langevin) with given f,t: FP64 update, Philox noise."""
    s = system
    n = len(s["type"])
    rng = np.random.default_rng(7)
    f = rng.normal(0, 3, (n, 3)); t = rng.normal(0, 1, (n, 3))
    seed = 12345
    sim, lip, prot = make_sim(s, precision, kBT=kT, seed=seed)
    sim.set_forces(orbc.LIPID, f[lip], t[lip])
    sim.set_forces(orbc.PROTEIN, f[prot], t[prot])
    if precision == orbc.MIXED:      # the FP32 accumulators hold f,t rounded to FP32
        f = f.astype(np.float32).astype(np.float64); t = t.astype(np.float32).astype(np.float64)
    sim.integrate("verlet_langevin")
    got = gather(sim, lip, prot, "state")
    st = {k: s[k].copy() for k in ("x", "v", "n", "o")}
    st.update(f=f.copy(), t=t.copy(), type=s["type"].astype(np.int32), id=s["id"].astype(np.uint32))
    oracle.langevin(st, kT, 0.01, 0.01, seed, 0)
    for k in ("x", "v", "n", "o"):
        assert np.abs(got[k] - st[k]).max() < 1e-12, k
    ff, tt = gather(sim, lip, prot)
    assert np.abs(ff).max() == 0 and np.abs(tt).max() == 0


def test_nose_hoover_halves_vs_oracle(system, oracle):
    s = system
    n = len(s["type"])
    rng = np.random.default_rng(3)
    f = rng.normal(0, 3, (n, 3)); t = rng.normal(0, 1, (n, 3))
    sim, lip, prot = make_sim(s, orbc.FP64, zeta=0.05, Q=7.0, kBT=0.22)
    sim.set_forces(orbc.LIPID, f[lip], t[lip]); sim.set_forces(orbc.PROTEIN, f[prot], t[prot])
    sim.integrate("post_toque_final_update")
    st = {k: s[k].copy() for k in ("x", "v", "n", "o")}
    st.update(f=f.copy(), t=t.copy(), type=s["type"].astype(np.int32))
    ke = oracle.nh_second(st, 0.01, 0.05)
    zeta = oracle.nh_zeta(0.05, 0.01, 7.0, ke, n, 0.22)
    got = gather(sim, lip, prot, "state")
    for k in ("v", "o"):
        assert np.abs(got[k] - st[k]).max() < 1e-12
    assert abs(sim.observables()["zeta"] - zeta) < 1e-12 * max(1, abs(zeta))
    sim.integrate("verlet_initial_bounce_clearforce_update")
    ke = oracle.nh_first(st, 0.01, zeta)
    zeta = oracle.nh_zeta(zeta, 0.01, 7.0, ke, n, 0.22)
    got = gather(sim, lip, prot, "state")
    for k in ("x", "v", "n", "o"):
        assert np.abs(got[k] - st[k]).max() < 1e-12, k
    assert abs(sim.observables()["zeta"] - zeta) < 1e-12 * max(1, abs(zeta))


def test_temperature(system, oracle):
    s = system
    sim, lip, prot = make_sim(s, orbc.MIXED)
    T = sim.compute_temperature()
    assert abs(T - oracle.temperature(s["v"], s["type"])) < 1e-12


def test_radix_sort_stable():
    sim = orbc.Simulation()
    rng = np.random.default_rng(0)
    for n, bits in ((1, 8), (1000, 5), (70001, 19), (300000, 32)):
        keys = rng.integers(0, 1 << bits, n, dtype=np.uint64).astype(np.uint32)
        vals = np.arange(n, dtype=np.uint32)
        k, v = sim.sort_pairs(keys, vals, bits)
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(v, order)
        assert np.array_equal(k, keys[order])


def test_morton_keys_reference(oracle):
    sim = orbc.Simulation()
    rng = np.random.default_rng(1)
    pts = rng.uniform(-600, 600, (5000, 3)).astype(np.float32)
    assert np.array_equal(sim.morton_keys(pts), oracle.morton_ref(pts))


def test_nearest_and_within_brute(oracle):
    sim = orbc.Simulation()
    rng = np.random.default_rng(2)
    pts = rng.uniform(-50, 50, (3000, 3)).astype(np.float32)
    q = rng.uniform(-60, 60, (2000, 3)).astype(np.float32)
    idx = sim.nearest(pts, q)
    assert np.array_equal(idx, oracle.partition(q.astype(np.float64), pts))
    got = sim.within(pts, q[:300], 9.0)
    for k in range(300):
        d = ((pts - q[k]) ** 2)
        d2 = (d[:, 0] + d[:, 1]) + d[:, 2]
        exp = np.where(d2 < np.float32(81.0))[0]
        assert np.array_equal(got[k], exp)


def test_langevin_run_tracks_oracle(oracle):
    """orbc_run (fused step kernel, FP64) vs the oracle stepping the same
    pairs with the same counter-based noise: trajectories agree to roundoff."""
    s = patch(12, 12, seed=11)
    seed = 99
    sim, lip, prot = make_sim(s, orbc.FP64, seed=seed)
    st = {k: s[k].copy() for k in ("x", "v", "n", "o")}
    n = len(s["type"])
    st.update(type=s["type"].astype(np.int32), id=s["id"].astype(np.uint32))
    nsteps = 10
    sim.run(nsteps)
    for step in range(nsteps):
        f, t, _, _ = oracle.forces(st["x"], st["n"], st["type"], bonds=s["bonds"])
        st["f"], st["t"] = f, t
        oracle.langevin(st, 0.22, 0.01, 0.01, seed, step)
    got = gather(sim, lip, prot, "state")
    assert np.abs(got["x"] - st["x"]).max() < 1e-9
    assert np.abs(got["n"] - st["n"]).max() < 1e-9
    assert sim.observables()["step"] == nsteps


@pytest.mark.parametrize("precision", [orbc.FP64, orbc.MIXED])
def test_run_with_skin_rebuilds(precision, oracle):
    """A small skin forces rebuilds inside the batched run (device-side skin
    check + sticky stop flag); the trajectory must still follow the oracle
    stepping the exact pair set."""
    s = patch(10, 10, seed=4)
    seed = 5
    sim, lip, prot = make_sim(s, precision, seed=seed, skin=0.03)
    st = {k: s[k].copy() for k in ("x", "v", "n", "o")}
    st.update(type=s["type"].astype(np.int32), id=s["id"].astype(np.uint32))
    nsteps = 70
    sim.run(nsteps)
    for step in range(nsteps):
        f, t, _, _ = oracle.forces(st["x"], st["n"], st["type"], bonds=s["bonds"])
        st["f"], st["t"] = f, t
        oracle.langevin(st, 0.22, 0.01, 0.01, seed, step)
    got = gather(sim, lip, prot, "state")
    tol = 1e-8 if precision == orbc.FP64 else 1e-3
    assert np.abs(got["x"] - st["x"]).max() < tol
    obs = sim.observables()
    assert obs["step"] == nsteps
    assert obs["rebuilds"] > 3

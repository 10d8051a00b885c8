"""GPU parity: liborbc_hip.so (through the C ABI) against the oracle.

Bars: pair sets bit-exact (same selection arithmetic as the device: FP64, or
FP32 anchor-relative for the mixed path); FP64 forces within 1e-11 of the
oracle (relative to the rms force); mixed forces within eps_F <= 1e-5
(max |dF| / rms |F|, the metric of the force-error target); integrators within
FP64 rounding.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

import openrbc_amd as orbc
from _systems import patch, periodic_patch

pytestmark = pytest.mark.gpu

EPS_MIXED = 1e-5      # eps_F target of the mixed path: max |dF| / rms |F| against FP64
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
def test_drifting_lipids_keep_the_list_valid(precision, oracle):
    """Lipids without list entries get a skin budget from their distance to
    everything they could interact with (k_budget): one lipid leaves the sheet
    fast, another comes in fast from above and ends up interacting.  The FP64
    run must still follow the oracle stepping the exact pair set every step,
    and the fast lipids must not force a rebuild every few steps."""
    base = patch(10, 10, seed=4)
    s = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in base.items()}
    c = base["x"][base["type"] == 0].mean(0)
    extra_x = np.array([c + [0.0, 0.0, 5.0], c + [3.0, 2.0, 4.0]])
    extra_v = np.array([[0.0, 0.0, 3.0], [0.0, 0.0, -2.0]])
    s["x"] = np.vstack([base["x"], extra_x])
    s["n"] = np.vstack([base["n"], [[0, 0, 1.0], [0, 0, 1.0]]])
    s["v"] = np.vstack([base["v"], extra_v])
    s["o"] = np.vstack([base["o"], np.zeros((2, 3))])
    s["type"] = np.concatenate([base["type"], [0, 0]]).astype(np.int32)
    s["id"] = np.concatenate([base["id"], [base["id"].max() + 1, base["id"].max() + 2]]).astype(np.uint32)
    seed = 5
    nsteps = 150
    runs = {}
    for name, sysd in (("drift", s), ("base", base)):
        sim, lip, prot = make_sim(sysd, precision, seed=seed, skin=0.4)
        sim.run(nsteps)
        runs[name] = (gather(sim, lip, prot, "state"), sim.observables(), sim, lip, prot)
    got, obs, sim, lip, prot = runs["drift"]
    # the leaving lipid moves 0.03 per step, past skin/2 = 0.2 every 7 steps
    assert obs["rebuilds"] <= runs["base"][1]["rebuilds"] + 6, (obs["rebuilds"], runs["base"][1]["rebuilds"])
    if precision == orbc.FP64:
        st = {k: s[k].copy() for k in ("x", "v", "n", "o")}
        st.update(type=s["type"], id=s["id"])
        touched = False
        for step in range(nsteps):
            f, t, _, _ = oracle.forces(st["x"], st["n"], st["type"], bonds=s["bonds"])
            touched = touched or np.abs(f[-1]).max() > 0
            st["f"], st["t"] = f, t
            oracle.langevin(st, 0.22, 0.01, 0.01, seed, step)
        assert touched, "the incoming lipid never reached the sheet"
        assert np.abs(got["x"] - st["x"]).max() < 1e-8
    # (mixed: the same budgets; its trajectory separates chaotically from the
    # FP64 oracle's, so only the rebuild count is checked)


@pytest.mark.parametrize("precision", [orbc.FP64, orbc.MIXED])
def test_far_lipids_get_wide_budgets(precision, oracle):
    """A lipid far above the sheet (nothing within the type's two rings of
    cells) gets its budget from the widened cube of cells: one falls from 17
    above at 0.1 per step and lands on the sheet, one crosses high above it.
    The FP64 run must follow the oracle stepping the exact pair set (a budget
    too wide would miss the landing), and the two lipids may cost only a few
    rebuilds."""
    base = patch(10, 10, seed=4)
    s = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in base.items()}
    c = base["x"][base["type"] == 0].mean(0)
    extra_x = np.array([c + [0.0, 0.0, 17.0], c + [-20.0, 0.0, 25.0]])
    extra_v = np.array([[0.0, 0.0, -10.0], [8.0, 0.0, 0.0]])
    s["x"] = np.vstack([base["x"], extra_x])
    s["n"] = np.vstack([base["n"], [[0, 0, 1.0], [0, 0, 1.0]]])
    s["v"] = np.vstack([base["v"], extra_v])
    s["o"] = np.vstack([base["o"], np.zeros((2, 3))])
    s["type"] = np.concatenate([base["type"], [0, 0]]).astype(np.int32)
    s["id"] = np.concatenate([base["id"], [base["id"].max() + 1, base["id"].max() + 2]]).astype(np.uint32)
    seed = 5
    nsteps = 190          # the falling lipid reaches the sheet at step ~160
    runs = {}
    for name, sysd in (("far", s), ("base", base)):
        sim, lip, prot = make_sim(sysd, precision, seed=seed, skin=0.4)
        sim.run(nsteps)
        runs[name] = (gather(sim, lip, prot, "state"), sim.observables())
    got, obs = runs["far"]
    # widened cubes: the falling lipid halves its gap at each rebuild (17 -> 9.8
    # -> 6.2 -> 4.4 -> 3.5 -> 3.1: 6 rebuilds), the crossing one needs one.
    # Two rings only (budget <= 1.7) would cost about 18.
    assert obs["rebuilds"] <= runs["base"][1]["rebuilds"] + 9, (obs["rebuilds"], runs["base"][1]["rebuilds"])
    if precision == orbc.FP64:
        st = {k: s[k].copy() for k in ("x", "v", "n", "o")}
        st.update(type=s["type"], id=s["id"])
        touched = False
        for step in range(nsteps):
            f, t, _, _ = oracle.forces(st["x"], st["n"], st["type"], bonds=s["bonds"])
            touched = touched or np.abs(f[-2]).max() > 0
            st["f"], st["t"] = f, t
            oracle.langevin(st, 0.22, 0.01, 0.01, seed, step)
        assert touched, "the falling lipid never reached the sheet"
        assert np.abs(got["x"] - st["x"]).max() < 1e-8


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


# ---- Voronoi cells, stencil pair set, constrain_volume ---------------------------------
def _voro_sim(s, mode=orbc.NEIGHBOR_EXACT, precision=orbc.FP64):
    sim, lip, prot = make_sim(s, precision, neighbor_mode=mode)
    rng = np.random.default_rng(1)
    lipx = s["x"][lip]
    c = lipx[rng.choice(len(lipx), max(1, len(lipx) // 14), replace=False)].astype(np.float32)
    return sim, lip, prot, c


def _aff(sim, lip, prot, n):
    _, al, ap = sim.voronoi_get()
    a = np.zeros(n, np.int32)
    a[lip] = al
    a[prot] = ap
    return a


def test_voronoi_partition_and_centroids(system, oracle):
    s = system
    sim, lip, prot, c = _voro_sim(s)
    sim.voronoi_set_centroids(c)
    aff = _aff(sim, lip, prot, len(s["type"]))
    assert np.array_equal(aff, oracle.partition(s["x"], c))
    sim.voronoi_update()              # centroids <- member means, then re-partition
    c2, _, _ = sim.voronoi_get()
    member = aff.copy()
    member[s["type"] != 0] = -1
    exp = c.copy()
    for k in range(len(c)):
        m = member == k
        if m.any():
            exp[k] = s["x"][m].mean(0)
    assert np.abs(c2 - exp).max() < 2e-6 * max(1.0, np.abs(exp).max())
    aff2 = _aff(sim, lip, prot, len(s["type"]))
    assert np.array_equal(aff2, oracle.partition(s["x"], c2))


def test_voronoi_lloyd_init_is_consistent(system, oracle):
    s = system
    sim, lip, prot, _ = _voro_sim(s)
    sim.voronoi_init(0, 8)
    c, _, _ = sim.voronoi_get()
    assert len(c) == len(lip) // 14
    aff = _aff(sim, lip, prot, len(s["type"]))
    assert np.array_equal(aff, oracle.partition(s["x"], c))


@pytest.mark.parametrize("precision", [orbc.FP64, orbc.MIXED])
def test_voronoi_stencil_pair_set(system, precision, oracle):
    """ORBC_NEIGHBOR_VORONOI: the reference's stencil pair set (centroid
    distance < 6 / 8 / 9 for ll / pl / pp), src/compute_pairwise_fused.h:253-316."""
    s = system
    sim, lip, prot, c = _voro_sim(s, orbc.NEIGHBOR_VORONOI, precision)
    sim.voronoi_set_centroids(c)
    aff = _aff(sim, lip, prot, len(s["type"]))
    got = sim.pairs()
    if precision == orbc.FP64:
        exp = id_pairs(s, oracle.pairs_voronoi(s["x"], s["type"], aff, c))
        assert len(got) > 0 and np.array_equal(got, exp)
        full = id_pairs(s, oracle.pairs_brute(s["x"], s["type"]))
    else:
        pm = oracle.pairs_mixed(s["x"], s["type"], sim.observables()["anchor_h"])
        full = id_pairs(s, pm)
        # the stencil test of the neighbour build in FP32: the squared
        # distance of the two cells' FP32 centroids, summed unfused, below
        # 6^2 / 8^2 / 9^2 for lipid-lipid / one lipid / no lipid
        c32 = np.asarray(c, np.float32)
        d = c32[aff[pm[:, 1]]] - c32[aff[pm[:, 0]]]
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        nl = (s["type"][pm[:, 0]] == 0).astype(int) + (s["type"][pm[:, 1]] == 0)
        R = np.where(nl == 2, 6.0, np.where(nl == 1, 8.0, 9.0)).astype(np.float32)
        exp = id_pairs(s, pm[d2 < R * R])
        assert len(got) > 0 and np.array_equal(got, exp)
    # the stencil keeps only in-cutoff pairs, and may drop some of them
    fs = set(map(tuple, full.tolist()))
    assert 0 < len(got) <= len(full) and all(tuple(p) in fs for p in got.tolist())


def test_constrain_volume_vs_oracle(system, oracle):
    s = system
    sim, lip, prot, c = _voro_sim(s)
    sim.voronoi_set_centroids(c)
    aff = _aff(sim, lip, prot, len(s["type"]))
    sim.compute_pairwise(zero_first=True)
    f0, _ = gather(sim, lip, prot)
    V = sim.constrain_volume(3.0, 0.05)
    f1, _ = gather(sim, lip, prot)
    fo = np.zeros_like(f0)
    Vo = oracle.constrain_volume(s["n"], s["type"], aff, c, 3.0, 0.05, fo)
    assert abs(V - Vo) < 1e-10 * max(1.0, abs(Vo))
    assert np.abs((f1 - f0) - fo).max() < 1e-10 * max(1.0, np.abs(fo).max())


def test_volume_of_a_lipid_sphere():
    """Cone-sum volume of a lipid sphere (radial directors) against the closed form."""
    sphere = orbc.random_sphere(r=40.0, seed=3)
    sim = orbc.Simulation(precision=orbc.FP64)
    orbc.system_to_simulation(sphere, sim)
    sim.voronoi_init(0, 6)
    V = sim.constrain_volume(1.0, 0.0)
    # every cell contributes h * count * (pi 1.26 / 4) / 3 * 1e-6 with h ~ r (cap sagitta < 0.1 %)
    n = len(sphere["lipid_x"])
    exp = 40.0 * n * 3.1415926 * 1.26 / 12.0 * 1e-6
    assert abs(V / exp - 1) < 0.01


# ---- C1: icosphere level 3 through the initialiser (225,384 particles) -----------------
@pytest.fixture(scope="module")
def c1():
    v, e, f = orbc.icosphere(3)
    s = orbc.init_rbc(v, e, f, seed=1)
    rng = np.random.default_rng(1)
    from openrbc_amd.systems import tilt_directors
    ln = tilt_directors(s["lipid_n"], 0.01, rng)
    x = np.vstack([s["lipid_x"], s["protein_x"]])
    n = np.vstack([ln, s["protein_n"]])
    t = np.concatenate([np.zeros(len(s["lipid_x"]), np.int32), s["protein_type"]])
    ids = np.concatenate([s["lipid_id"], s["protein_id"]])
    idx = {int(k): i for i, k in enumerate(ids)}
    bi = np.array([idx[int(k)] for k in s["bond_i"]]); bj = np.array([idx[int(k)] for k in s["bond_j"]])
    return dict(x=x, n=n, type=t, id=ids, v=np.zeros_like(x), o=np.zeros_like(x),
                bonds=(s["bond_type"], bi, bj))


def test_c1_cell_sums_stored_and_atomic_agree(c1, oracle):
    """Per-cell sums over 110 blocks of 2048 particles: the first volume
    evaluation after a partition adds every block's part atomically and
    records the cell counts; later ones store the sums of the cells a block
    holds whole.  Both must give the oracle's volume and push, and the
    centroid update that follows (stored sums too) the member means."""
    s = c1
    sim, lip, prot, c = _voro_sim(s)
    sim.voronoi_set_centroids(c)
    aff = _aff(sim, lip, prot, len(s["type"]))
    fo = np.zeros_like(s["x"])
    Vo = oracle.constrain_volume(s["n"], s["type"], aff, c, 3.0, 0.05, fo)
    sim.compute_pairwise(zero_first=True)
    f_prev, _ = gather(sim, lip, prot)
    for call in range(3):            # 0: atomic after the partition; 1, 2: counts current
        V = sim.constrain_volume(3.0, 0.05)
        f, _ = gather(sim, lip, prot)
        assert abs(V - Vo) < 1e-10 * max(1.0, abs(Vo)), call
        assert np.abs((f - f_prev) - fo).max() < 1e-10 * max(1.0, np.abs(fo).max()), call
        f_prev = f
    sim.voronoi_update()
    c2, _, _ = sim.voronoi_get()
    member = np.where(s["type"] == 0, aff, -1)
    cnt = np.bincount(member[member >= 0], minlength=len(c))
    exp = c.astype(np.float64).copy()
    for d in range(3):
        sx = np.bincount(member[member >= 0], weights=s["x"][member >= 0, d], minlength=len(c))
        exp[cnt > 0, d] = sx[cnt > 0] / cnt[cnt > 0]
    assert np.abs(c2 - exp).max() < 2e-6 * max(1.0, np.abs(exp).max())


@pytest.mark.parametrize("precision,tol", [(orbc.FP64, EPS_FP64), (orbc.MIXED, 5e-5)])
def test_c1_forces_vs_oracle(c1, precision, tol, oracle):
    """C1 as the initialiser leaves it: lipids placed at random 0.55 apart, so
    pair terms reach thousands while the rms force is tens, and the FP32
    rounding of the mirrors is amplified accordingly (mixed: 5e-5 here; the
    relaxed state below is held to the 1e-5 target)."""
    s = c1
    sim, lip, prot = make_sim(s, precision)
    sim.compute_pairwise(zero_first=True)
    sim.compute_bonded()
    f, t = gather(sim, lip, prot)
    pairs = oracle.pairs_cell(s["x"], s["type"])
    got = sim.pairs()
    if precision == orbc.FP64:
        assert np.array_equal(got, id_pairs(s, pairs))
    else:
        # FP32 selection may flip pairs that sit on a cutoff to within FP32 rounding
        e = set(map(tuple, id_pairs(s, pairs).tolist()))
        g = set(map(tuple, got.tolist()))
        flips = np.array(sorted(e ^ g), np.int64).reshape(-1, 2)
        assert len(flips) <= 1e-5 * len(e)
        order = np.argsort(s["id"])
        to_idx = lambda q: order[np.searchsorted(s["id"][order], q)]
        if len(flips):
            a, b = to_idx(flips[:, 0]), to_idx(flips[:, 1])
            r = np.linalg.norm(s["x"][a] - s["x"][b], axis=1)
            ff = oracle.ff()
            ta, tb = s["type"][a], s["type"][b]
            cuts = np.stack([np.sqrt(ff["cut2_poly"][ta, tb]), np.sqrt(ff["cut2_lj"][ta, tb])], 1)
            assert np.all(np.min(np.abs(r[:, None] / cuts - 1), 1) < 1e-5)
        pairs = np.stack([to_idx(got[:, 0]), to_idx(got[:, 1])], 1)
    of, ot, _, _ = oracle.forces(s["x"], s["n"], s["type"], pairs=pairs, bonds=s["bonds"])
    assert eps_f(f, of) < tol
    assert eps_f(t, ot) < tol


def _rounding_bounds(sim, nsteps, dt=0.01):
    """Mixed precision: the fused step sums the pair terms in FP64 and uses
    them at once, the per-call sequence stores them in the FP32 force arrays
    (and adds bonds and the volume push there), so every force differs by a
    few FP32 roundings of its own size.  Over k steps that moves v by at most
    k dt (4 u |F|max / m_min) and o by k dt 4 u |t|max (u = 2^-24), x by
    k dt times the velocity bound and n by k dt times the o bound; the bounds
    below are 4x those (neighbouring forces respond to the moved positions
    within these few steps by far less).  Over tens of steps the trajectories
    separate chaotically from that rounding, so mixed runs are compared over
    a few steps only."""
    sim.compute_pairwise(zero_first=True)
    sim.compute_bonded()
    f, t = sim.get_forces(orbc.LIPID)
    u = 2.0 ** -24
    dv = 4 * nsteps * dt * 4 * u * np.linalg.norm(f, axis=1).max() / 1.0
    do = 4 * nsteps * dt * 4 * u * np.linalg.norm(t, axis=1).max()
    return dict(v=dv, x=nsteps * dt * dv, o=do, n=nsteps * dt * do)


@pytest.mark.parametrize("precision,tol", [(orbc.FP64, 1e-8), (orbc.MIXED, None)])
@pytest.mark.parametrize("skin", [0.4, 0.03])
def test_run_with_volume_constraint_matches_call_sequence(skin, precision, tol):
    """orbc_run with the volume constraint == the per-step calls of the
    reference loop (src/openrbc.cpp:200-233 with constrain_volume enabled):
    Voronoi update every freq_voronoi steps, forces, constrain_volume, then
    integrate(verlet_langevin).  skin=0.03 forces rebuilds inside the batches,
    so the skip gate of the cell/volume kernels is exercised too.  Mixed
    precision: the fused step sums the pair terms in FP64 and uses them at
    once, the per-call sequence stores them in the FP32 force arrays first,
    so the two differ by that rounding (bounds in _rounding_bounds)."""
    sphere = orbc.random_sphere(r=12.0, seed=2)

    def make(target):
        sim = orbc.Simulation(precision=precision, seed=11, skin=skin, kBT=0.22, volume_target=target,
                              volume_strength=0.5, freq_cleanup=0)
        orbc.system_to_simulation(sphere, sim, kT=0.22, tilt=0.02, seed=3)
        sim.voronoi_init(0, 4)
        return sim

    call_sim = make(0.0)
    # target 2% above the initial enclosed volume, so that the push is not zero
    V0 = 1.02 * call_sim.constrain_volume(1.0, 0.0)
    run_sim = make(V0)
    nsteps = 40 if precision == orbc.FP64 else 8
    bound = None if precision == orbc.FP64 else _rounding_bounds(make(V0), nsteps)
    run_sim.run(nsteps)
    for step in range(nsteps):
        if step % 2 == 0:
            call_sim.voronoi_update()
        call_sim.compute_pairwise(zero_first=True)
        call_sim.compute_bonded()
        call_sim.constrain_volume(V0, 0.5)
        call_sim.integrate("verlet_langevin")
    a = run_sim.get_state(orbc.LIPID)
    b = call_sim.get_state(orbc.LIPID)
    for k in ("x", "v", "n", "o"):
        err = np.abs(a[k] - b[k]).max()
        assert err < (tol * max(1.0, np.abs(b[k]).max()) if bound is None else bound[k]), (k, err, bound)
    ca, _, _ = run_sim.voronoi_get()
    cb, _, _ = call_sim.voronoi_get()
    assert np.abs(ca - cb).max() < 1e-4
    if skin < 0.1:
        assert run_sim.observables()["rebuilds"] > (3 if precision == orbc.FP64 else 1)


# ---- stray-lipid cleanup (delete_lipid, src/cleanup.h:29-91) ----------------------------
def _lifted_sphere(r=12.0, seed=2, n_lift=12, lift=4.0):
    sphere = orbc.random_sphere(r=r, seed=seed)
    rng = np.random.default_rng(seed + 100)
    idx = rng.choice(len(sphere["lipid_x"]), n_lift, replace=False)
    sphere["lipid_x"][idx] *= (r + lift) / r              # radially out of the membrane
    return sphere, idx


@pytest.mark.parametrize("precision", [orbc.FP64, orbc.MIXED])
def test_stray_cleanup_vs_oracle(precision, oracle):
    """orbc_delete_stray_lipids removes exactly the lipids the oracle's
    delete_lipid removes on the device's own cells; the survivors keep their
    upload order and state; the compacted store evaluates forces as the
    oracle does on the surviving system."""
    sphere, lifted = _lifted_sphere()
    sim = orbc.Simulation(precision=precision, seed=5, freq_cleanup=0)
    orbc.system_to_simulation(sphere, sim)
    # relax the random sphere's near-overlapping lipids (the lifted ones have
    # no neighbours and stay where they are)
    sim.minimize(200)
    sim.voronoi_init(0, 4)
    c, aff, _ = sim.voronoi_get()
    before = sim.get_state(orbc.LIPID)
    x0 = before["x"]
    exp = oracle.stray_delete(x0, np.zeros(len(x0), np.int32), aff, c, 2.5)
    # most lifted lipids are strays (a lifted lipid that pulled its own cell's
    # centroid, or shares a cell with another, may stay)
    assert exp[lifted].sum() >= len(lifted) // 2
    nd = sim.delete_stray_lipids(2.5)
    assert nd == exp.sum() and sim.n[orbc.LIPID] == len(x0) - nd
    obs = sim.observables()
    assert obs["n_lipid"] == len(x0) - nd and obs["lipids_deleted"] == nd
    after = sim.get_state(orbc.LIPID)
    for k in ("x", "v", "n", "o"):
        assert np.array_equal(after[k], before[k][~exp]), k
    # forces of the compacted store (lipids only) against the oracle on the survivors
    sim.compute_pairwise(zero_first=True)
    f, t = sim.get_forces(orbc.LIPID)
    xs, ns = x0[~exp], before["n"][~exp]
    of, ot, _, _ = oracle.forces(xs, ns, np.zeros(len(xs), np.int32))
    if precision == orbc.FP64:
        assert eps_f(f, of) < EPS_FP64 and eps_f(t, ot) < EPS_FP64
    else:
        assert eps_f(f, of) < EPS_MIXED and eps_f(t, ot) < EPS_MIXED, (eps_f(f, of), eps_f(t, ot))
    # nothing left to remove with the cells recomputed
    sim.voronoi_update()
    assert sim.delete_stray_lipids(2.5) >= 0


@pytest.mark.parametrize("precision,tol", [(orbc.FP64, 1e-8), (orbc.MIXED, None)])
def test_run_with_cleanup_matches_call_sequence(precision, tol):
    """orbc_run removes strays every freq_cleanup steps on Voronoi-update
    steps, before the update (src/openrbc.cpp:200-202): equal to the per-call
    sequence delete_stray_lipids, voronoi_update, forces, constrain_volume,
    integrate.  The FP64 run crosses two cleanup steps (0 and 20); the mixed
    run crosses the first and is held to the bounds of _rounding_bounds."""
    sphere, lifted = _lifted_sphere(seed=4)

    def make(fc, target):
        sim = orbc.Simulation(precision=precision, seed=11, skin=0.3, kBT=0.22, volume_target=target,
                              volume_strength=0.5, freq_cleanup=fc)
        orbc.system_to_simulation(sphere, sim, kT=0.22, tilt=0.02, seed=3)
        sim.voronoi_init(0, 4)
        return sim

    call_sim = make(0, 0.0)
    V0 = 1.02 * call_sim.constrain_volume(1.0, 0.0)
    run_sim = make(20, V0)
    nsteps = 30 if precision == orbc.FP64 else 6
    bound = None
    if precision != orbc.FP64:
        bsim = make(0, V0)
        bsim.delete_stray_lipids(2.5)
        bound = _rounding_bounds(bsim, nsteps)
    run_sim.run(nsteps)
    removed = 0
    for step in range(nsteps):
        if step % 2 == 0:
            if step % 20 == 0:
                removed += call_sim.delete_stray_lipids(2.5)
            call_sim.voronoi_update()
        call_sim.compute_pairwise(zero_first=True)
        call_sim.compute_bonded()
        call_sim.constrain_volume(V0, 0.5)
        call_sim.integrate("verlet_langevin")
    assert removed >= len(lifted) // 2
    assert run_sim.observables()["lipids_deleted"] == removed
    a = run_sim.get_state(orbc.LIPID)
    b = call_sim.get_state(orbc.LIPID)
    assert a["x"].shape == b["x"].shape
    for k in ("x", "v", "n", "o"):
        err = np.abs(a[k] - b[k]).max()
        assert err < (tol * max(1.0, np.abs(b[k]).max()) if bound is None else bound[k]), (k, err, bound)


# ---- trajectory / topology files: byte for byte the reference's writers ----------------
def test_save_frame_and_topology_match_reference(tmp_path):
    """orbc_save_frame / orbc_save_topology on the state the reference's
    save_frame / save_topology wrote in tests/golden/frame.npz: the same
    bytes (lipids then proteins; FP32 fields; cells from the same centroids)."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "frame.npz"))
    nl = len(g["lx"])
    f64 = lambda a: np.asarray(a, np.float64)
    sim = orbc.Simulation(precision=orbc.FP64)
    lids = np.arange(int(g["lbase"]), int(g["lbase"]) + nl, dtype=np.uint32)
    sim.set_state(orbc.LIPID, f64(g["lx"]), f64(g["ln"]), lids, v=f64(g["lv"]))
    sim.set_state(orbc.PROTEIN, f64(g["px"]), f64(g["pn"]), g["ptag"].astype(np.uint32), v=f64(g["pv"]),
                  type=g["ptype"])
    sim.set_bonds(g["bt"], g["bi"].astype(np.uint32), g["bj"].astype(np.uint32))
    sim.voronoi_set_centroids(g["centroids"])
    for sp, m in ((orbc.LIPID, nl), (orbc.PROTEIN, len(g["px"]))):
        sim.set_forces(sp, np.zeros((m, 3)), np.zeros((m, 3)))
    fp = tmp_path / "f.orbc"
    sim.save_frame(str(fp), 31, append=False)
    ref = g["frame"].tobytes()
    end = ref.index(b"FRAMEEND") + 8                 # the first of the two reference frames (step 0)
    assert fp.read_bytes() == ref[:end]
    tp = tmp_path / "t.data"
    sim.save_topology(str(tp))
    assert tp.read_bytes() == g["topo"].tobytes()
    back = orbc.read_frame(str(fp), 0)
    assert np.array_equal(back["x"], np.concatenate([g["lx"], g["px"]]))


# ---- restart files (resume; no reference counterpart) -----------------------------------
@pytest.mark.parametrize("thermostat", [orbc.RUN_LANGEVIN, orbc.RUN_NOSE_HOOVER])
def test_restart_resumes_the_run(system, thermostat, tmp_path):
    """30 steps == 15 steps, save_restart, load_restart into a fresh context, 15
    steps.  Same noise (keyed by id and step) and the same Nose-Hoover state;
    the neighbour order after the reload may differ, so FP64 rounding only."""
    kw = dict(seed=9, kBT=0.22, thermostat=thermostat, skin=0.3, zeta=0.05, Q=7.0)
    a, _, _ = make_sim(system, orbc.FP64, **kw)
    a.run(30)
    b, _, _ = make_sim(system, orbc.FP64, **kw)
    b.run(15)
    path = str(tmp_path / "r.rst")
    b.save_restart(path)
    c = orbc.Simulation()
    c.load_restart(path)
    assert c.cfg.thermostat == thermostat and c.observables()["step"] == 15
    c.run(15)
    for sp in (orbc.LIPID, orbc.PROTEIN):
        sa, sc = a.get_state(sp), c.get_state(sp)
        for k in ("x", "v", "n", "o"):
            assert np.abs(sa[k] - sc[k]).max() < 1e-9 * max(1.0, np.abs(sa[k]).max()), (sp, k)
    assert abs(a.observables()["zeta"] - c.observables()["zeta"]) < 1e-9


def test_restart_with_cells_and_volume(tmp_path):
    """The Voronoi centroids and every particle's cell come back, so a run with
    the in-loop volume constraint resumes along the same trajectory."""
    sphere = orbc.random_sphere(r=12.0, seed=2)

    def make():
        sim = orbc.Simulation(precision=orbc.FP64, seed=11, skin=0.3, kBT=0.22, volume_target=-1.0,
                              volume_strength=0.5, freq_cleanup=0)
        orbc.system_to_simulation(sphere, sim, kT=0.22, tilt=0.02, seed=3)
        sim.voronoi_init(0, 4)
        return sim

    a = make()
    a.run(21)
    b = make()
    b.run(11)                     # stops between two Voronoi updates
    path = str(tmp_path / "v.rst")
    b.save_restart(path)
    c = orbc.Simulation()
    c.load_restart(path)
    c.run(10)
    sa, sc = a.get_state(orbc.LIPID), c.get_state(orbc.LIPID)
    for k in ("x", "v", "n"):
        assert np.abs(sa[k] - sc[k]).max() < 1e-9 * max(1.0, np.abs(sa[k]).max()), k
    ca, la, _ = a.voronoi_get()
    cc, lc, _ = c.voronoi_get()
    assert np.abs(ca - cc).max() < 1e-5 and np.array_equal(la, lc)


def test_run_volume_needs_cells():
    sphere = orbc.random_sphere(r=8.0, seed=2)
    sim = orbc.Simulation(precision=orbc.FP64, volume_target=-1.0)
    orbc.system_to_simulation(sphere, sim)
    with pytest.raises(orbc.OrbcError):
        sim.run(2)


def test_partition_exact_with_drifted_centroid_index(oracle):
    """The centroid grid is rebuilt only at neighbour rebuilds; between them
    the nearest-centroid search widens its reach by the measured centroid
    drift.  After many Lloyd updates with a large skin (few rebuilds) the
    affiliation must still equal the brute-force nearest centroid."""
    sphere = orbc.random_sphere(r=14.0, seed=5)
    sim = orbc.Simulation(precision=orbc.FP64, seed=3, skin=1.5, kBT=0.22, volume_target=-1.0, volume_strength=0.5)
    orbc.system_to_simulation(sphere, sim, kT=0.22, tilt=0.02, seed=4)
    sim.voronoi_init(0, 2)                    # few Lloyd steps: the centroids keep moving
    sim.run(60)
    sim.voronoi_update()
    c, al, _ = sim.voronoi_get()
    x = sim.get_state(orbc.LIPID)["x"]
    assert np.array_equal(al, oracle.partition(x, c))


# ---- periodic boxes (planar patch workloads; no reference counterpart) ----------------
@pytest.fixture(scope="module", params=[0, 7])
def psystem(request):
    return periodic_patch(18, 24, seed=request.param)      # at least 5 cells per periodic axis


def make_psim(s, precision, **kw):
    """Periodic simulation, built (so that positions are wrapped into the box)."""
    lo, hi, per = s["box"]
    sim, lip, prot = make_sim(s, precision, box_lo=list(lo), box_hi=list(hi), periodic=[int(p) for p in per], **kw)
    sim.build_neighbors()
    return sim, lip, prot


@pytest.fixture
def pbox(oracle):
    yield oracle
    oracle.set_box(None)


@pytest.mark.parametrize("precision", [orbc.FP64, orbc.MIXED])
def test_periodic_pair_set_bit_exact(psystem, precision, pbox):
    s = psystem
    sim, lip, prot = make_psim(s, precision)
    x = gather(sim, lip, prot, "state")["x"]          # wrapped into the box by the build
    lo, hi, per = s["box"]
    assert np.all(x[:, :2] >= lo[:2]) and np.all(x[:, :2] < hi[:2])
    pbox.set_box(s["box"])
    if precision == orbc.FP64:
        exp = id_pairs(s, pbox.pairs_brute(x, s["type"]))
    else:
        exp = id_pairs(s, pbox.pairs_mixed_anchor(x, s["type"], sim.observables()))
    got = sim.pairs()
    assert len(got) == len(exp) > 0 and np.array_equal(got, exp)
    pbox.set_box(None)
    assert len(id_pairs(s, pbox.pairs_brute(x, s["type"]))) < len(got)    # pairs across the edges exist


@pytest.mark.parametrize("precision,tol", [(orbc.FP64, EPS_FP64), (orbc.MIXED, EPS_MIXED)])
def test_periodic_forces_vs_oracle(psystem, precision, tol, pbox):
    s = psystem
    sim, lip, prot = make_psim(s, precision)
    sim.compute_pairwise(zero_first=True, energy=True)
    sim.compute_bonded(energy=True)
    f, t, e, w = gather(sim, lip, prot, energy=True)
    x = gather(sim, lip, prot, "state")["x"]
    pbox.set_box(s["box"])
    of, ot, oe, ow = pbox.forces(x, s["n"], s["type"], bonds=s["bonds"])
    assert eps_f(f, of) < tol
    assert eps_f(t, ot) < tol
    assert np.abs(e - oe).max() / np.sqrt(np.mean(oe ** 2)) < 10 * tol


@pytest.mark.parametrize("precision,tol", [(orbc.FP64, 1e-10), (orbc.MIXED, 2e-5)])
def test_periodic_translation_invariance(psystem, precision, tol):
    s = psystem
    lo, hi, _ = s["box"]
    out = []
    for shift in (np.zeros(3), np.array([0.37 * (hi[0] - lo[0]), 0.61 * (hi[1] - lo[1]), 0.0])):
        t = dict(s)
        t["x"] = s["x"] + shift
        sim, lip, prot = make_psim(t, precision)
        sim.compute_pairwise(zero_first=True)
        sim.compute_bonded()
        out.append(gather(sim, lip, prot))
    (f0, t0), (f1, t1) = out
    assert eps_f(f1, f0) < tol and eps_f(t1, t0) < tol


def _min_image_diff(a, b, box):
    lo, hi, per = box
    d = a - b
    for k in range(3):
        if per[k]:
            L = hi[k] - lo[k]
            d[:, k] -= L * np.round(d[:, k] / L)
    return d


def test_periodic_langevin_run_tracks_oracle(pbox):
    """Small skin: rebuilds (and wraps) inside the batched run."""
    s = periodic_patch(16, 22, seed=3)
    seed = 9
    sim, lip, prot = make_psim(s, orbc.FP64, seed=seed, skin=0.03)
    pbox.set_box(s["box"])
    st = {k: s[k].copy() for k in ("x", "v", "n", "o")}
    st.update(type=s["type"].astype(np.int32), id=s["id"].astype(np.uint32))
    nsteps = 40
    sim.run(nsteps)
    for step in range(nsteps):
        f, t, _, _ = pbox.forces(st["x"], st["n"], st["type"], bonds=s["bonds"])
        st["f"], st["t"] = f, t
        pbox.langevin(st, 0.22, 0.01, 0.01, seed, step)
    got = gather(sim, lip, prot, "state")
    assert np.abs(_min_image_diff(got["x"], st["x"], s["box"])).max() < 1e-8
    assert np.abs(got["n"] - st["n"]).max() < 1e-8
    assert sim.observables()["rebuilds"] > 3


def test_periodic_nose_hoover_run_tracks_oracle(pbox):
    """orbc_run with the Nose-Hoover thermostat (the planar-patch workload's
    integrator) against the oracle stepping the fused NH halves, src/openrbc.cpp:176-243."""
    s = periodic_patch(16, 22, seed=5)
    n = len(s["type"])
    Q, kT, dt, zeta = 0.01 * n, 0.22, 0.01, 0.0
    sim, lip, prot = make_psim(s, orbc.FP64, thermostat=orbc.RUN_NOSE_HOOVER, Q=Q, kBT=kT, skin=0.05)
    pbox.set_box(s["box"])
    st = {k: s[k].copy() for k in ("x", "v", "n", "o")}
    st.update(type=s["type"].astype(np.int32))
    big = (-1e9, -1e9, -1e9), (1e9, 1e9, 1e9)           # periodic axes have no walls; z walls far away
    nsteps = 30
    sim.run(nsteps)

    def forces():
        f, t, _, _ = pbox.forces(st["x"], st["n"], st["type"], bonds=s["bonds"])
        st["f"], st["t"] = f, t

    forces()
    ke = pbox.nh_second(st, dt, zeta)
    zeta = pbox.nh_zeta(zeta, dt, Q, ke, n, kT)
    for _ in range(nsteps):
        ke = pbox.nh_first(st, dt, zeta, *big)
        zeta = pbox.nh_zeta(zeta, dt, Q, ke, n, kT)
        forces()
        ke = pbox.nh_second(st, dt, zeta)
        zeta = pbox.nh_zeta(zeta, dt, Q, ke, n, kT)
    got = gather(sim, lip, prot, "state")
    assert np.abs(_min_image_diff(got["x"], st["x"], s["box"])).max() < 1e-8
    assert np.abs(got["v"] - st["v"]).max() < 1e-8
    assert abs(sim.observables()["zeta"] - zeta) < 1e-9 * max(1.0, abs(zeta))


def test_periodic_box_rejects_voronoi():
    s = periodic_patch(16, 22, seed=1)
    sim, _, _ = make_psim(s, orbc.FP64)
    with pytest.raises(orbc.OrbcError):
        sim.voronoi_init(0, 2)


# ---- batched ensembles (one store, one launch per step, per-member kBT) -----------------
def _ensemble_sim(e, members, precision, seed=13, **kw):
    sel = np.isin(e["group"], members)
    sim = orbc.Simulation(precision=precision, seed=seed, **kw)
    x, nrm, ids, g = e["lipid_x"][sel], e["lipid_n"][sel], e["lipid_id"][sel], e["group"][sel]
    # Maxwell velocities drawn for the whole ensemble, so a member alone gets the same ones
    rng = np.random.default_rng(4)
    v = (rng.normal(0, 1, e["lipid_x"].shape) * np.sqrt(e["kT"][e["group"]])[:, None])[sel]
    sim.set_state(orbc.LIPID, x, nrm, ids, v=v)
    sim.set_state(orbc.PROTEIN, np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0, np.uint32), type=np.zeros(0, np.int32))
    return sim, sel, g, v


@pytest.mark.parametrize("precision,tol", [(orbc.FP64, 1e-9), (orbc.MIXED, 1e-5)])
def test_ensemble_equals_separate_runs(precision, tol):
    """Three vesicles with different kBT in one store evolve exactly as each
    vesicle alone at its own kBT (noise keyed by particle id and step)."""
    e = orbc.vesicle_ensemble(n_members=3, n_min=900, n_max=2500, gap=10.0, seed=2)
    nsteps = 25
    ens, sel, g, v = _ensemble_sim(e, [0, 1, 2], precision)
    remap = {m: k for k, m in enumerate([0, 1, 2])}
    ens.set_groups(orbc.LIPID, [remap[m] for m in g], e["kT"])
    ens.run(nsteps)
    xe = ens.get_state(orbc.LIPID)["x"]
    ids_e = e["lipid_id"][sel]
    # no pair between members
    p = ens.pairs()
    gid = dict(zip(ids_e.tolist(), g.tolist()))
    assert all(gid[a] == gid[b] for a, b in p.tolist())
    for m in range(3):
        one, sel1, _, _ = _ensemble_sim(e, [m], precision, kBT=float(e["kT"][m]))
        one.run(nsteps)
        x1 = one.get_state(orbc.LIPID)["x"]
        xe_m = xe[g == m]
        assert np.abs(xe_m - x1).max() < tol * max(1.0, np.abs(x1).max()), m


def _mesh_member_sim(e, members, precision, seed=13, **kw):
    """Members `members` of a vesicle_ensemble_mesh system (both species,
    bonds, per-member kBT) in one Simulation; velocities drawn for the whole
    ensemble so a member alone gets the same ones."""
    ls, ps = np.isin(e["lipid_group"], members), np.isin(e["protein_group"], members)
    rng = np.random.default_rng(4)
    lv = rng.normal(0, 1, e["lipid_x"].shape) * np.sqrt(e["kT"][e["lipid_group"]])[:, None]
    pv = orbc.systems.maxwell(e["protein_type"], 1.0, rng) * np.sqrt(e["kT"][e["protein_group"]])[:, None]
    sim = orbc.Simulation(precision=precision, seed=seed, **kw)
    sim.set_state(orbc.LIPID, e["lipid_x"][ls], e["lipid_n"][ls], e["lipid_id"][ls], v=lv[ls])
    sim.set_state(orbc.PROTEIN, e["protein_x"][ps], e["protein_n"][ps], e["protein_id"][ps], v=pv[ps],
                  type=e["protein_type"][ps])
    keep = np.isin(e["bond_i"], e["protein_id"][ps])
    sim.set_bonds(e["bond_type"][keep], e["bond_i"][keep], e["bond_j"][keep])
    return sim, ls, ps


@pytest.mark.parametrize("precision,tol", [(orbc.FP64, 1e-9), (orbc.MIXED, 1e-5)])
def test_mesh_ensemble_equals_separate_runs(precision, tol):
    """Cells from geodesic meshes through the initialiser (lipids, proteins,
    spectrin bonds) with different kBT in one store evolve as each cell alone
    at its own kBT."""
    e = orbc.vesicle_ensemble_mesh(n_members=3, n_min=3000, n_max=16000, gap=10.0, seed=2)
    assert len(e["protein_x"]) > 0 and len(e["bond_type"]) > 0
    nsteps = 25
    ens, ls, ps = _mesh_member_sim(e, [0, 1, 2], precision)
    ens.set_groups(orbc.LIPID, e["lipid_group"], e["kT"])
    ens.set_groups(orbc.PROTEIN, e["protein_group"], e["kT"])
    ens.run(nsteps)
    le, pe = ens.get_state(orbc.LIPID)["x"], ens.get_state(orbc.PROTEIN)["x"]
    for m in range(3):
        one, ls1, ps1 = _mesh_member_sim(e, [m], precision, kBT=float(e["kT"][m]))
        one.run(nsteps)
        for xe, x1 in ((le[e["lipid_group"] == m], one.get_state(orbc.LIPID)["x"]),
                       (pe[e["protein_group"] == m], one.get_state(orbc.PROTEIN)["x"])):
            assert np.abs(xe - x1).max() < tol * max(1.0, np.abs(x1).max()), m


def test_mesh_ensemble_full_size_properties():
    """The batched-ensemble workload of the benchmark at full shape (64 cells
    from geodesic meshes, 5k-300k particles each, ~4.5M particles, 0.8M
    bonds, mixed precision): after minimisation and 500 Langevin steps every
    state is finite, no pair joins two members, each member's temperature
    follows its own kBT, lists last many steps (no ejected lipids force
    rebuilds) and the stray cleanup removes few lipids."""
    e = orbc.vesicle_ensemble_mesh(n_members=64, seed=0)
    sim, ls, ps = _mesh_member_sim(e, list(range(64)), orbc.MIXED, skin=0.6)
    sim.set_groups(orbc.LIPID, e["lipid_group"], e["kT"])
    sim.set_groups(orbc.PROTEIN, e["protein_group"], e["kT"])
    sim.minimize(100)
    sim.voronoi_init(0, 4)
    r0 = sim.observables()["rebuilds"]
    sim.run(500)
    o = sim.observables()
    assert o["rebuilds"] - r0 < 60, o["rebuilds"] - r0
    assert o["lipids_deleted"] < 0.001 * len(e["lipid_x"]), o["lipids_deleted"]
    g = e["lipid_group"]
    st = sim.get_state(orbc.LIPID)
    assert np.isfinite(st["x"]).all() and np.isfinite(st["v"]).all()
    assert np.isfinite(sim.get_state(orbc.PROTEIN)["x"]).all()
    # member of each surviving lipid by the nearest cell centre (the cells are
    # 10 apart; the cleanup closes up the slots of the removed lipids)
    assert len(st["x"]) == len(g) - o["lipids_deleted"]
    _, gs = cKDTree(e["centre"]).query(st["x"])
    T = np.bincount(gs, weights=np.sum(st["v"] ** 2, 1), minlength=64) / (3 * np.bincount(gs, minlength=64))
    ratio = T / e["kT"]
    med = np.median(ratio)
    assert 0.5 < med < 1.1, ratio
    assert np.all(np.abs(ratio / med - 1) < 0.2), ratio
    # member kBT spans 0.2-0.24: with one shared thermostat T would not track it
    assert np.corrcoef(T, e["kT"])[0, 1] > 0.7, (T, e["kT"])
    pg = e["protein_group"]
    p = sim.pairs()
    gid = np.zeros(int(max(e["lipid_id"].max(), e["protein_id"].max())) + 1, np.int32)
    gid[e["lipid_id"]] = g
    gid[e["protein_id"]] = pg
    assert np.array_equal(gid[p[:, 0]], gid[p[:, 1]])


def test_ensemble_rejects_nose_hoover():
    e = orbc.vesicle_ensemble(n_members=2, n_min=500, n_max=600, seed=1)
    sim, sel, g, _ = _ensemble_sim(e, [0, 1], orbc.FP64, thermostat=orbc.RUN_NOSE_HOOVER)
    sim.set_groups(orbc.LIPID, g, e["kT"])
    with pytest.raises(orbc.OrbcError):
        sim.run(2)


def test_ensemble_full_size_properties():
    """The batched-ensemble workload at its full shape (64 vesicles, 5k-300k
    lipids each, ~4.5M particles, mixed precision): after minimisation and
    500 Langevin steps no pair joins two members, every state is finite, and
    each member's temperature follows its own kBT.  500 steps are short of the
    Langevin relaxation time at eta = 0.01, so every member is still sharing
    kinetic energy with its potential from the same minimised start: the
    members share one T / kBT below 1 rather than reaching 1."""
    e = orbc.vesicle_ensemble(n_members=64, seed=0)
    g = e["group"]
    sim, sel, _, _ = _ensemble_sim(e, list(range(64)), orbc.MIXED, skin=0.6)
    sim.set_groups(orbc.LIPID, g, e["kT"])
    sim.minimize(100)
    rng = np.random.default_rng(5)
    st = sim.get_state(orbc.LIPID)
    v = rng.normal(0, 1, st["x"].shape) * np.sqrt(e["kT"][g])[:, None]
    sim.set_state(orbc.LIPID, st["x"], st["n"], e["lipid_id"], v=v)
    sim.run(500)
    st = sim.get_state(orbc.LIPID)
    assert np.isfinite(st["x"]).all() and np.isfinite(st["v"]).all()
    T = np.bincount(g, weights=np.sum(st["v"] ** 2, 1), minlength=64) / (3 * np.bincount(g, minlength=64))
    ratio = T / e["kT"]
    med = np.median(ratio)
    assert 0.6 < med < 1.1, ratio
    assert np.all(np.abs(ratio / med - 1) < 0.2), ratio
    # member kBT spans 0.2-0.24: with one shared thermostat T would not track it
    assert np.corrcoef(T, e["kT"])[0, 1] > 0.7, (T, e["kT"])
    p = sim.pairs()
    gid = np.zeros(len(g) + 2, np.int32)
    gid[e["lipid_id"]] = g
    assert np.array_equal(gid[p[:, 0]], gid[p[:, 1]])


# ---- the planar membrane with its spectrin network (periodic, Nose-Hoover workload) ------
def _flat(sysd):
    """systems-format dict -> the combined arrays of these tests (lipids first)."""
    x = np.vstack([sysd["lipid_x"], sysd["protein_x"]])
    n = np.vstack([sysd["lipid_n"], sysd["protein_n"]])
    t = np.concatenate([np.zeros(len(sysd["lipid_x"]), np.int32), sysd["protein_type"]]).astype(np.int32)
    ids = np.concatenate([sysd["lipid_id"], sysd["protein_id"]]).astype(np.uint32)
    order = np.argsort(ids)
    idx = lambda q: order[np.searchsorted(ids[order], q)]
    bonds = (sysd["bond_type"], idx(sysd["bond_i"]), idx(sysd["bond_j"]))
    return dict(x=x, n=n, type=t, id=ids, v=np.zeros_like(x), o=np.zeros_like(x), bonds=bonds, box=sysd["box"])


@pytest.mark.parametrize("precision,tol", [(orbc.FP64, EPS_FP64), (orbc.MIXED, EPS_MIXED)])
def test_planar_membrane_forces_vs_oracle(precision, tol, pbox):
    s = _flat(orbc.planar_membrane(64, 64, seed=2))
    sim, lip, prot = make_psim(s, precision)
    sim.compute_pairwise(zero_first=True)
    sim.compute_bonded()
    f, t = gather(sim, lip, prot)
    x = gather(sim, lip, prot, "state")["x"]
    pbox.set_box(s["box"])
    pairs = pbox.pairs_cell(x, s["type"])
    if precision == orbc.FP64:
        assert np.array_equal(sim.pairs(), id_pairs(s, pairs))
    else:
        assert np.array_equal(sim.pairs(), id_pairs(s, pbox.pairs_mixed_anchor(x, s["type"], sim.observables())))
    of, ot, _, _ = pbox.forces(x, s["n"], s["type"], pairs=pairs, bonds=s["bonds"])
    assert eps_f(f, of) < tol and eps_f(t, ot) < tol


def test_planar_membrane_full_size_nose_hoover():
    """The planar-patch workload at its full shape (4096 x 4096 lattice,
    20.5M particles, spectrin network beneath, periodic in x and y, mixed
    precision, Nose-Hoover): 200 steps after a short minimisation stay finite,
    inside the box walls, at a bounded temperature."""
    sysd = orbc.planar_membrane(seed=1)
    lo, hi, per = sysd["box"]
    sim = orbc.Simulation(precision=orbc.MIXED, thermostat=orbc.RUN_NOSE_HOOVER, skin=0.6, box_lo=list(lo),
                          box_hi=list(hi), periodic=[int(p) for p in per])
    orbc.system_to_simulation(sysd, sim)
    sim.minimize(50)
    sim.run(200)
    obs = sim.observables()
    assert obs["step"] == 200 and np.isfinite(obs["zeta"])
    assert 0.0 < obs["temperature"] < 2.0
    st = sim.get_state(orbc.PROTEIN)
    assert np.isfinite(st["x"]).all() and np.abs(st["x"][:, 2]).max() < 100


@pytest.mark.parametrize("precision,tol", [(orbc.FP64, EPS_FP64), (orbc.MIXED, EPS_MIXED)])
def test_c1_minimised_forces_vs_oracle(c1, precision, tol, oracle):
    """C1 after 100 FP64 minimisation steps (the reference minimises before
    every run, src/openrbc.cpp:88-146): forces at the eps_F target."""
    s = c1
    sim, lip, prot = make_sim(s, orbc.FP64)
    sim.minimize(100)
    st = gather(sim, lip, prot, "state")
    sim.close()
    r = dict(s)
    r["x"], r["n"] = st["x"], st["n"]
    r["v"], r["o"] = np.zeros_like(st["x"]), np.zeros_like(st["x"])
    sim, lip, prot = make_sim(r, precision)
    sim.compute_pairwise(zero_first=True)
    sim.compute_bonded()
    f, t = gather(sim, lip, prot)
    pairs = oracle.pairs_cell(r["x"], r["type"])
    got = sim.pairs()
    exp = id_pairs(r, pairs)
    if precision == orbc.FP64:
        assert np.array_equal(got, exp)
    else:
        e = set(map(tuple, exp.tolist()))
        g = set(map(tuple, got.tolist()))
        assert len(e ^ g) <= 5
        order = np.argsort(r["id"])
        to_idx = lambda q: order[np.searchsorted(r["id"][order], q)]
        pairs = np.stack([to_idx(got[:, 0]), to_idx(got[:, 1])], 1)
    of, ot, _, _ = oracle.forces(r["x"], r["n"], r["type"], pairs=pairs, bonds=r["bonds"])
    assert eps_f(f, of) < tol
    assert eps_f(t, ot) < tol

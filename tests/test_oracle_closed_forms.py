"""Pins of the oracle that do not need the reference: closed forms, finite
differences, known answers, invariants, brute force.  Each check is chosen so
that a plausible slip (a dropped term, a sign, a transposed operand) fails it.
"""
import numpy as np
import pytest

from _systems import patch

def unit(v):
    return v / np.linalg.norm(v)


def pair_energy(oracle, ti, tj, xi, xj, ni, nj):
    hit, f, t, u, w = oracle.pair(ti, tj, xi - xj, ni, nj)
    return u if hit else 0.0


@pytest.mark.parametrize("ti,tj", [(0, 0), (0, 1), (0, 2), (0, 3), (2, 0), (1, 2), (2, 3), (3, 3), (0, 4), (0, 5),
                                   (1, 4), (2, 4), (2, 5), (5, 1)])
def test_force_is_minus_gradient(oracle, ti, tj):
    """F_i = -dU/dx_i and t_i = -dU/dn_i (central differences)."""
    ff = oracle.ff()
    rc = max(ff["cut_poly"][ti, tj], np.sqrt(ff["cut2_lj"][ti, tj]))
    rng = np.random.default_rng(10 * ti + tj)
    h = 1e-6
    checked = 0
    for _ in range(40):
        ni, nj = unit(rng.normal(size=3)), unit(rng.normal(size=3))
        d = unit(rng.normal(size=3)) * rng.uniform(0.75 * rc, 0.99 * rc)
        xi, xj = rng.normal(size=3), None
        xj = xi - d
        hit, f, t, u, w = oracle.pair(ti, tj, d, ni, nj)
        if not hit:
            continue
        checked += 1
        g = np.zeros(3); gn = np.zeros(3)
        for k in range(3):
            e = np.zeros(3); e[k] = h
            g[k] = (pair_energy(oracle, ti, tj, xi + e, xj, ni, nj) - pair_energy(oracle, ti, tj, xi - e, xj, ni, nj)) / (2 * h)
            gn[k] = (pair_energy(oracle, ti, tj, xi, xj, ni + e, nj) - pair_energy(oracle, ti, tj, xi, xj, ni - e, nj)) / (2 * h)
        scale = max(1.0, np.abs(f).max())
        assert np.abs(f + g).max() < 1e-5 * scale
        assert np.abs(t + gn).max() < 1e-5 * max(1.0, np.abs(t).max())
        assert abs(w - np.dot(d, f)) < 1e-9 * max(1.0, abs(w))
    assert checked > 0


def test_aligned_lipids_at_req_known_answer(oracle):
    """n_i = n_j perpendicular to e_ij at r = req: a = A = 1, F = 0, U = -eps,
    and the torque n x t vanishes (t parallel to n)."""
    n = np.array([0.0, 0.0, 1.0])
    for req, eps, tj in ((1.1225, 1.2, 0), (1.685, 1.4, 1), (1.685, 2.8, 2), (1.1225, 2.8, 3)):
        d = np.array([req, 0.0, 0.0])
        hit, f, t, u, w = oracle.pair(0, tj, d, n, n)
        assert hit
        assert np.abs(f).max() < 1e-12
        assert abs(u + eps) < 1e-12
        assert np.abs(np.cross(n, t)).max() < 1e-12


def test_zero_at_cutoff(oracle):
    n1, n2 = unit(np.array([0.2, 0.1, 1.0])), unit(np.array([-0.1, 0.3, 1.0]))
    for tj in range(4):
        d = unit(np.array([1.0, 0.5, 0.2])) * (2.6 * (1 - 1e-9))
        hit, f, t, u, w = oracle.pair(0, tj, d, n1, n2)
        assert hit and abs(u) < 1e-12 and np.abs(f).max() < 1e-12 and np.abs(t).max() < 1e-12
        hit, *_ = oracle.pair(0, tj, d * 1.001, n1, n2)
        assert not hit
    # below 1e-5 in r^2 nothing interacts (src/compute_pairwise_fused.h:210)
    hit, *_ = oracle.pair(0, 0, np.array([3e-3, 0, 0]), n1, n2)
    assert not hit


def test_lj_textbook(oracle):
    """WCA: U = 4 eps ((s/r)^12 - (s/r)^6), zero force at 2^(1/6) sigma."""
    n = np.array([0.0, 0.0, 1.0])
    for ti, tj, sig in ((0, 4, 1.0), (0, 5, 1.0), (1, 4, 3.4), (2, 4, 3.4), (2, 5, 1.0)):
        r0 = 2 ** (1 / 6) * sig
        hit, f, t, u, w = oracle.pair(ti, tj, np.array([0, r0, 0.0]), n, n)
        assert hit and np.abs(f).max() < 1e-9 and abs(u + 1.0) < 1e-12
        r = 0.95 * sig
        hit, f, t, u, w = oracle.pair(ti, tj, np.array([0, 0, r]), n, n)
        assert abs(u - 4 * ((sig / r) ** 12 - (sig / r) ** 6)) < 1e-9 * abs(u)
        assert abs(f[2] - 24 * (2 * sig ** 12 / r ** 13 - sig ** 6 / r ** 7)) < 1e-9 * abs(f[2])
        assert np.abs(t).max() == 0


def test_newton_antisymmetry(oracle):
    rng = np.random.default_rng(3)
    for _ in range(50):
        ti, tj = rng.integers(0, 6, 2)
        d = unit(rng.normal(size=3)) * rng.uniform(0.8, 3.0)
        ni, nj = unit(rng.normal(size=3)), unit(rng.normal(size=3))
        h1, f1, t1, u1, _ = oracle.pair(ti, tj, d, ni, nj)
        h2, f2, t2, u2, _ = oracle.pair(tj, ti, -d, nj, ni)
        assert h1 == h2
        if h1:
            assert np.abs(f1 + f2).max() < 1e-12 * max(1, np.abs(f1).max())
            assert abs(u1 - u2) < 1e-12 * max(1, abs(u1))


def test_bond_hooke(oracle):
    x = np.array([[0.0, 0, 0], [2.5, 0, 0], [0, 0, 0], [0, 2.25, 0]])
    f, e, w = oracle.bonded(x, (np.array([0, 0]), np.array([0, 2]), np.array([1, 3])))
    assert abs(f[0, 0] - 57 * (2.5 - 2.25)) < 1e-12 and abs(f[1, 0] + 57 * 0.25) < 1e-12
    assert np.abs(f[2:]).max() < 1e-12                 # at r0: no force
    assert abs(e[:2].sum() - 0.5 * 57 * 0.25 ** 2) < 1e-12


def test_momentum_and_angular_momentum(oracle):
    """sum f = 0 and sum (x cross f + n cross t) = 0 over a patch with bonds."""
    s = patch(12, 12, seed=1)
    f, t, e, w = oracle.forces(s["x"], s["n"], s["type"], bonds=s["bonds"])
    scale = np.abs(f).max()
    assert np.abs(f.sum(0)).max() < 1e-10 * scale * len(f)
    L = np.cross(s["x"], f).sum(0) + np.cross(s["n"], t).sum(0)
    assert np.abs(L).max() < 1e-9 * scale * len(f)


def test_rotation_covariance(oracle):
    s = patch(10, 10, seed=2)
    f, t, e, _ = oracle.forces(s["x"], s["n"], s["type"], bonds=s["bonds"])
    q, _ = np.linalg.qr(np.random.default_rng(0).normal(size=(3, 3)))
    f2, t2, e2, _ = oracle.forces(s["x"] @ q.T, s["n"] @ q.T, s["type"], bonds=s["bonds"])
    assert np.abs(f2 - f @ q.T).max() < 1e-9 * np.abs(f).max()
    assert np.abs(t2 - t @ q.T).max() < 1e-9 * np.abs(t).max()
    assert np.abs(e2 - e).max() < 1e-9 * np.abs(e).max()


def test_energy_sums_to_pair_energies(oracle):
    s = patch(9, 9, seed=4, proteins=False)
    pairs = oracle.pairs_brute(s["x"], s["type"])
    _, _, e, _ = oracle.forces(s["x"], s["n"], s["type"], pairs=pairs)
    tot = 0.0
    for i, j in pairs:
        tot += oracle.pair(0, 0, s["x"][i] - s["x"][j], s["n"][i], s["n"][j])[3]
    assert abs(e.sum() - tot) < 1e-10 * abs(tot)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_pair_sets_brute_equals_cell(oracle, seed):
    s = patch(14, 14, seed=seed)
    a = oracle.pairs_brute(s["x"], s["type"])
    b = oracle.pairs_cell(s["x"], s["type"])
    assert np.array_equal(a, b)
    # mixed (FP32 anchor) geometry agrees except for pairs within rounding of a cutoff
    m = oracle.pairs_mixed(s["x"], s["type"], 1.0)
    ka = set(map(tuple, a.tolist())); km = set(map(tuple, m.tolist()))
    assert len(ka ^ km) <= 2


def test_philox_known_answers(oracle):
    """Random123 Philox-4x32-10 known-answer vectors."""
    assert [hex(v) for v in oracle.philox([0, 0, 0, 0], [0, 0])] == ["0x6627e8d5", "0xe169c58d", "0xbc57ac4c",
                                                                      "0x9b00dbd8"]
    assert [hex(v) for v in oracle.philox([0xffffffff] * 4, [0xffffffff] * 2)] == ["0x408f276d", "0x41c83b0e",
                                                                                     "0xa20bc7c6", "0x6d5451fd"]
    assert [hex(v) for v in oracle.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                                          [0xa4093822, 0x299f31d0])] == ["0xd16cfe09", "0x94fdcceb", "0x5001e420",
                                                                         "0x24126ea1"]


def test_noise_distribution(oracle):
    """Uniform [-1,1): mean 0, variance 1/3 (the reference's u11 noise, F9)."""
    r = np.array([oracle.noise(7, i, s) for i in range(300) for s in range(30)])
    assert r.min() >= -1 and r.max() < 1
    assert abs(r.mean()) < 0.02 and abs(r.var() - 1 / 3) < 0.02


def test_langevin_single_particle_closed_form(oracle):
    """kT = 0, f = t = o = 0: v' = v (1 - gamma dt/m), x' = x + v' dt, n unchanged."""
    for ty in range(6):
        ff = oracle.ff()
        m = ff["mass"][ty]
        gamma = 6 * np.pi * 0.01 * ff["radius"][ty]
        s = dict(x=np.array([[1.0, 2.0, 3.0]]), v=np.array([[0.3, -0.2, 0.1]]), n=np.array([[0, 0, 1.0]]),
                 o=np.zeros((1, 3)), f=np.zeros((1, 3)), t=np.zeros((1, 3)), type=np.array([ty], np.int32),
                 id=np.array([1], np.uint32))
        v0 = s["v"].copy(); x0 = s["x"].copy()
        oracle.langevin(s, 0.0, 0.01, 0.01, 1, 0)
        v1 = v0 * (1 - gamma * 0.01 / m)
        assert np.abs(s["v"] - v1).max() < 1e-15
        assert np.abs(s["x"] - (x0 + v1 * 0.01)).max() < 1e-15
        assert np.abs(s["n"] - [[0, 0, 1]]).max() == 0


def test_langevin_rotation_closed_form(oracle):
    """t perpendicular to n: o += dt (n x t); n rotated by o x n dt and renormalised."""
    s = dict(x=np.zeros((1, 3)), v=np.zeros((1, 3)), n=np.array([[0, 0, 1.0]]), o=np.zeros((1, 3)),
             f=np.zeros((1, 3)), t=np.array([[1.0, 0, 0]]), type=np.array([0], np.int32), id=np.array([1], np.uint32))
    oracle.langevin(s, 0.0, 0.01, 0.01, 1, 0)
    o = 0.01 * np.cross([0, 0, 1.0], [1.0, 0, 0])      # (0, 0.01, 0)
    n = np.array([0, 0, 1.0]) + np.cross(o, [0, 0, 1.0]) * 0.01
    assert np.abs(s["o"][0] - o).max() < 1e-16
    assert np.abs(s["n"][0] - n / np.linalg.norm(n)).max() < 1e-15


def test_temperature_closed_form(oracle):
    rng = np.random.default_rng(5)
    v = rng.normal(size=(100, 3)); t = rng.integers(0, 6, 100).astype(np.int32)
    m = np.array([1, 4, 4, 1, 10, 10.0])[t]
    assert abs(oracle.temperature(v, t) - (m[:, None] * v * v).sum() / 300) < 1e-12


def test_nh_zeta_closed_form(oracle):
    assert abs(oracle.nh_zeta(0.1, 0.01, 5.0, 30.0, 20, 0.22) - (0.1 + 0.5 * 0.01 / 5.0 * (30.0 - 1.5 * 20 * 0.22))) < 1e-15


def test_morton_brute_interleave(oracle):
    """Bits 0..10 of uint(2x + 2000) per axis, interleaved x,y,z from bit 0 and
    truncated to 32 bits (the reference's bit_space3 drops bit 11 and up)."""
    rng = np.random.default_rng(1)
    for p in rng.uniform(-500, 500, (200, 3)).astype(np.float32):
        u = [int(2.0 * np.float32(c) + 2000.0) for c in p]
        exp = 0
        for b in range(11):
            for d in range(3):
                exp |= ((u[d] >> b) & 1) << (3 * b + d)
        assert oracle.morton_ref(p[None])[0] == exp & 0xffffffff


def test_partition_and_centroids_definition(oracle):
    rng = np.random.default_rng(2)
    x = rng.uniform(-10, 10, (500, 3)); c = rng.uniform(-10, 10, (30, 3)).astype(np.float32)
    aff = oracle.partition(x, c)
    d = ((x.astype(np.float32)[:, None, :] - c[None]) ** 2).sum(-1)
    assert np.array_equal(aff, d.argmin(1))
    cc = oracle.centroids(x, aff, 30, init=c)
    for k in range(30):
        if (aff == k).any():
            assert np.abs(cc[k] - x[aff == k].mean(0)).max() < 1e-5


def test_volume_of_sphere(oracle):
    """Cone-sum volume of radially oriented lipids on a sphere of radius R:
    V = 1e-6 * (pi 1.26 / 4) * sum_cells h N_cell / 3 -> (4/3) pi R^3 * rho_a a_l 1e-6."""
    rng = np.random.default_rng(3)
    R, n = 40.0, 20000
    u = rng.normal(size=(n, 3)); u /= np.linalg.norm(u, axis=1, keepdims=True)
    x = R * u
    ncell = n // 14
    c = x[rng.choice(n, ncell, replace=False)].astype(np.float32)
    for _ in range(5):
        aff = oracle.partition(x, c)
        c = oracle.centroids(x, aff, ncell, init=c)
    aff = oracle.partition(x, c)
    f = np.zeros((n, 3))
    vol = oracle.constrain_volume(u, np.zeros(n, np.int32), aff, c, 1.0, 0.0, f)
    exp = 4 / 3 * np.pi * R ** 3 * (n / (4 * np.pi * R * R)) * (np.pi * 1.26 / 4) * 1e-6
    assert abs(vol / exp - 1) < 0.01
    # push: f = strength (V0 - V)/V0 m n_cell, along the outward normal
    f = np.zeros((n, 3))
    oracle.constrain_volume(u, np.zeros(n, np.int32), aff, c, 2 * vol, 0.5, f)
    assert np.all(np.einsum("ij,ij->i", f, u) > 0)
    assert abs(np.linalg.norm(f, axis=1).mean() - 0.25) < 0.01

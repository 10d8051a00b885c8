"""orbc_run's Langevin batches replayed from captured hipGraphs (cfg.graphs = 1)
against the same batches launched kernel by kernel (cfg.graphs = 0).

A graph holds whole timesteps -- the gated Voronoi update and volume
evaluation, the fused force + bond + Langevin kernel -- with the step number
read from a device counter, so one graph serves every batch of its length and
parity.  The runs below replay the cached graphs many times, with neighbour
rebuilds, stray-lipid cleanups and host synchronisations between the
replays, over several orbc_run calls; every state after every call must be
bit-identical to the kernel-by-kernel run (same kernels, same arguments).
"""
import numpy as np
import pytest

import openrbc_amd as orbc
from test_gpu_halo import _rbc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rbc():
    return _rbc(2)


def _sim(s, graphs, precision, **kw):
    sim = orbc.Simulation(precision=precision, graphs=graphs, **kw)
    t = s["type"]
    lip, prot = np.where(t == 0)[0], np.where(t != 0)[0]
    sim.set_state(orbc.LIPID, s["x"][lip], s["n"][lip], s["id"][lip], v=s["v"][lip], o=s["o"][lip])
    sim.set_state(orbc.PROTEIN, s["x"][prot], s["n"][prot], s["id"][prot], v=s["v"][prot], o=s["o"][prot],
                  type=t[prot])
    bt, bi, bj = s["bonds"]
    sim.set_bonds(bt, s["id"][bi], s["id"][bj])
    return sim


def _states(sim):
    return {sp: sim.get_state(sp) for sp in (orbc.LIPID, orbc.PROTEIN)}


@pytest.mark.parametrize("precision", [orbc.MIXED, orbc.FP64])
@pytest.mark.parametrize("skin,volume", [(0.05, True), (0.4, True), (0.05, False)])
def test_graph_batches_bit_identical(rbc, precision, skin, volume):
    print(f"graphs: icosphere-2 cell, n={len(rbc['type'])}, precision {precision}, skin {skin}, "
          f"volume {volume}", flush=True)
    kw = dict(seed=5, skin=skin, volume_target=-1.0 if volume else 0.0, volume_strength=0.2, freq_cleanup=20)
    sims = [_sim(rbc, g, precision, **kw) for g in (1, 0)]
    for sim in sims:
        sim.minimize(20)
        sim.voronoi_init(0, 4)
    for call, nsteps in enumerate((7, 33, 45, 1)):
        for sim in sims:
            sim.run(nsteps)
        a, b = (_states(sim) for sim in sims)
        for sp in a:
            for k in a[sp]:
                assert np.array_equal(a[sp][k], b[sp][k]), (call, sp, k)
    og, oe = (sim.observables() for sim in sims)
    assert og["step"] == oe["step"] == 86 and og["rebuilds"] == oe["rebuilds"]
    assert og["lipids_deleted"] == oe["lipids_deleted"]
    assert oe["graph_replays"] == 0 and og["graph_replays"] > og["graph_captures"] > 0
    for sim in sims:
        sim.close()


def test_graph_ensemble_members_bit_identical():
    """Members with their own kBT in one store (the batched ensemble): the
    member noise amplitudes are baked into the captured steps the same way."""
    e = orbc.vesicle_ensemble(n_members=6, n_min=800, n_max=3000, seed=1)
    res = []
    for graphs in (1, 0):
        sim = orbc.Simulation(precision=orbc.MIXED, seed=7, skin=0.3, graphs=graphs)
        rng = np.random.default_rng(7)
        v = rng.normal(0, 1, e["lipid_x"].shape) * np.sqrt(e["kT"][e["group"]])[:, None]
        sim.set_state(orbc.LIPID, e["lipid_x"], e["lipid_n"], e["lipid_id"], v=v)
        sim.set_state(orbc.PROTEIN, np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0, np.uint32),
                      type=np.zeros(0, np.int32))
        sim.set_groups(orbc.LIPID, e["group"], e["kT"])
        sim.minimize(10)
        sim.run(24)
        sim.run(17)
        res.append((sim.get_state(orbc.LIPID), sim.observables()))
        sim.close()
    for k in res[0][0]:
        assert np.array_equal(res[0][0][k], res[1][0][k]), k
    assert res[0][1]["graph_replays"] > 0


def test_graph_timing_falls_back_to_launches():
    """Kernel timing (HIP events around every fused kernel) launches the
    steps one by one; the run continues from the graph-replayed steps with
    the same state as a run without graphs."""
    s = _rbc(1)
    res = []
    for graphs in (1, 0):
        sim = _sim(s, graphs, orbc.MIXED, seed=3, skin=0.3)
        sim.run(25)
        replays = sim.observables()["graph_replays"]
        sim.set_timing(True)
        sim.run(40)
        tm = sim.timing()
        obs = sim.observables()
        assert obs["graph_replays"] == replays and (replays > 0) == (graphs == 1)
        assert tm["step_launches"] == 40 and tm["step_kernel_ms"] > 0
        res.append(sim.get_state(orbc.LIPID))
        sim.close()
    for k in res[0]:
        assert np.array_equal(res[0][k], res[1][k]), k


@pytest.mark.parametrize("precision", [orbc.MIXED, orbc.FP64])
@pytest.mark.parametrize("skin", [0.05, 0.4])
def test_nose_hoover_batches_equal_per_step_loop(rbc, precision, skin, monkeypatch):
    """The Nose-Hoover run in batches (skin check on the device between the
    two halves, a rebuild between the halves of the step that spends the list,
    one host synchronisation per batch) against the loop that synchronises
    every step (ORBC_NH_PER_STEP=1): same kernels in the same order, so the
    states are bit-identical, through rebuilds and over several orbc_run calls."""
    print(f"NH batches: precision {precision}, skin {skin}", flush=True)
    res = []
    for per_step in ("0", "1"):
        monkeypatch.setenv("ORBC_NH_PER_STEP", per_step)
        sim = _sim(rbc, 1, precision, seed=5, skin=skin, thermostat=orbc.RUN_NOSE_HOOVER, kBT=0.22)
        sim.minimize(10)
        for n in (9, 30, 1):
            sim.run(n)
        res.append((_states(sim), sim.observables()))
        sim.close()
    (a, oa), (b, ob) = res
    for sp in a:
        for k in a[sp]:
            assert np.array_equal(a[sp][k], b[sp][k]), (sp, k)
    assert oa["zeta"] == ob["zeta"] and oa["step"] == ob["step"] == 40 and oa["rebuilds"] == ob["rebuilds"]
    if skin < 0.1:
        assert oa["rebuilds"] > 3

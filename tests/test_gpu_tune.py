"""orbc_autotune: the run-time autotuner of orbc_run (skin, block halos,
centroid-grid edge, batch length measured on the uploaded system).

The tuner must leave the state as it found it (every trial restores it) and
apply a configuration from its candidate sets; a run afterwards follows a run
of a fresh context with the same knobs from the same state (to FP64 rounding:
the trials permute the store, so neighbour sums can be ordered differently).
"""
import json

import numpy as np
import pytest

import openrbc_amd as orbc
from test_gpu_halo import _rbc

pytestmark = pytest.mark.gpu


def _sim(s, precision, **kw):
    sim = orbc.Simulation(precision=precision, **kw)
    t = s["type"]
    lip, prot = np.where(t == 0)[0], np.where(t != 0)[0]
    sim.set_state(orbc.LIPID, s["x"][lip], s["n"][lip], s["id"][lip], v=s["v"][lip], o=s["o"][lip])
    sim.set_state(orbc.PROTEIN, s["x"][prot], s["n"][prot], s["id"][prot], v=s["v"][prot], o=s["o"][prot],
                  type=t[prot])
    bt, bi, bj = s["bonds"]
    sim.set_bonds(bt, s["id"][bi], s["id"][bj])
    return sim


@pytest.mark.parametrize("precision", [orbc.FP64, orbc.MIXED])
def test_autotune_restores_state_and_applies_a_candidate(tmp_path, precision):
    s = _rbc(2)
    kw = dict(seed=3, skin=0.4, volume_target=-1.0, volume_strength=0.2, freq_cleanup=20)
    sim = _sim(s, precision, **kw)
    sim.minimize(10)
    sim.voronoi_init(0, 4)
    sim.run(6)
    before = {sp: sim.get_state(sp) for sp in (orbc.LIPID, orbc.PROTEIN)}
    o0 = sim.observables()
    table = tmp_path / "tuning.jsonl"
    res = sim.autotune(trial_steps=12, table=str(table), warm_steps=3)
    after = {sp: sim.get_state(sp) for sp in (orbc.LIPID, orbc.PROTEIN)}
    for sp in before:
        for k in before[sp]:
            assert np.array_equal(before[sp][k], after[sp][k]), (sp, k)
    o1 = sim.observables()
    assert o1["step"] == o0["step"] and o1["n"] == o0["n"] and o1["lipids_deleted"] == o0["lipids_deleted"]
    tun = sim.tuning()
    assert tun["skin"] in (0.4, 0.5, 0.6, 0.7, 0.8) and tun["block_halo"] in (0, 1, 2)
    assert tun["centroid_cell"] in (6.0, 8.0, 10.0) and tun["batch"] in (8, 16, 32)
    assert res["ms_per_step"] <= res["ms_per_step_start"] and res["trials"] >= 10
    line = json.loads(table.read_text().strip().splitlines()[-1])
    assert len(line["trials"]) == res["trials"] and line["best"]["skin"] == tun["skin"]
    assert line["n_lipid"] == o0["n_lipid"] and line["precision"] == ("fp64" if precision == orbc.FP64 else "mixed")
    # continue: the same trajectory as a fresh context that did the same set-up
    # and then took the chosen knobs (to FP64 rounding: the trials left the
    # store in another order, so neighbour sums can be ordered differently)
    sim.run(15)
    ref = _sim(s, precision, **kw)
    ref.minimize(10)
    ref.voronoi_init(0, 4)
    ref.run(6)
    ref.set_tuning(**tun)
    ref.run(15)
    a, b = sim.get_state(orbc.LIPID), ref.get_state(orbc.LIPID)
    # mixed: a rounding-level difference of an FP64 position can move the
    # FP32 mirror of that particle by one FP32 ulp (6e-8 of an O(1) offset),
    # which the next steps carry into the trajectory at that level
    tol = 1e-8 if precision == orbc.FP64 else 1e-6
    for k in ("x", "n"):
        assert np.abs(a[k] - b[k]).max() < tol, (k, np.abs(a[k] - b[k]).max())
    sim.close()
    ref.close()

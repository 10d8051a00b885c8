"""Langevin statistics against the reference's own run.

tests/golden/langevin_stats.npz holds a 1900-lipid vesicle (minimised),
Maxwell velocities at kT = 0.22, and the time series of temperature, energy
per particle and area per lipid of 30000 steps of the reference's
verlet_langevin (src/integrate_langevin.h:99-149) with exact pairs, every 20
steps, computed by tests/_stats.py on the reference's snapshots.  The same
run through orbc_run (Philox noise instead of the reference's per-thread
MT19937 / xorshift streams, so only the statistics can agree) must give the
same means within their combined statistical error (block averages over the
24000 steps after the first 6000).
"""
import os

import numpy as np
import pytest

import openrbc_amd as orbc
from _stats import area_per_lipid, block_mean

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("precision", [orbc.MIXED, orbc.FP64])
def test_langevin_statistics_match_reference(precision):
    z = np.load(os.path.join(GOLD, "langevin_stats.npz"))
    x, n, v = (z[k].astype(np.float64) for k in ("x", "n", "v"))
    N = len(x)
    every, nsteps, relax = int(z["every"]), int(z["nsteps"]), int(z["relax"])
    sim = orbc.Simulation(precision=precision, kBT=float(z["kT"]), eta=float(z["eta"]), dt=float(z["dt"]), seed=2024,
                          skin=0.4)
    sim.set_state(orbc.LIPID, x, n, np.arange(1, N + 1, dtype=np.uint32), v=v)
    sim.set_state(orbc.PROTEIN, np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0, np.uint32),
                  type=np.zeros(0, np.int32))
    T, E, A = [], [], []
    for _ in range(nsteps // every):
        sim.run(every)
        obs = sim.observables()
        T.append(obs["temperature"])
        E.append(obs["potential"] / N)
        A.append(area_per_lipid(sim.get_state(orbc.LIPID)["x"]))
    k = relax // every
    for name, ours in (("T", T), ("E", E), ("A", A)):
        m1, e1 = block_mean(ours[k:])
        m0, e0 = block_mean(z[name][k:])
        assert np.isfinite(m1)
        # the error bars are small enough to mean something
        assert e0 < 0.01 * abs(m0) and e1 < 0.01 * abs(m1), (name, e0, e1)
        assert abs(m1 - m0) < 4 * np.hypot(e0, e1), (name, m0, e0, m1, e1)
    # fluctuation-dissipation: the reference and this run both sit at kT
    assert abs(block_mean(T[k:])[0] / float(z["kT"]) - 1) < 0.03

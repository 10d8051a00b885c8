"""Write the golden fixtures of tests/golden/ from the OpenRBC reference itself.

The reference (oracle/_ref/libref_openrbc.so, compiled in place from the
reference sources by oracle/Makefile) is driven through oracle/ref_harness.cpp
on small synthetic inputs; inputs and the reference's outputs are stored as
.npz data.  Run from the repository root:

    python tests/golden/make_golden.py

The fixtures are data only (inputs and expected outputs, FP32 as the shipped
reference computes them); the tests compare the FP64 oracle against them.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle import Reference  # noqa: E402
from _systems import patch, split  # noqa: E402


def containers(s):
    lip, prot = split(s)
    bt, bi, bj = s["bonds"]
    tag = s["id"].astype(np.int32)
    return lip, prot, (bt.astype(np.int32), tag[bi], tag[bj])


def main():
    R = Reference()
    out = {}
    ff = R.ff_tables()
    np.savez_compressed(os.path.join(HERE, "ff_tables.npz"), **ff)

    # compute_pairwise_fused + compute_bonded on two membrane patches
    for seed in (3, 8):
        s = patch(12, 12, seed=seed)
        lip, prot, bonds = containers(s)
        lf, lt, pf, pt = R.forces(s["x"][lip], s["n"][lip], s["x"][prot], s["n"][prot], s["type"][prot],
                                  s["id"][prot], bonds)
        np.savez_compressed(os.path.join(HERE, f"forces_patch{seed}.npz"),
                            lx=s["x"][lip].astype(np.float32), ln=s["n"][lip].astype(np.float32),
                            px=s["x"][prot].astype(np.float32), pn=s["n"][prot].astype(np.float32),
                            ptype=s["type"][prot].astype(np.int32), ptag=s["id"][prot].astype(np.int32),
                            bt=bonds[0], bi=bonds[1], bj=bonds[2], lf=lf, lt=lt, pf=pf, pt=pt)

    # verlet_langevin at kT = 0 (deterministic) and the two fused NH halves
    s = patch(8, 8, seed=2)
    lip, prot, _ = containers(s)
    rng = np.random.default_rng(4)
    n = len(s["type"])
    f = rng.normal(0, 3, (n, 3)); t = rng.normal(0, 1, (n, 3))

    def state(idx):
        return {k: np.ascontiguousarray(a[idx], dtype=np.float32)
                for k, a in (("x", s["x"]), ("v", s["v"]), ("n", s["n"]), ("o", s["o"]), ("f", f), ("t", t))}

    L, P = state(lip), state(prot)
    inp = {f"l{k}": v.copy() for k, v in L.items()}
    inp.update({f"p{k}": v.copy() for k, v in P.items()})
    ptype = s["type"][prot].astype(np.int32)
    R.langevin(L, P, ptype, 0.0, 0.01, 0.01)
    np.savez_compressed(os.path.join(HERE, "langevin_T0.npz"), ptype=ptype, **inp,
                        **{f"out_l{k}": v for k, v in L.items()}, **{f"out_p{k}": v for k, v in P.items()})
    for which in (0, 1):
        L, P = state(lip), state(prot)
        z = R.nh(which, L, P, ptype, 0.22, 0.01, 0.05, 7.0)
        np.savez_compressed(os.path.join(HERE, f"nh{which}.npz"), ptype=ptype, zeta_in=0.05, Q=7.0, kT=0.22,
                            zeta_out=z, **inp, **{f"out_l{k}": v for k, v in L.items()},
                            **{f"out_p{k}": v for k, v in P.items()})

    # temperature
    T = R.temperature(s["v"][lip].astype(np.float32), s["v"][prot].astype(np.float32), ptype)
    np.savez_compressed(os.path.join(HERE, "temperature.npz"), lv=s["v"][lip].astype(np.float32),
                        pv=s["v"][prot].astype(np.float32), ptype=ptype, T=T)

    # Morton keys and the reorder_morton permutation
    pts = rng.uniform(-600, 600, (2000, 3)).astype(np.float32)
    np.savez_compressed(os.path.join(HERE, "morton.npz"), pts=pts, keys=R.morton(pts),
                        reordered=R.reorder_morton(pts[:500]))

    # k-d tree queries and Voronoi partition
    cen = rng.uniform(-30, 30, (400, 3)).astype(np.float32)
    q = rng.uniform(-35, 35, (1500, 3)).astype(np.float32)
    near = R.kdtree_nearest(cen, q)
    within = R.kdtree_within(cen, q[:200], 9.0)
    counts = np.array([len(w) for w in within], np.int32)
    flat = np.concatenate(within).astype(np.int32)
    aff = R.partition(q, cen)
    np.savez_compressed(os.path.join(HERE, "voronoi.npz"), centroids=cen, queries=q, nearest=near,
                        within_counts=counts, within_flat=flat, partition=aff)

    # stray-lipid cleanup (delete_lipid) on a jittered sheet with lifted lipids
    sheet = patch(40, 40, seed=11, proteins=False)["x"]
    lift = rng.choice(len(sheet), 24, replace=False)
    sheet[lift, 2] += rng.uniform(0.4, 6.0, len(lift))
    xs = sheet.astype(np.float32)
    cen = xs[7::14].copy()
    cen[:, 2] = 0.0
    aff, keep = R.stray(xs, cen, 2.5)
    # no lipid within 1e-4 (relative) of its cell's threshold: FP32 and FP64 decide alike
    d2 = np.sum((xs.astype(np.float64) - cen[aff].astype(np.float64)) ** 2, 1)
    thr = np.zeros(len(cen))
    for k in range(len(cen)):
        m = np.sort(d2[aff == k])
        thr[k] = m[len(m) // 2] * 6.25 if len(m) else 0.0
    assert np.min(np.abs(d2 / np.maximum(thr[aff], 1e-300) - 1)) > 1e-4
    np.savez_compressed(os.path.join(HERE, "stray.npz"), x=xs, centroids=cen, tol=2.5, aff=aff, keep=keep)
    print("stray cleanup: %d of %d lipids removed" % (int((~keep).sum()), len(keep)))

    # trajectory frame (all five fields) and LAMMPS topology written by the reference
    import tempfile
    s = patch(12, 12, seed=2)
    lip, prot = split(s)
    f32 = lambda a: np.ascontiguousarray(a, np.float32)
    lx, ln, lv = f32(s["x"][lip]), f32(s["n"][lip]), f32(s["v"][lip])
    px, pn, pv = f32(s["x"][prot]), f32(s["n"][prot]), f32(s["v"][prot])
    ptype = s["type"][prot].astype(np.int32)
    ptag = s["id"][prot].astype(np.int32)
    lbase = int(len(prot) + 1)
    assert np.array_equal(s["id"][lip], np.arange(lbase, lbase + len(lip)))
    cen = lx[3::14].copy()
    laff, paff = R.partition(lx, cen), R.partition(px, cen)
    bt, bi, bj = s["bonds"]
    bt = bt.astype(np.int32); bit = s["id"][bi].astype(np.int32); bjt = s["id"][bj].astype(np.int32)
    with tempfile.TemporaryDirectory() as tmp:
        fp, tp = os.path.join(tmp, "f.orbc"), os.path.join(tmp, "t.data")
        R.save_frame(fp, lx, lv, ln, laff, lbase, px, pv, pn, paff, ptype, ptag, 0, 31)
        R.save_frame(fp, lx, lv, ln, laff, lbase, px, pv, pn, paff, ptype, ptag, 5, 7, append=True)
        R.save_topology(tp, lx, lbase, px, ptype, ptag, bt, bit, bjt)
        frame = np.frombuffer(open(fp, "rb").read(), np.uint8)
        topo = np.frombuffer(open(tp, "rb").read(), np.uint8)
    np.savez_compressed(os.path.join(HERE, "frame.npz"), lx=lx, ln=ln, lv=lv, px=px, pn=pn, pv=pv, ptype=ptype,
                        ptag=ptag, lbase=lbase, centroids=cen, laff=laff, paff=paff, bt=bt, bi=bit, bj=bjt,
                        frame=frame, topo=topo)
    print("frame fixture: %d bytes, topology %d bytes" % (len(frame), len(topo)))

    # the reference's orbc-util (oracle/_ref/orbc-util) on a trajectory whose
    # lipid count shrinks, as after stray cleanup: frame 0 with every field,
    # frame 5 with the default fields, frame 9 with 5 lipids fewer
    import subprocess
    tool = os.path.join(ROOT, "oracle", "_ref", "orbc-util")
    util = {}
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "traj.orbc")
        R.save_frame(src, lx, lv, ln, laff, lbase, px, pv, pn, paff, ptype, ptag, 0, 31)
        R.save_frame(src, lx * 1.01, lv, ln, laff, lbase, px, pv, pn, paff, ptype, ptag, 5, 7, append=True)
        k = len(lx) - 5
        R.save_frame(src, lx[:k] * 0.99, lv[:k], ln[:k], laff[:k], lbase, px, pv, pn, paff, ptype, ptag, 9, 3,
                     append=True)
        util["orbc"] = np.frombuffer(open(src, "rb").read(), np.uint8)
        run = lambda *a: subprocess.run([tool, *a], cwd=tmp, check=True, capture_output=True)
        run("convert", "traj.orbc", "traj.lammpstrj")
        util["lammpstrj"] = np.frombuffer(open(os.path.join(tmp, "traj.lammpstrj"), "rb").read(), np.uint8)
        run("convert", "traj.orbc", "traj.%d.namdbin")
        for step in (0, 5, 9):
            util[f"namdbin_{step}"] = np.frombuffer(open(os.path.join(tmp, f"traj.{step}.namdbin"), "rb").read(),
                                                    np.uint8)
        for sel in (("5",), ("0", "9")):
            run("extract", "traj.orbc", "ext.orbc", *sel)
            util["extract_" + "_".join(sel)] = np.frombuffer(open(os.path.join(tmp, "ext.orbc"), "rb").read(),
                                                             np.uint8)
    np.savez_compressed(os.path.join(HERE, "orbc_util.npz"), **util)
    print("orbc-util fixture:", {k: len(v) for k, v in util.items()})
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()

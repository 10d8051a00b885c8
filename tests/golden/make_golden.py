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

from oracle import REF_OPENRBC, Oracle, Reference  # noqa: E402
from _systems import contact_patch, patch, split  # noqa: E402


def containers(s):
    lip, prot = split(s)
    bt, bi, bj = s["bonds"]
    tag = s["id"].astype(np.int32)
    return lip, prot, (bt.astype(np.int32), tag[bi], tag[bj])


def main():
    R = Reference()
    if len(sys.argv) > 1:          # only the named fixture groups
        for name in sys.argv[1:]:
            if name == "minimize_fixture":
                minimize_fixture()
            else:
                globals()[name](R)
        return
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
    contact_fixtures(R)
    voronoi_stencil_fixture(R)
    init_rbc_fixture(R)
    minimize_fixture()
    langevin_stats_fixture(R)
    print("wrote fixtures to", HERE)


def contact_fixtures(R):
    """compute_pairwise_fused + compute_bonded on the contact patch, whose
    pairs include every WCA and protein-protein kind (one cell: all pairs)."""
    s = contact_patch(seed=0)
    lip, prot, bonds = containers(s)
    lf, lt, pf, pt = R.forces(s["x"][lip], s["n"][lip], s["x"][prot], s["n"][prot], s["type"][prot],
                              s["id"][prot], bonds)
    np.savez_compressed(os.path.join(HERE, "forces_contacts.npz"),
                        lx=s["x"][lip].astype(np.float32), ln=s["n"][lip].astype(np.float32),
                        px=s["x"][prot].astype(np.float32), pn=s["n"][prot].astype(np.float32),
                        ptype=s["type"][prot].astype(np.int32), ptag=s["id"][prot].astype(np.int32),
                        bt=bonds[0], bi=bonds[1], bj=bonds[2], lf=lf, lt=lt, pf=pf, pt=pt)
    print("contact patch: %d lipids, %d proteins" % (len(lip), len(prot)))


def voronoi_stencil_fixture(R):
    """compute_pairwise_fused over 20+ Voronoi cells with given centroids (the
    reference's cell-stencil sweep, radii 9 / 8 / 6): cells of ~30 lipids, so
    that every radius drops pairs that lie within the cutoff."""
    s = contact_patch(seed=3)
    x = s["x"].astype(np.float32)
    t = s["type"]
    rng = np.random.default_rng(5)
    lipids = np.where(t == 0)[0]
    cen = x[rng.choice(lipids, len(lipids) // 30, replace=False)].copy()
    aff = R.partition(x, cen)
    lip, prot = split(s)
    lo = lip[np.argsort(aff[lip], kind="stable")]
    po = prot[np.argsort(aff[prot], kind="stable")]
    ncell = len(cen)
    ls = np.searchsorted(aff[lo], np.arange(ncell + 1)).astype(np.int32)
    ps = np.searchsorted(aff[po], np.arange(ncell + 1)).astype(np.int32)
    bt, bi, bj = s["bonds"]
    tag = s["id"].astype(np.int32)
    bonds = (bt.astype(np.int32), tag[bi], tag[bj])
    lf, lt, pf, pt = R.forces_cells(x[lo], s["n"][lo], x[po], s["n"][po], t[po], s["id"][po], bonds, cen, ls, ps)
    np.savez_compressed(os.path.join(HERE, "voronoi_forces.npz"),
                        lx=x[lo], ln=s["n"][lo].astype(np.float32), px=x[po], pn=s["n"][po].astype(np.float32),
                        ptype=t[po].astype(np.int32), ptag=s["id"][po].astype(np.int32), bt=bonds[0], bi=bonds[1],
                        bj=bonds[2], centroids=cen, laff=aff[lo].astype(np.int32), paff=aff[po].astype(np.int32),
                        lcell_start=ls, pcell_start=ps, lf=lf, lt=lt, pf=pf, pt=pt)
    print("voronoi stencil: %d cells" % ncell)


def _write_mesh(prefix, v, e, f):
    np.savetxt(prefix + ".vert.txt", v, fmt="%.17g")
    np.savetxt(prefix + ".bond.txt", e, fmt="%d")
    np.savetxt(prefix + ".face.txt", f, fmt="%d")


def init_rbc_fixture(R):
    """init_rbc on the icosphere of subdivision 1 (42 vertices), seed 1."""
    import tempfile
    import openrbc_amd.systems as systems
    v, e, f = systems.icosphere(1)
    with tempfile.TemporaryDirectory() as d:
        prefix = os.path.join(d, "ico1")
        _write_mesh(prefix, v, e, f)
        g = R.init_rbc(prefix, 1, 1.05)
    np.savez_compressed(os.path.join(HERE, "init_rbc_ico1.npz"), vert=v, edge=e.astype(np.int32),
                        face=f.astype(np.int32), seed=1, rho=1.05, area=g["area"],
                        **{k: g[k] for k in ("lx", "ln", "px", "pn", "ptype", "ptag", "bt", "bi", "bj")})
    print("init_rbc: %d lipids, %d proteins, %d bonds" % (len(g["lx"]), len(g["px"]), len(g["bt"])))


def _run_openrbc(v, e, f, extra):
    """The reference program on a mesh: init, Voronoi init, one minimisation
    step, frames before and after it with every field."""
    import subprocess
    import tempfile
    import openrbc_amd as orbc
    with tempfile.TemporaryDirectory() as d:
        _write_mesh(os.path.join(d, "mesh"), v, e, f)
        cmd = [REF_OPENRBC, "-i", "trimesh", "-m", "mesh", "-E", "1", "-t", "0", "-d", "1", "--dump-field", "31",
               "--random-seed", "1"] + extra
        subprocess.run(cmd, cwd=d, check=True, capture_output=True, env=dict(os.environ, OMP_NUM_THREADS="1"))
        return [orbc.read_frame(os.path.join(d, "cell.orbc"), k) for k in range(2)]


def _match_frames(a, b):
    """Particles of frame a in frame b: proteins by tag; lipids, whose tags
    are positional in the reference, by nearest position (they move less than
    the minimiser's 0.05 cap)."""
    from scipy.spatial import cKDTree
    nl = int((a["type"] == 0).sum())
    ia = np.argsort(a["tag"][nl:]) + nl
    ib = np.argsort(b["tag"][nl:]) + nl
    dd, jj = cKDTree(b["x"][:nl]).query(a["x"][:nl])
    assert len(np.unique(jj)) == nl and dd.max() < 0.2
    return np.concatenate([np.arange(nl), ia]), np.concatenate([jj, ib])


def minimize_fixture():
    """The capped steepest-descent step of the reference's minimisation loop
    (src/openrbc.cpp:88-146), run by the reference program itself on the
    icosphere of subdivision 0: frame 0 (x0, n0) and frame 1 (x1, n1 and the
    force f0 that moved it).  'capped' uses the default caps (0.05 / 0.05),
    'free' caps of 1e9 (every step is dt; its lipids move too far to be
    matched by position, so only proteins are kept)."""
    import openrbc_amd.systems as systems
    v, e, f = systems.icosphere(0)
    out = dict(vert=v, edge=e.astype(np.int32), face=f.astype(np.int32))
    for name, extra in (("capped", []), ("free", ["--opt-max-move", "1e9", "--opt-max-rotate", "1e9"])):
        a, b = _run_openrbc(v, e, f, extra)
        if name == "capped":
            A, B = _match_frames(a, b)
        else:
            nl = int((a["type"] == 0).sum())
            A, B = np.argsort(a["tag"][nl:]) + nl, np.argsort(b["tag"][nl:]) + nl
        out.update({f"{name}_x0": a["x"][A], f"{name}_n0": a["n"][A], f"{name}_x1": b["x"][B],
                    f"{name}_n1": b["n"][B], f"{name}_f0": b["f"][B], f"{name}_type": a["type"][A].astype(np.int32),
                    f"{name}_tag": np.where(a["type"][A] == 0, -1, a["tag"][A]).astype(np.int32)})
    np.savez_compressed(os.path.join(HERE, "minimize.npz"), **out)
    print("minimize: %d particles (capped), %d proteins (free)" % (len(out["capped_x0"]), len(out["free_x0"])))


def vesicle_observables(O, xs, ns, vs):
    """Per snapshot of a lipid vesicle: temperature sum m v^2 / 3N (m = 1),
    potential energy per particle (closed-form pair energies) and area per
    lipid (tests/_stats.py)."""
    from _stats import area_per_lipid
    T, E, A = [], [], []
    for x, n, v in zip(xs, ns, vs):
        x = x.astype(np.float64); n = n.astype(np.float64); v = v.astype(np.float64)
        N = len(x)
        T.append(np.sum(v * v) / (3 * N))
        _, _, e, _ = O.forces(x, n, np.zeros(N, np.int32))
        E.append(e.sum() / N)
        A.append(area_per_lipid(x))
    return np.array(T), np.array(E), np.array(A)


def langevin_stats_fixture(R):
    """Reference Langevin statistics: a lipid vesicle of 1900 lipids (radius
    12, minimised by the oracle's capped descent), Maxwell velocities at
    kT = 0.22, then 30000 steps of the reference's verlet_langevin with exact
    pairs; T, E/N and area per lipid every 20 steps (the first 6000 steps are
    relaxation)."""
    import openrbc_amd.systems as systems
    O = Oracle()
    sp = systems.random_sphere(r=12.0, seed=7)
    x, n = sp["lipid_x"].copy(), sp["lipid_n"].copy()
    N = len(x)
    ty = np.zeros(N, np.int32)
    for _ in range(400):
        f, t, _, _ = O.forces(x, n, ty)
        st = dict(x=x, n=n, f=f, t=np.cross(n, t), type=ty)
        O.minimize_move(st, 0.05, 0.05, 0.01)
        x, n = st["x"], st["n"]
    rng = np.random.default_rng(3)
    v = rng.normal(0, np.sqrt(0.22), (N, 3))
    x32, n32, v32 = x.astype(np.float32), n.astype(np.float32), v.astype(np.float32)
    every, nsteps = 20, 30000
    xs, ns, vs = R.langevin_lipids(x32, n32, v32, 0.22, 0.01, 0.01, 11, nsteps, every)
    T, E, A = vesicle_observables(O, xs, ns, vs)
    np.savez_compressed(os.path.join(HERE, "langevin_stats.npz"), x=x32, n=n32, v=v32, kT=0.22, eta=0.01, dt=0.01,
                        every=every, nsteps=nsteps, relax=6000, T=T, E=E, A=A)
    k = 6000 // every
    print("langevin stats: <T> %.5f <E/N> %.5f <A/N> %.5f" % (T[k:].mean(), E[k:].mean(), A[k:].mean()))


if __name__ == "__main__":
    main()

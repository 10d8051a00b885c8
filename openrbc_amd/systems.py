"""Synthetic inputs: meshes and particle systems.

* ``init_rbc``: mesh -> glycophorin, actin, band-3, spectrin and lipids with
  the rules of src/init_rbc.h:30-337 (C++ in liborbc_hip, ``orbc_init_rbc``).
* ``icosphere``: subdivided icosahedron (level 3: 642 V / 1920 E / 1280 F),
  written like example-large's vert/bond/face arrays (0-based).
* ``evans_fung``: the biconcave Evans-Fung RBC surface, icosphere-triangulated.
* ``load_example_large``: the full-size Evans-Fung RBC mesh that OpenRBC ships
  (example-large/rbc.{vert,bond,face}.txt, 9128 V), stored as
  openrbc_amd/data/rbc_example_large.npz.
* ``random_sphere``: init_random_sphere (src/init_random.h:27-66), lipids only.
* ``planar_patch``: hexagonal lipid lattice (periodic patch workloads).
* ``vesicle_ensemble``: many lipid vesicles of uneven size (batched ensembles).
* ``planar_membrane``: periodic planar lipid lattice with a spectrin network beneath.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def load_example_large():
    z = np.load(os.path.join(_DATA, "rbc_example_large.npz"))
    return z["vert"].astype(np.float64), z["edge"].astype(np.int32), z["face"].astype(np.int32)


def icosphere(level: int = 3, radius: float = 1.0):
    """Icosphere mesh (vert (nv,3), edge (ne,2), face (nf,3)), faces CCW seen from outside."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    faces = list(f)
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts) * radius
    F = np.array(faces, dtype=np.int32)
    E = set()
    for a, b, c in F:
        for p, q in ((a, b), (b, c), (c, a)):
            E.add((min(p, q), max(p, q)))
    return V, np.array(sorted(E), dtype=np.int32), F


def _icosahedron():
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    return np.array(v, np.float64), f


def geodesic(nu: int, radius: float = 1.0):
    """Class-I geodesic sphere of frequency nu: every icosahedron face split
    into nu^2 triangles, points projected to the sphere.  10 nu^2 + 2 vertices,
    30 nu^2 edges, 20 nu^2 faces (CCW seen from outside)."""
    if nu < 1:
        raise ValueError("nu >= 1")
    V0, F0 = _icosahedron()
    V0 /= np.linalg.norm(V0, axis=1, keepdims=True)
    index = {}
    verts = []

    def vid(fa, i, j):
        # lattice point A + i (B - A)/nu + j (C - A)/nu of face fa = (A, B, C), keyed by its
        # exact barycentric position over the icosahedron's vertex ids so shared points merge
        a, b, c = fa
        w = {a: nu - i - j, b: i, c: j}
        key = tuple(sorted((k, x) for k, x in w.items() if x))
        if key not in index:
            p = sum(V0[k] * x for k, x in key) / nu
            verts.append(p / np.linalg.norm(p))
            index[key] = len(verts) - 1
        return index[key]

    faces = []
    for fa in F0:
        for i in range(nu):
            for j in range(nu - i):
                p, q, r = vid(fa, i, j), vid(fa, i + 1, j), vid(fa, i, j + 1)
                faces.append((p, q, r))
                if i + j < nu - 1:
                    faces.append((q, vid(fa, i + 1, j + 1), r))
    V = np.array(verts) * radius
    F = np.array(faces, dtype=np.int32)
    E = np.unique(np.sort(np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]]), axis=1), axis=0)
    return V, E.astype(np.int32), F


def evans_fung(nu: int = 34, D0: float = 7.82, a0: float = 0.0518, a1: float = 2.0026, a2: float = -4.491,
               prefactor: float | None = None):
    """Biconcave Evans-Fung surface z = +-P sqrt(1 - 4r^2/D0^2) (a0 + a1 r^2/D0^2 + a2 r^4/D0^4),
    P = D0 by default (the shape OpenRBC's example-large mesh follows; P = D0/2
    gives a cell half as thick), triangulated by mapping a geodesic sphere of
    frequency nu (10 nu^2 + 2 vertices) onto it: each hemisphere goes onto the
    disc r <= D0/2 area-preservingly ((2r/D0)^2 = 1 - |cos theta|), keeping the
    connectivity and the outward face orientation.  Units are microns; the
    initialiser rescales to its 80 nm mean edge (src/init_rbc.h:70-82)."""
    P = D0 if prefactor is None else prefactor
    v, e, f = geodesic(nu)
    c = np.abs(v[:, 2])
    rho = np.sqrt(np.clip(1.0 - c, 0.0, 1.0))        # 2r / D0
    phi = np.arctan2(v[:, 1], v[:, 0])
    r = 0.5 * D0 * rho
    q = rho * rho
    z = np.sign(v[:, 2]) * P * np.sqrt(np.clip(1.0 - q, 0.0, 1.0)) * (a0 + a1 * q / 4.0 + a2 * q * q / 16.0)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1), e, f


def init_rbc(vert, edge, face, rho: float = 1.05, seed: int = 1):
    """Particle system on a triangle mesh (src/init_rbc.h rules).

    Returns a dict with lipid/protein containers (FP64 x, n; ids) and bonds
    (type, id_i, id_j), ids following the reference tags."""
    from . import load_library
    from . import _System  # noqa: F401
    L = load_library()
    v = np.ascontiguousarray(vert, dtype=np.float64)
    e = np.ascontiguousarray(edge, dtype=np.int32)
    f = np.ascontiguousarray(face, dtype=np.int32)
    s = _system_struct()
    rc = L.orbc_init_rbc(v.ctypes.data_as(C.POINTER(C.c_double)), len(v), e.ctypes.data_as(C.POINTER(C.c_int)),
                         len(e), f.ctypes.data_as(C.POINTER(C.c_int)), len(f), float(rho), int(seed), C.byref(s))
    if rc != 0:
        raise RuntimeError(f"orbc_init_rbc failed ({rc})")
    try:
        nl, np_, nb = s.n_lipid, s.n_protein, s.n_bond

        def arr(p, n, dt):
            return np.ctypeslib.as_array(p, shape=(n,)).astype(dt, copy=True) if n > 0 else np.zeros(0, dt)

        out = dict(
            lipid_x=arr(s.lipid_x, 3 * nl, np.float64).reshape(-1, 3),
            lipid_n=arr(s.lipid_n, 3 * nl, np.float64).reshape(-1, 3),
            lipid_id=arr(s.lipid_id, nl, np.uint32),
            protein_x=arr(s.protein_x, 3 * np_, np.float64).reshape(-1, 3),
            protein_n=arr(s.protein_n, 3 * np_, np.float64).reshape(-1, 3),
            protein_type=arr(s.protein_type, np_, np.int32),
            protein_id=arr(s.protein_id, np_, np.uint32),
            bond_type=arr(s.bond_type, nb, np.int32),
            bond_i=arr(s.bond_i, nb, np.uint32),
            bond_j=arr(s.bond_j, nb, np.uint32),
            area=s.area,
            clutter_capped=s.clutter_capped,
        )
    finally:
        L.orbc_system_free(C.byref(s))
    return out


def _system_struct():
    from . import _System
    return _System()


def random_sphere(r: float = 100.0, rho: float = 1.05, seed: int = 0):
    """init_random_sphere: int(rho 4 pi r^2) lipids uniform on a sphere, n = x/|x|."""
    rng = np.random.default_rng(seed)
    n = int(rho * 4 * 3.14159265358 * r * r)
    theta = 2 * np.pi * rng.random(n)
    u = 2 * rng.random(n) - 1
    s = np.sqrt(1 - u * u)
    x = np.stack([r * s * np.cos(theta), r * s * np.sin(theta), r * u], 1)
    return dict(lipid_x=x, lipid_n=x / np.linalg.norm(x, axis=1, keepdims=True),
                lipid_id=np.arange(1, n + 1, dtype=np.uint32))


def vesicle_ensemble(n_members: int = 64, n_min: int = 5000, n_max: int = 300000, rho: float = 1.05,
                     kT: float = 0.22, spread: float = 0.1, gap: float = 12.0, seed: int = 0):
    """Batched-ensemble workload: lipid-only vesicles (the init_random_sphere
    rule, src/init_random.h:27-66, uniform points on a sphere, directors
    radial) with particle counts log-uniform in [n_min, n_max], member
    temperatures kT * U(1 - spread, 1 + spread), centred on a cubic lattice
    whose spacing leaves `gap` between the largest vesicles.  Ids are unique
    across members.  Returns lipid_x, lipid_n, lipid_id, group (member of each
    lipid), kT (per member), radius and centre (per member)."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(n_min), np.log(n_max), n_members)).astype(np.int64)
    radius = np.sqrt(counts / (4 * np.pi * rho))
    kts = kT * rng.uniform(1 - spread, 1 + spread, n_members)
    side = int(np.ceil(n_members ** (1 / 3)))
    pitch = 2 * radius.max() + gap
    xs, ns, gs, centres = [], [], [], []
    for m in range(n_members):
        c = pitch * np.array([m % side, (m // side) % side, m // (side * side)], np.float64)
        n = int(counts[m])
        u = 2 * rng.random(n) - 1
        phi = 2 * np.pi * rng.random(n)
        s = np.sqrt(1 - u * u)
        d = np.stack([s * np.cos(phi), s * np.sin(phi), u], 1)
        xs.append(c + radius[m] * d)
        ns.append(d)
        gs.append(np.full(n, m, np.int32))
        centres.append(c)
    x = np.vstack(xs)
    return dict(lipid_x=x, lipid_n=np.vstack(ns), lipid_id=np.arange(1, len(x) + 1, dtype=np.uint32),
                group=np.concatenate(gs), kT=kts, radius=radius, centre=np.array(centres))


def vesicle_ensemble_mesh(n_members: int = 64, n_min: int = 5000, n_max: int = 300000, rho: float = 1.05,
                          kT: float = 0.22, spread: float = 0.1, gap: float = 12.0, seed: int = 0):
    """Batched-ensemble workload on meshes: spherical cells from class-I
    geodesic meshes (icosahedron faces split nu^2 ways) through the
    initialiser (init_rbc rules, src/init_rbc.h:30-337: lipids, glycophorin,
    actin, band-3, spectrin and their bonds).  init_rbc scales a mesh to a
    mean edge of 80 nm, so a member of frequency nu holds about 3,540 nu^2
    particles: nu is drawn so that the counts are log-uniform in [n_min,
    n_max] (nu >= 1).  Member temperatures kT * U(1 - spread, 1 + spread);
    members centred on a cubic lattice whose spacing leaves `gap` between
    the largest ones; each member has its own placement seed.  Ids (and the
    bond ids) are offset so they are unique across members.  Returns the
    species arrays of init_rbc, lipid_group / protein_group (member of each
    particle), kT, nu, radius and centre (per member)."""
    rng = np.random.default_rng(seed)
    target = np.exp(rng.uniform(np.log(n_min), np.log(n_max), n_members))
    nus = np.maximum(1, np.rint(np.sqrt(target / 3540.0))).astype(np.int64)
    kts = kT * rng.uniform(1 - spread, 1 + spread, n_members)
    meshes = {nu: geodesic(int(nu)) for nu in np.unique(nus)}
    members = [init_rbc(*meshes[nu], rho=rho, seed=int(seed) * 100003 + m + 1) for m, nu in enumerate(nus)]
    radius = np.array([np.linalg.norm(s["lipid_x"], axis=1).max() for s in members])
    side = int(np.ceil(n_members ** (1 / 3)))
    pitch = 2 * radius.max() + gap
    out = {k: [] for k in ("lipid_x", "lipid_n", "lipid_id", "protein_x", "protein_n", "protein_type", "protein_id",
                           "bond_type", "bond_i", "bond_j", "lipid_group", "protein_group")}
    centres = []
    base = 0
    for m, s in enumerate(members):
        c = pitch * np.array([m % side, (m // side) % side, m // (side * side)], np.float64)
        centres.append(c)
        for sp in ("lipid", "protein"):
            out[f"{sp}_x"].append(s[f"{sp}_x"] + c)
            out[f"{sp}_n"].append(s[f"{sp}_n"])
            out[f"{sp}_id"].append(s[f"{sp}_id"].astype(np.int64) + base)
            out[f"{sp}_group"].append(np.full(len(s[f"{sp}_x"]), m, np.int32))
        out["protein_type"].append(s["protein_type"])
        out["bond_type"].append(s["bond_type"])
        out["bond_i"].append(s["bond_i"].astype(np.int64) + base)
        out["bond_j"].append(s["bond_j"].astype(np.int64) + base)
        base += len(s["lipid_x"]) + len(s["protein_x"])
    res = {k: np.concatenate(v) for k, v in out.items()}
    for k in ("lipid_id", "protein_id", "bond_i", "bond_j"):
        res[k] = res[k].astype(np.uint32)
    res.update(kT=kts, nu=nus, radius=radius, centre=np.array(centres))
    return res


def planar_membrane(nx: int = 4096, ny: int = 4096, a: float = 1.0487, edge: float = 16.84, tilt: float = 0.02,
                    jitter: float = 0.02, layer: float = 2.25, seed: int = 0):
    """Large periodic planar membrane: an nx x ny hexagonal lipid lattice
    (lattice constant a, directors +z tilted by a Gaussian of width `tilt`
    radians, positions jittered by jitter * a) tiling a box periodic in x and
    y, with a triangular spectrin network of edge ~`edge` tethered beneath it
    by the composition of src/init_rbc.h:122-233: per network vertex a
    glycophorin in the membrane and an actin `layer` below it (bond type 0);
    per edge an immobile band-3 at the midpoint in the membrane and a chain of
    19 spectrin beads between the two actins (types 3, 2 to the actins, 1 from
    the middle bead to the band-3).  Membrane proteins take the place of the
    nearest lattice lipid.  Periodic bonds are given by their nearest image.
    Returns the system dict (lipid_*, protein_*, bond_*) and box = (lo, hi,
    periodic)."""
    assert ny % 2 == 0
    rng = np.random.default_rng(seed)
    Lx, Ly = nx * a, ny * a * np.sqrt(3) / 2
    # network vertices on a triangular lattice that tiles the box
    mx = max(3, int(round(Lx / edge)))
    my = max(4, int(round(Ly / (edge * np.sqrt(3) / 2))) // 2 * 2)
    ex, ey = Lx / mx, Ly / my
    pp, qq = np.meshgrid(np.arange(mx), np.arange(my), indexing="ij")
    pp, qq = pp.ravel(), qq.ravel()
    vx = ex * (pp + 0.5 * (qq % 2))
    vy = ey * qq
    nv = len(vx)
    vid = lambda p, q: (p % mx) * my + (q % my)
    # three edges per vertex: right, and the two upward neighbours
    up_l = np.where(qq % 2 == 0, pp - 1, pp)
    up_r = np.where(qq % 2 == 0, pp, pp + 1)
    e0 = np.concatenate([vid(pp, qq)] * 3)
    e1 = np.concatenate([vid(pp + 1, qq), vid(up_l, qq + 1), vid(up_r, qq + 1)])
    ne = len(e0)

    def mi(d, L):
        return d - L * np.round(d / L)

    def wrap(v, L):
        return v - L * np.floor(v / L)

    gly = np.stack([vx, vy, np.zeros(nv)], 1)
    act = gly - np.array([0.0, 0.0, layer])
    d_e = np.stack([mi(vx[e1] - vx[e0], Lx), mi(vy[e1] - vy[e0], Ly), np.zeros(ne)], 1)
    mid = gly[e0] + 0.5 * d_e
    mid[:, 0] = wrap(mid[:, 0], Lx); mid[:, 1] = wrap(mid[:, 1], Ly)
    # spectrin: 19 beads between the actins, 1.0 in from each end (src/init_rbc.h:201-214)
    res = 19
    beg = act[e0]
    inc = d_e / (res + 1)
    u = inc / np.linalg.norm(inc, axis=1, keepdims=True)
    beg = beg + u
    inc = (d_e - 2 * u) / (res + 1)
    j = np.arange(1, res + 1)
    spec = beg[:, None, :] + inc[:, None, :] * j[None, :, None] + 0.05 * rng.uniform(-1, 1, (ne, res, 3))
    spec = spec.reshape(-1, 3)
    spec[:, 0] = wrap(spec[:, 0], Lx); spec[:, 1] = wrap(spec[:, 1], Ly)
    # proteins in tag order: glycophorin, actin, band-3, spectrin (src/init_rbc.h:122-233)
    px = np.vstack([gly, act, mid, spec])
    ptype = np.concatenate([np.full(nv, 3), np.full(nv, 4), np.full(ne, 2), np.full(ne * res, 5)]).astype(np.int32)
    pn = np.zeros_like(px); pn[:, 2] = 1.0
    np_ = len(px)
    tag = np.arange(1, np_ + 1, dtype=np.uint32)
    t_gly, t_act, t_b3 = tag[:nv], tag[nv:2 * nv], tag[2 * nv:2 * nv + ne]
    t_sp = tag[2 * nv + ne:].reshape(ne, res)
    bt = [np.zeros(nv, np.int32), np.full(ne * (res - 1), 3, np.int32), np.full(ne, 2, np.int32),
          np.full(ne, 1, np.int32), np.full(ne, 2, np.int32)]
    bi = [t_gly, t_sp[:, :-1].ravel(), t_act[e0], t_b3, t_act[e1]]
    bj = [t_act, t_sp[:, 1:].ravel(), t_sp[:, 0], t_sp[:, res // 2], t_sp[:, -1]]
    # lipid lattice minus the sites nearest to the membrane proteins
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    taken = np.zeros((nx, ny), bool)
    for q in (gly, mid):
        jq = np.round(q[:, 1] / (a * np.sqrt(3) / 2)).astype(np.int64) % ny
        iq = np.round(q[:, 0] / a - 0.5 * (jq % 2)).astype(np.int64) % nx
        taken[iq, jq] = True
    keep = ~taken.ravel()
    lx = np.stack([(a * (ii + 0.5 * (jj % 2))).ravel()[keep], (a * (np.sqrt(3) / 2) * jj).ravel()[keep]], 1)
    nl = len(lx)
    lx = np.hstack([lx, np.zeros((nl, 1))]) + rng.normal(0, jitter * a, (nl, 3)) * np.array([1.0, 1.0, 0.0])
    ln = np.zeros((nl, 3)); ln[:, 2] = 1.0
    ln += rng.normal(0, tilt, (nl, 3)) * np.array([1.0, 1.0, 0.0])
    ln /= np.linalg.norm(ln, axis=1, keepdims=True)
    lo = np.array([0.0, 0.0, -1000.0]); hi = np.array([Lx, Ly, 1000.0])
    return dict(lipid_x=lx, lipid_n=ln, lipid_id=np.arange(np_ + 1, np_ + 1 + nl, dtype=np.uint32),
                protein_x=px, protein_n=pn, protein_type=ptype, protein_id=tag,
                bond_type=np.concatenate(bt), bond_i=np.concatenate(bi).astype(np.uint32),
                bond_j=np.concatenate(bj).astype(np.uint32), box=(lo, hi, np.array([1, 1, 0], np.int32)),
                n_vertices=nv, n_edges=ne)


def planar_patch(nx: int, ny: int, a: float = 1.0487, tilt: float = 0.02, jitter: float = 0.02, seed: int = 0):
    """Hexagonal lipid lattice nx * ny in the z=0 plane, directors +z with
    Gaussian tilt (radians) and positions jittered by jitter * a."""
    rng = np.random.default_rng(seed)
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    x = a * (i + 0.5 * (j % 2))
    y = a * (np.sqrt(3) / 2) * j
    p = np.stack([x.ravel(), y.ravel(), np.zeros(x.size)], 1)
    p[:, :2] += rng.normal(0, jitter * a, (len(p), 2))
    n = np.zeros_like(p)
    n[:, 2] = 1
    n[:, :2] = rng.normal(0, tilt, (len(p), 2))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    return dict(lipid_x=p, lipid_n=n, lipid_id=np.arange(1, len(p) + 1, dtype=np.uint32))


def tilt_directors(n, sigma, rng):
    """Rotate unit vectors by Gaussian tilt of width sigma (radians)."""
    g = rng.normal(0, sigma, n.shape)
    m = n + g - (g * n).sum(1, keepdims=True) * n
    return m / np.linalg.norm(m, axis=1, keepdims=True)


def maxwell(types, kT, rng):
    """Maxwell-Boltzmann velocities at kT (masses of src/forcefield_canonical.h:40)."""
    mass = np.array([1, 4, 4, 1, 10, 10], dtype=np.float64)[np.asarray(types)]
    return rng.normal(0, 1, (len(mass), 3)) * np.sqrt(kT / mass)[:, None]


def system_to_simulation(sys: dict, sim, kT: float | None = None, tilt: float = 0.0, seed: int = 0):
    """Upload a system dict into a Simulation (optionally tilt directors and
    draw Maxwell velocities at kT)."""
    from . import LIPID, PROTEIN
    rng = np.random.default_rng(seed)
    ln = sys["lipid_n"]
    if tilt > 0:
        ln = tilt_directors(ln, tilt, rng)
    nl = len(sys["lipid_x"])
    lv = maxwell(np.zeros(nl, np.int32), kT, rng) if kT else None
    sim.set_state(LIPID, sys["lipid_x"], ln, sys["lipid_id"], v=lv)
    if "protein_x" in sys and len(sys["protein_x"]):
        pv = maxwell(sys["protein_type"], kT, rng) if kT else None
        sim.set_state(PROTEIN, sys["protein_x"], sys["protein_n"], sys["protein_id"], v=pv,
                      type=sys["protein_type"])
        if len(sys["bond_type"]):
            sim.set_bonds(sys["bond_type"], sys["bond_i"], sys["bond_j"])
    else:
        sim.set_state(PROTEIN, np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0, np.uint32),
                      type=np.zeros(0, np.int32))
    return sim

"""Small synthetic membrane patches for the parity tests (numpy only).

A jittered hexagonal lipid sheet (directors near +z, the reference's lipid
geometry) with glycophorin (type 3) and immobile band-3 (type 2) in the
sheet, actin (type 4) 2.25 below, and spectrin chains (type 5) bonded the way
src/init_rbc.h:140-233 bonds them.  Ids follow the reference's tag rule:
proteins 1..np, lipids np+1.. (src/container.h:122-126, src/init_rbc.h:306).
"""
import numpy as np


def patch(nx=12, ny=12, seed=0, proteins=True, tilt=0.05, jitter=0.05, spacing=1.0487, z0=0.0):
    rng = np.random.default_rng(seed)
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    x = spacing * (ii + 0.5 * (jj % 2))
    y = spacing * (np.sqrt(3) / 2) * jj
    pts = np.stack([x.ravel(), y.ravel(), np.full(x.size, z0)], 1)
    pts += rng.normal(0, jitter * spacing, pts.shape)
    nrm = np.tile([0.0, 0.0, 1.0], (len(pts), 1)) + rng.normal(0, tilt, (len(pts), 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    types = np.zeros(len(pts), np.int32)
    bonds = []  # (type, i, j) by index into the combined array
    extra_x, extra_n, extra_t = [], [], []
    if proteins:
        # replace a few sheet sites by glycophorin / band-3; put actin below glycophorin
        sites = rng.choice(len(pts), size=max(2, len(pts) // 40), replace=False)
        gly = sites[: len(sites) // 2]
        b3 = sites[len(sites) // 2:]
        types[gly] = 3
        types[b3] = 2
        n0 = len(pts)
        actins = []
        for g in gly:
            a = n0 + len(extra_x)
            extra_x.append(pts[g] - 2.25 * np.array([0, 0, 1.0]) + rng.normal(0, 0.05, 3))
            extra_n.append(nrm[g])
            extra_t.append(4)
            bonds.append((0, g, a))
            actins.append(a)
        # spectrin chains of 4 beads between consecutive actins, band-3 tethered mid-chain
        for k in range(len(actins) - 1):
            p0 = np.array(extra_x[actins[k] - n0]); p1 = np.array(extra_x[actins[k + 1] - n0])
            prev = actins[k]
            for s in range(4):
                q = p0 + (p1 - p0) * (s + 1) / 5.0 + rng.normal(0, 0.05, 3)
                idx = n0 + len(extra_x)
                extra_x.append(q); extra_n.append(q.copy()); extra_t.append(5)
                bonds.append((2 if s == 0 else 3, prev, idx))
                if s == 1 and k < len(b3):
                    bonds.append((1, b3[k], idx))
                prev = idx
            bonds.append((2, actins[k + 1], prev))
    if extra_x:
        pts = np.vstack([pts, np.array(extra_x)])
        nrm = np.vstack([nrm, np.array(extra_n)])
        types = np.concatenate([types, np.array(extra_t, np.int32)])
    n = len(pts)
    vel = rng.normal(0, 0.3, (n, 3))
    omg = rng.normal(0, 0.1, (n, 3))
    # reference tags: proteins first (1..np), then lipids (np+1..)
    prot = np.where(types != 0)[0]
    lip = np.where(types == 0)[0]
    ids = np.zeros(n, np.uint32)
    ids[prot] = np.arange(1, len(prot) + 1)
    ids[lip] = np.arange(len(prot) + 1, len(prot) + 1 + len(lip))
    b = np.array(bonds, dtype=np.int64).reshape(-1, 3)
    return dict(x=pts.astype(np.float64), n=nrm, v=vel, o=omg, type=types, id=ids,
                bonds=(b[:, 0].astype(np.int32), b[:, 1], b[:, 2]))


def split(s):
    """Combined system -> reference containers (lipid arrays, protein arrays)."""
    t = s["type"]
    lip = np.where(t == 0)[0]
    prot = np.where(t != 0)[0]
    # proteins ordered by tag so tag2idx is the identity shift
    prot = prot[np.argsort(s["id"][prot])]
    return lip, prot


def periodic_patch(nx=16, ny=16, seed=0, spacing=1.0487, origin=(-3.0, 5.0), tilt=0.05, jitter=0.05):
    """Hexagonal lipid sheet that tiles a periodic box in x and y (ny even),
    z open: glycophorin (type 3) and band-3 (type 2) replace sheet sites
    (lipids nearer than 1.685 to them are removed),
    actin (type 4) sits 2.25 below each glycophorin with a type-0 bond.
    Positions are jittered and may fall outside the box (the device wraps
    them).  Returns the system dict plus box = (lo, hi, periodic)."""
    assert ny % 2 == 0
    rng = np.random.default_rng(seed)
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    x = spacing * (ii + 0.5 * (jj % 2)) + origin[0]
    y = spacing * (np.sqrt(3) / 2) * jj + origin[1]
    pts = np.stack([x.ravel(), y.ravel(), np.zeros(x.size)], 1)
    pts += rng.normal(0, jitter * spacing, pts.shape)
    nrm = np.tile([0.0, 0.0, 1.0], (len(pts), 1)) + rng.normal(0, tilt, (len(pts), 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    types = np.zeros(len(pts), np.int32)
    nprot = max(2, len(pts) // 30)
    Lb = np.array([nx * spacing, ny * spacing * np.sqrt(3) / 2])

    def pdist(a, b):                      # in-plane minimum-image distances
        d = a[:, None, :2] - b[None, :, :2]
        d -= Lb * np.round(d / Lb)
        return np.linalg.norm(d, axis=2)

    def spread(cand, k, dmin):            # up to k sites of cand, pairwise farther than dmin
        out = []
        for c in rng.permutation(cand):
            if len(out) == k:
                break
            if not out or pdist(pts[[c]], pts[out]).min() > dmin:
                out.append(int(c))
        return np.array(out, dtype=np.int64)

    # membrane proteins spaced as on the reference's mesh (vertices and edge
    # midpoints, src/init_rbc.h:122-172): no two within the protein-protein
    # cutoff 2.6, band-3 more than 4 from a glycophorin (a band-3 next to the
    # actin under a glycophorin would sit deep inside their sigma-3.4 WCA core)
    gly = spread(np.arange(len(pts)), nprot // 2, 2.6)
    far = np.where(pdist(pts, pts[gly]).min(axis=1) > 4.0)[0]
    b3 = spread(far, nprot - len(gly), 2.6)
    types[gly] = 3
    types[b3] = 2
    # lipids closer than reqlp = 1.685 to a protein are removed, as the
    # reference's init does (src/init_rbc.h:243-282)
    mem = np.concatenate([gly, b3])
    keep = (types != 0) | (pdist(pts, pts[mem]).min(axis=1) >= 1.685)
    idx = np.cumsum(keep) - 1
    pts, nrm, types = pts[keep], nrm[keep], types[keep]
    gly, b3 = idx[gly], idx[b3]
    n0 = len(pts)
    act = pts[gly] - np.array([0, 0, 2.25]) + rng.normal(0, 0.05, (len(gly), 3))
    pts = np.vstack([pts, act])
    nrm = np.vstack([nrm, nrm[gly]])
    types = np.concatenate([types, np.full(len(gly), 4, np.int32)])
    bonds = [(0, g, n0 + k) for k, g in enumerate(gly)]
    n = len(pts)
    prot = np.where(types != 0)[0]
    lip = np.where(types == 0)[0]
    ids = np.zeros(n, np.uint32)
    ids[prot] = np.arange(1, len(prot) + 1)
    ids[lip] = np.arange(len(prot) + 1, len(prot) + 1 + len(lip))
    b = np.array(bonds, dtype=np.int64).reshape(-1, 3)
    lo = np.array([origin[0], origin[1], -1000.0])
    hi = np.array([origin[0] + nx * spacing, origin[1] + ny * spacing * np.sqrt(3) / 2, 1000.0])
    return dict(x=pts, n=nrm, v=rng.normal(0, 0.3, (n, 3)), o=rng.normal(0, 0.1, (n, 3)), type=types, id=ids,
                bonds=(b[:, 0].astype(np.int32), b[:, 1], b[:, 2]),
                box=(lo, hi, np.array([1, 1, 0], np.int32)))


# Cutoffs of every type pair (App. of the force field, src/forcefield_canonical.h:40-99):
# polynomial rc and WCA cut 1.1225 sigma, 0 where the pair has no such term.
_POLY_RC = np.zeros((6, 6))
_POLY_RC[:4, :4] = 2.6
_LJ_RC = np.zeros((6, 6))
for _a, _b, _s in ((0, 4, 1.0), (0, 5, 1.0), (1, 4, 3.4), (1, 5, 3.4), (2, 4, 3.4), (2, 5, 1.0)):
    _LJ_RC[_a, _b] = _LJ_RC[_b, _a] = 1.1225 * _s


def pair_kinds(x, types, pairs=None):
    """Counts of the interacting unordered pairs of each kind: {(ta, tb, 'poly'|'lj'): count}
    with ta <= tb (the polynomial term is tested first, as the reference does)."""
    n = len(x)
    d = np.linalg.norm(x[:, None, :] - x[None, :, :], axis=2)
    iu, ju = np.triu_indices(n, 1)
    r = d[iu, ju]
    ta, tb = types[iu], types[ju]
    poly = r < _POLY_RC[ta, tb]
    lj = ~poly & (r < _LJ_RC[ta, tb])
    out = {}
    for m, kind in ((poly, "poly"), (lj, "lj")):
        for a, b in zip(np.minimum(ta[m], tb[m]), np.maximum(ta[m], tb[m])):
            out[(int(a), int(b), kind)] = out.get((int(a), int(b), kind), 0) + 1
    return out


def _clear_cutoffs(x, types, rng, margin=2e-3, movable=None, iters=200):
    """Nudges particles until no pair distance lies within `margin` (relative)
    of a cutoff of its type pair, so FP32 and FP64 select the same pairs."""
    for _ in range(iters):
        d = np.linalg.norm(x[:, None, :] - x[None, :, :], axis=2)
        np.fill_diagonal(d, np.inf)
        bad = np.zeros(len(x), bool)
        for rc in (_POLY_RC, _LJ_RC):
            c = rc[types[:, None], types[None, :]]
            near = (c > 0) & (np.abs(d / np.where(c > 0, c, 1.0) - 1.0) < margin)
            bad |= near.any(1)
        if movable is not None:
            bad &= movable
        if not bad.any():
            return x
        k = np.where(bad)[0]
        x[k] += rng.normal(0, 0.02, (len(k), 3))
    raise RuntimeError("could not clear the cutoffs")


def contact_patch(nx=30, ny=30, seed=0, spacing=1.0487, tilt=0.05, jitter=0.04):
    """A lipid sheet with protein clusters that exercise every non-bonded term
    of the force field, each kept clear of its cutoff:

    * lipid-protein anisotropic terms with types 1, 2, 3 (lipids kept beyond 1.55);
    * the repulsive protein-protein polynomial among types 1-3 for every type
      pair (1-1, 1-2, 1-3, 2-2, 2-3, 3-3 within 2.6);
    * WCA Lennard-Jones: lipid-actin and lipid-spectrin (an unbonded actin and
      spectrin bead 1.0 below a lipid), band-3 (types 1, 2)-actin and
      type 1-spectrin at sigma 3.4, type 2-spectrin at sigma 1 (the bead
      tethered 1.1 below a band-3).
    Bonds: glycophorin-actin (0), band-3-spectrin (1), actin-spectrin (2),
    spectrin-spectrin (3), near their rest lengths.  Ids: proteins 1..np,
    lipids np+1.. (the reference's tag rule)."""
    rng = np.random.default_rng(seed)
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    lx = np.stack([(spacing * (ii + 0.5 * (jj % 2))).ravel(), (spacing * (np.sqrt(3) / 2) * jj).ravel(),
                   np.zeros(ii.size)], 1)
    lx += rng.normal(0, jitter * spacing, lx.shape)
    ext = lx.max(0)
    # clusters on a coarse grid inside the sheet
    cx = np.arange(7.0, ext[0] - 6.0, 13.0)
    cy = np.arange(7.0, ext[1] - 6.0, 13.0)
    P, T, bonds = [], [], []

    def add(p, t):
        P.append(np.asarray(p, np.float64))
        T.append(t)
        return len(P) - 1

    for a in cx:
        for b in cy:
            c = np.array([a, b, 0.0]) + np.r_[rng.normal(0, 0.3, 2), 0.0]
            z = np.array([0.0, 0.0, 1.0])
            g = add(c, 3)                                   # glycophorin
            act = add(c - 2.25 * z, 4)                      # its actin
            bonds.append((0, g, act))
            b2 = add(c + [2.35, 0.0, 0.0], 2)               # band-3: 2-3 at 2.35, 2-actin at 3.25
            add(c + [2.35 + 2.35, 0.4, 0.0], 2)             # 2-2 at 2.38
            b1 = add(c + [-1.2, 2.0, 0.0], 1)               # mobile band-3: 1-3 at 2.33, 1-actin at 3.24
            add(c + [-1.2 - 2.4, 2.3, 0.0], 1)              # 1-1 at 2.42
            add(c + [1.15, 2.55, 0.0], 3)                   # 1-3 / 2-3 near 2.6
            add(c + [4.7 + 0.5, 0.4 + 2.3, 0.0], 1)         # 1-2 at 2.35 (with the second band-3)
            g2 = add(c + [0.3, -1.9, 0.0], 3)               # 3-3 at 1.92
            a2 = add(c + [0.3, -1.9, -2.25], 4)
            bonds.append((0, g2, a2))
            t = add(c + [2.35, 0.0, -1.1], 5)               # tether bead: 2-5 WCA at 1.1
            bonds.append((1, b2, t))
            bonds.append((2, act, t))
            prev = act
            u = np.array([-0.59, -0.81, 0.0])
            for k in range(1, 4):                            # spectrin chain from the actin, away from band-3
                s = add(c + 2.22 * k * u - 2.25 * z, 5)
                bonds.append((2 if k == 1 else 3, prev, s))
                prev = s
            add(c + [-1.2, 2.0, -3.4], 5)                    # 1-5 WCA at sigma 3.4 (r = 3.4)
            _ = b1
    prot = np.array(P)
    ptype = np.array(T, np.int32)
    mem = ptype <= 3
    # remove lipids within 1.55 of a membrane protein
    d = np.linalg.norm(lx[:, None, :] - prot[None, mem, :], axis=2)
    lx = lx[d.min(1) > 1.55]
    # unbonded actin and spectrin beads 1.0 below chosen lipids (lipid-actin / lipid-spectrin WCA),
    # away from the band-3 (so that their sigma-3.4 terms stay outside the core)
    far = np.linalg.norm(lx[:, None, :2] - prot[None, ptype <= 2, :2], axis=2).min(1) > 4.5
    cand = np.where(far & (lx[:, 0] > 2) & (lx[:, 1] > 2) & (lx[:, 0] < ext[0] - 2) & (lx[:, 1] < ext[1] - 2))[0]
    pick = rng.choice(cand, 2 * len(cx) * len(cy), replace=False)
    extra_x, extra_t = [], []
    for k, li in enumerate(pick):
        extra_x.append(lx[li] - [0, 0, 0.97 + 0.05 * rng.random()])
        extra_t.append(4 if k % 2 == 0 else 5)
    prot = np.vstack([prot, np.array(extra_x)])
    ptype = np.concatenate([ptype, np.array(extra_t, np.int32)])
    prot += rng.normal(0, 0.01, prot.shape)
    x = np.vstack([lx, prot])
    types = np.concatenate([np.zeros(len(lx), np.int32), ptype])
    nl = len(lx)
    # the contact beads stay put (their distance to the lipid above sets the WCA contact)
    x = _clear_cutoffs(x, types, rng)
    nrm = np.tile([0.0, 0.0, 1.0], (len(x), 1)) + rng.normal(0, tilt, (len(x), 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    n = len(x)
    ids = np.zeros(n, np.uint32)
    ids[nl:] = np.arange(1, n - nl + 1)
    ids[:nl] = np.arange(n - nl + 1, n + 1)
    b = np.array(bonds, np.int64).reshape(-1, 3)
    return dict(x=x, n=nrm, v=rng.normal(0, 0.3, (n, 3)), o=rng.normal(0, 0.1, (n, 3)), type=types, id=ids,
                bonds=(b[:, 0].astype(np.int32), b[:, 1] + nl, b[:, 2] + nl))

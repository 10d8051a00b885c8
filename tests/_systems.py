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

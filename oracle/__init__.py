"""ctypes front end for the checker libraries.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this package.  The product (openrbc_amd) never does.

* ``Oracle``: the CPU restatement, oracle/_build/liborbc_oracle.so (FP64).
* ``Reference``: the OpenRBC reference compiled in place,
  oracle/_ref/libref_openrbc.so (FP32, fast-math, as shipped); present only
  where it was built.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liborbc_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref_openrbc.so")
REF_OPENRBC = os.path.join(HERE, "_ref", "openrbc")

_d = C.POINTER(C.c_double)
_f = C.POINTER(C.c_float)
_i32 = C.POINTER(C.c_int32)
_u32 = C.POINTER(C.c_uint32)
_i64 = C.POINTER(C.c_int64)
_u8 = C.POINTER(C.c_uint8)


def build(ref: bool = True) -> None:
    """Compile the checker libraries (make -C oracle)."""
    target = "all" if ref else os.path.join(HERE, "_build", "liborbc_oracle.so")
    subprocess.run(["make", "-s", "-C", HERE, target], check=True)


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class _FF(C.Structure):
    _fields_ = [(k, C.c_double * 36) for k in ("cut_poly", "cut2_poly", "rep", "att", "alpha")] + \
               [("aniso", C.c_int * 36)] + \
               [(k, C.c_double * 36) for k in ("cut2_lj", "lj1", "lj2")] + \
               [("mass", C.c_double * 6), ("radius", C.c_double * 6), ("bond_r0", C.c_double * 4),
                ("bond_K", C.c_double * 4)]


class Oracle:
    """FP64 CPU restatement (oracle/orbc_oracle.c)."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            build(ref=False)
        L = C.CDLL(path)
        self.L = L
        L.orc_ff_get.restype = C.POINTER(_FF)
        L.orc_ff_get.argtypes = []
        L.orc_pair.restype = C.c_int
        L.orc_pair.argtypes = [C.c_int, C.c_int, _d, _d, _d, _d, _d, _d, _d]
        L.orc_pairs_brute_fp64.restype = C.c_int64
        L.orc_pairs_brute_fp64.argtypes = [C.c_int64, _d, _i32, C.c_int64, _i64]
        L.orc_pairs_cell_fp64.restype = C.c_int64
        L.orc_pairs_cell_fp64.argtypes = [C.c_int64, _d, _i32, C.c_int64, _i64]
        L.orc_pairs_brute_mixed.restype = C.c_int64
        L.orc_pairs_brute_mixed.argtypes = [C.c_int64, _d, _d, _i32, C.c_double, C.c_int64, _i64]
        L.orc_forces_pairs.restype = None
        L.orc_forces_pairs.argtypes = [C.c_int64, _d, _d, _i32, C.c_int64, _i64, _d, _d, _d, _d]
        L.orc_bonded.restype = None
        L.orc_bonded.argtypes = [C.c_int64, _i32, _i64, _i64, _d, _d, _d, _d]
        L.orc_philox4x32_10.restype = None
        L.orc_philox4x32_10.argtypes = [_u32, _u32, _u32]
        L.orc_langevin_noise.restype = None
        L.orc_langevin_noise.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, _d]
        L.orc_langevin.restype = None
        L.orc_langevin.argtypes = [C.c_int64, _d, _d, _d, _d, _d, _d, _i32, _u32,
                                   C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_uint64]
        L.orc_nh_first.restype = C.c_double
        L.orc_nh_first.argtypes = [C.c_int64, _d, _d, _d, _d, _d, _d, _i32, C.c_double, C.c_double, _d, _d]
        L.orc_nh_second.restype = C.c_double
        L.orc_nh_second.argtypes = [C.c_int64, _d, _d, _d, _d, _d, _d, _i32, C.c_double, C.c_double]
        L.orc_nh_zeta.restype = C.c_double
        L.orc_nh_zeta.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int64, C.c_double]
        L.orc_temperature.restype = C.c_double
        L.orc_temperature.argtypes = [C.c_int64, _d, _i32]
        L.orc_minimize_move.restype = None
        L.orc_minimize_move.argtypes = [C.c_int64, _d, _d, _d, _d, _i32, C.c_double, C.c_double, C.c_double]
        L.orc_morton_ref.restype = C.c_uint32
        L.orc_morton_ref.argtypes = [C.c_float, C.c_float, C.c_float, C.c_double]
        L.orc_partition.restype = None
        L.orc_partition.argtypes = [C.c_int64, _d, C.c_int64, _f, _i32]
        L.orc_centroids.restype = None
        L.orc_centroids.argtypes = [C.c_int64, _d, _i32, C.c_int64, _f]
        L.orc_stencil.restype = C.c_int64
        L.orc_stencil.argtypes = [C.c_int64, _f, C.c_int64, C.c_float, C.c_int64, _i64]
        L.orc_pairs_voronoi_fp64.restype = C.c_int64
        L.orc_pairs_voronoi_fp64.argtypes = [C.c_int64, _d, _i32, _i32, C.c_int64, _f, C.c_int64, _i64]
        L.orc_stray_delete.restype = C.c_int64
        L.orc_stray_delete.argtypes = [C.c_int64, _d, _i32, _i32, C.c_int64, _f, C.c_double, _u8]
        L.orc_constrain_volume.restype = C.c_double
        L.orc_constrain_volume.argtypes = [C.c_int64, _d, _i32, _i32, C.c_int64, _f, C.c_double, C.c_double, _d]
        L.orc_set_box.restype = None
        L.orc_set_box.argtypes = [_d, _d, _i32]
        L.orc_wrap.restype = None
        L.orc_wrap.argtypes = [C.c_int64, _d]
        L.orc_pairs_brute_mixed_anchor.restype = C.c_int64
        L.orc_pairs_brute_mixed_anchor.argtypes = [C.c_int64, _d, _d, _i32, _d, _d, _d, _i32, C.c_int64, _i64]
        self.set_box(None)

    def set_box(self, box):
        """box = (lo[3], hi[3], periodic[3]) or None (no periodic axis).  All
        FP64 pair/force/bond functions take minimum images on periodic axes."""
        if box is None:
            lo = np.zeros(3); L = np.zeros(3); per = np.zeros(3, np.int32)
        else:
            lo = np.asarray(box[0], np.float64); L = np.asarray(box[1], np.float64) - lo
            per = np.asarray(box[2], np.int32)
        self.L.orc_set_box(_p(_c(lo, np.float64), _d), _p(_c(L, np.float64), _d), _p(_c(per, np.int32), _i32))

    def wrap(self, x):
        x = np.array(x, np.float64, order="C", copy=True)
        self.L.orc_wrap(len(x), _p(x, _d))
        return x

    def pairs_mixed_anchor(self, x, type_, obs, xa=None):
        """Mixed pair set with the device's per-axis anchor grid (observables of a Simulation)."""
        x = _c(x, np.float64); t = _c(type_, np.int32)
        xa = x if xa is None else _c(xa, np.float64)
        lo = np.asarray(obs["anchor_lo"], np.float64); h = np.asarray(obs["anchor_spacing"], np.float64)
        A0 = np.asarray(obs["anchor_origin"], np.float64); bits = np.asarray(obs["anchor_bits"], np.int32)
        return self._pairs(self.L.orc_pairs_brute_mixed_anchor, len(t), _p(x, _d), _p(xa, _d), _p(t, _i32),
                           _p(lo, _d), _p(h, _d), _p(A0, _d), _p(bits, _i32))

    def ff(self) -> dict:
        """Force-field tables, 6x6 arrays indexed [owner type, neighbour type]."""
        f = self.L.orc_ff_get().contents
        out = {}
        for name, _ in _FF._fields_:
            a = np.array(getattr(f, name)[:])
            out[name] = a.reshape(6, 6) if a.size == 36 else a
        return out

    # --- pair sets ---------------------------------------------------------
    def _pairs(self, fn, *args):
        cnt = fn(*args, 0, None)
        out = np.zeros((max(cnt, 1), 2), dtype=np.int64)
        fn(*args, cnt, _p(out, _i64))
        return out[:cnt]

    def pairs_brute(self, x, type_):
        x = _c(x, np.float64); t = _c(type_, np.int32)
        return self._pairs(self.L.orc_pairs_brute_fp64, len(t), _p(x, _d), _p(t, _i32))

    def pairs_cell(self, x, type_):
        x = _c(x, np.float64); t = _c(type_, np.int32)
        return self._pairs(self.L.orc_pairs_cell_fp64, len(t), _p(x, _d), _p(t, _i32))

    def pairs_mixed(self, x, type_, h, xa=None):
        x = _c(x, np.float64); t = _c(type_, np.int32)
        xa = x if xa is None else _c(xa, np.float64)
        return self._pairs(self.L.orc_pairs_brute_mixed, len(t), _p(x, _d), _p(xa, _d), _p(t, _i32), float(h))

    def pairs_voronoi(self, x, type_, aff, centroids):
        x = _c(x, np.float64); t = _c(type_, np.int32); a = _c(aff, np.int32)
        c = _c(centroids, np.float32)
        return self._pairs(self.L.orc_pairs_voronoi_fp64, len(t), _p(x, _d), _p(t, _i32), _p(a, _i32),
                           len(c), _p(c, _f))

    def pair(self, ti, tj, dx, ni, nj):
        """One directed pair: returns (hit, f_i, director force t_i, u, w)."""
        dx = _c(dx, np.float64); ni = _c(ni, np.float64); nj = _c(nj, np.float64)
        fi = np.zeros(3); tq = np.zeros(3); u = np.zeros(1); w = np.zeros(1)
        hit = self.L.orc_pair(int(ti), int(tj), _p(dx, _d), _p(ni, _d), _p(nj, _d), _p(fi, _d), _p(tq, _d),
                              _p(u, _d), _p(w, _d))
        return hit, fi, tq, u[0], w[0]

    # --- forces --------------------------------------------------------------
    def forces(self, x, nrm, type_, pairs=None, bonds=None):
        """Pair (+ bond) forces.  bonds = (btype, bi, bj) by index.
        Returns f, t, e (per-particle energy), w (per-particle virial)."""
        x = _c(x, np.float64); nrm = _c(nrm, np.float64); t = _c(type_, np.int32)
        n = len(t)
        if pairs is None:
            pairs = self.pairs_cell(x, t)
        pairs = _c(pairs, np.int64)
        f = np.zeros((n, 3)); tq = np.zeros((n, 3)); e = np.zeros(n); w = np.zeros(n)
        self.L.orc_forces_pairs(n, _p(x, _d), _p(nrm, _d), _p(t, _i32), len(pairs), _p(pairs, _i64),
                                _p(f, _d), _p(tq, _d), _p(e, _d), _p(w, _d))
        if bonds is not None:
            bt, bi, bj = (_c(bonds[0], np.int32), _c(bonds[1], np.int64), _c(bonds[2], np.int64))
            self.L.orc_bonded(len(bt), _p(bt, _i32), _p(bi, _i64), _p(bj, _i64), _p(x, _d), _p(f, _d),
                              _p(e, _d), _p(w, _d))
        return f, tq, e, w

    def bonded(self, x, bonds):
        x = _c(x, np.float64)
        n = len(x)
        bt, bi, bj = (_c(bonds[0], np.int32), _c(bonds[1], np.int64), _c(bonds[2], np.int64))
        f = np.zeros((n, 3)); e = np.zeros(n); w = np.zeros(n)
        self.L.orc_bonded(len(bt), _p(bt, _i32), _p(bi, _i64), _p(bj, _i64), _p(x, _d), _p(f, _d),
                          _p(e, _d), _p(w, _d))
        return f, e, w

    # --- integrators ---------------------------------------------------------
    def philox(self, ctr, key):
        c = _c(ctr, np.uint32); k = _c(key, np.uint32); o = np.zeros(4, np.uint32)
        self.L.orc_philox4x32_10(_p(c, _u32), _p(k, _u32), _p(o, _u32))
        return o

    def noise(self, seed, pid, step):
        r = np.zeros(3)
        self.L.orc_langevin_noise(int(seed), int(pid), int(step), _p(r, _d))
        return r

    def langevin(self, s, kT, eta, dt, seed, step):
        """s: dict with x v n o f t type id (modified in place)."""
        n = len(s["type"])
        self.L.orc_langevin(n, *[_p(s[k], _d) for k in ("x", "v", "n", "o", "f", "t")],
                            _p(s["type"], _i32), _p(s["id"], _u32), kT, eta, dt, int(seed), int(step))

    def nh_first(self, s, dt, zeta, box_lo=(-1000.0,) * 3, box_hi=(1000.0,) * 3):
        n = len(s["type"])
        lo = np.array(box_lo, dtype=np.float64); hi = np.array(box_hi, dtype=np.float64)
        return self.L.orc_nh_first(n, *[_p(s[k], _d) for k in ("x", "v", "n", "o", "f", "t")],
                                   _p(s["type"], _i32), dt, zeta, _p(lo, _d), _p(hi, _d))

    def nh_second(self, s, dt, zeta):
        n = len(s["type"])
        return self.L.orc_nh_second(n, *[_p(s[k], _d) for k in ("x", "v", "n", "o", "f", "t")],
                                    _p(s["type"], _i32), dt, zeta)

    def nh_zeta(self, zeta, dt, Q, ke, n, kT):
        return self.L.orc_nh_zeta(zeta, dt, Q, ke, n, kT)

    def temperature(self, v, type_):
        v = _c(v, np.float64); t = _c(type_, np.int32)
        return self.L.orc_temperature(len(t), _p(v, _d), _p(t, _i32))

    def minimize_move(self, s, dr_opt=5e-2, dn_opt=5e-2, dt=1e-2):
        n = len(s["type"])
        self.L.orc_minimize_move(n, _p(s["x"], _d), _p(s["n"], _d), _p(s["f"], _d), _p(s["t"], _d),
                                 _p(s["type"], _i32), dr_opt, dn_opt, dt)

    # --- Voronoi -------------------------------------------------------------
    def morton_ref(self, pts, bsize=2000.0):
        pts = _c(pts, np.float32)
        return np.array([self.L.orc_morton_ref(float(p[0]), float(p[1]), float(p[2]), bsize) for p in pts],
                        dtype=np.uint32)

    def partition(self, x, centroids):
        x = _c(x, np.float64); c = _c(centroids, np.float32)
        aff = np.zeros(len(x), np.int32)
        self.L.orc_partition(len(x), _p(x, _d), len(c), _p(c, _f), _p(aff, _i32))
        return aff

    def centroids(self, x, aff, ncell, init=None):
        x = _c(x, np.float64); a = _c(aff, np.int32)
        c = np.zeros((ncell, 3), np.float32) if init is None else _c(init, np.float32).copy()
        self.L.orc_centroids(len(x), _p(x, _d), _p(a, _i32), ncell, _p(c, _f))
        return c

    def stencil(self, centroids, cell, rmax):
        c = _c(centroids, np.float32)
        out = np.zeros(len(c), np.int64)
        k = self.L.orc_stencil(len(c), _p(c, _f), int(cell), float(rmax), len(c), _p(out, _i64))
        return out[:k]

    def stray_delete(self, x, type_, aff, centroids, tol):
        """Boolean mask of the lipids delete_lipid removes (src/cleanup.h:29-60)."""
        x = _c(x, np.float64); t = _c(type_, np.int32); a = _c(aff, np.int32); c = _c(centroids, np.float32)
        d = np.zeros(len(t), np.uint8)
        self.L.orc_stray_delete(len(t), _p(x, _d), _p(t, _i32), _p(a, _i32), len(c), _p(c, _f), float(tol),
                                _p(d, _u8))
        return d.astype(bool)

    def constrain_volume(self, nrm, type_, aff, centroids, target, strength, f):
        nrm = _c(nrm, np.float64); t = _c(type_, np.int32); a = _c(aff, np.int32)
        c = _c(centroids, np.float32)
        return self.L.orc_constrain_volume(len(t), _p(nrm, _d), _p(t, _i32), _p(a, _i32), len(c), _p(c, _f),
                                           target, strength, _p(f, _d))


class Reference:
    """The OpenRBC reference's own functions (FP32), oracle/ref_harness.cpp."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = C.CDLL(path)
        self.L = L
        L.ref_ff_tables.argtypes = [_f]
        L.ref_forces.restype = C.c_int
        L.ref_forces.argtypes = [C.c_int, _f, _f, C.c_int, _f, _f, _i32, _i32, C.c_int, _i32, _i32, _i32,
                                 C.c_int, C.c_int, _f, _f, _f, _f]
        L.ref_langevin.restype = None
        L.ref_langevin.argtypes = [C.c_int] + [_f] * 6 + [C.c_int] + [_f] * 6 + [_i32, C.c_double, C.c_double,
                                                                              C.c_double, C.c_int]
        L.ref_nh.restype = C.c_double
        L.ref_nh.argtypes = [C.c_int, C.c_int] + [_f] * 6 + [C.c_int] + [_f] * 6 + [_i32, C.c_double, C.c_double,
                                                                                   C.c_double, C.c_double]
        L.ref_temperature.restype = C.c_double
        L.ref_temperature.argtypes = [C.c_int, _f, C.c_int, _f, _i32]
        L.ref_morton.argtypes = [C.c_int, _f, _u32]
        L.ref_reorder_morton.argtypes = [C.c_int, _f]
        L.ref_kdtree_nearest.argtypes = [C.c_int, _f, C.c_int, _f, _i32]
        L.ref_kdtree_within.argtypes = [C.c_int, _f, C.c_int, _f, C.c_float, C.c_int, _i32, _i32]
        L.ref_partition.argtypes = [C.c_int, _f, C.c_int, _f, _i32]
        L.ref_stray.argtypes = [C.c_int, _f, C.c_int, _f, C.c_double, _i32, _u8]
        L.ref_save_frame.argtypes = [C.c_char_p, C.c_int, _f, _f, _f, _i32, C.c_int, C.c_int, _f, _f, _f, _i32,
                                     _i32, _i32, C.c_int, C.c_int, C.c_int]
        L.ref_save_topology.argtypes = [C.c_char_p, C.c_int, _f, C.c_int, C.c_int, _f, _i32, _i32, C.c_int,
                                        _i32, _i32, _i32]
        L.ref_forces_cells.restype = C.c_int
        L.ref_forces_cells.argtypes = [C.c_int, _f, _f, C.c_int, _f, _f, _i32, _i32, C.c_int, _i32, _i32, _i32,
                                       C.c_int, _f, _i32, _i32, _f, _f, _f, _f]
        L.ref_init_rbc.restype = C.c_double
        L.ref_init_rbc.argtypes = [C.c_char_p, C.c_int, C.c_double, C.POINTER(C.c_long), _f, _f, _f, _f, _i32,
                                   _i32, _i32, _i32, _i32]
        L.ref_langevin_lipids.restype = None
        L.ref_langevin_lipids.argtypes = [C.c_int, _f, _f, _f, C.c_double, C.c_double, C.c_double, C.c_int,
                                          C.c_int, C.c_int, _f, _f, _f]
        L.ref_run.restype = C.c_double
        L.ref_run.argtypes = [C.c_int, _f, _f, C.c_int, _f, _f, _i32, _i32, C.c_int, _i32, _i32, _i32,
                              C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                              C.POINTER(C.c_long)]

    def ff_tables(self):
        out = np.zeros(6 * 7 + 36 * 6 + 8, np.float32)
        self.L.ref_ff_tables(_p(out, _f))
        k = 0
        res = {}
        for name, sz in (("mass", 6), ("radius", 6), ("cutsqlp", 6), ("replp", 6), ("attlp", 6), ("alphalp", 6),
                         ("cutlp", 6), ("cutsqpp", 36), ("reppp", 36), ("cutpp", 36), ("lj_cutsq", 36),
                         ("lj_lj1", 36), ("lj_lj2", 36), ("r0", 4), ("K", 4)):
            res[name] = out[k:k + sz].copy()
            k += sz
        return res

    def forces(self, lx, ln, px, pn, ptype, ptag, bonds_tag=None, do_pair=True, do_bond=True):
        """bonds_tag = (btype, tag_i, tag_j).  Returns lf, lt, pf, pt (FP32)."""
        lx = _c(lx, np.float32); ln = _c(ln, np.float32); px = _c(px, np.float32); pn = _c(pn, np.float32)
        pty = _c(ptype, np.int32); ptg = _c(ptag, np.int32)
        nl, np_ = len(lx), len(px)
        if bonds_tag is None:
            bt = bi = bj = np.zeros(0, np.int32)
        else:
            bt, bi, bj = (_c(b, np.int32) for b in bonds_tag)
        lf = np.zeros((nl, 3), np.float32); lt = np.zeros((nl, 3), np.float32)
        pf = np.zeros((np_, 3), np.float32); pt = np.zeros((np_, 3), np.float32)
        self.L.ref_forces(nl, _p(lx, _f), _p(ln, _f), np_, _p(px, _f), _p(pn, _f), _p(pty, _i32), _p(ptg, _i32),
                          len(bt), _p(bt, _i32), _p(bi, _i32), _p(bj, _i32), int(do_pair), int(do_bond),
                          _p(lf, _f), _p(lt, _f), _p(pf, _f), _p(pt, _f))
        return lf, lt, pf, pt

    def forces_cells(self, lx, ln, px, pn, ptype, ptag, bonds_tag, centroids, lcell_start, pcell_start):
        """compute_pairwise_fused over Voronoi cells (containers sorted by cell,
        cell_start offsets) + compute_bonded.  Returns lf, lt, pf, pt (FP32)."""
        lx = _c(lx, np.float32); ln = _c(ln, np.float32); px = _c(px, np.float32); pn = _c(pn, np.float32)
        pty = _c(ptype, np.int32); ptg = _c(ptag, np.int32)
        bt, bi, bj = (_c(b, np.int32) for b in bonds_tag)
        c = _c(centroids, np.float32); ls = _c(lcell_start, np.int32); ps = _c(pcell_start, np.int32)
        nl, np_ = len(lx), len(px)
        lf = np.zeros((nl, 3), np.float32); lt = np.zeros((nl, 3), np.float32)
        pf = np.zeros((np_, 3), np.float32); pt = np.zeros((np_, 3), np.float32)
        self.L.ref_forces_cells(nl, _p(lx, _f), _p(ln, _f), np_, _p(px, _f), _p(pn, _f), _p(pty, _i32),
                                _p(ptg, _i32), len(bt), _p(bt, _i32), _p(bi, _i32), _p(bj, _i32), len(c), _p(c, _f),
                                _p(ls, _i32), _p(ps, _i32), _p(lf, _f), _p(lt, _f), _p(pf, _f), _p(pt, _f))
        return lf, lt, pf, pt

    def init_rbc(self, prefix, seed, rho=1.05):
        """init_rbc on the mesh files prefix.{vert,bond,face}.txt (FP32 outputs)."""
        cnt = (C.c_long * 3)()
        self.L.ref_init_rbc(prefix.encode(), int(seed), float(rho), cnt, *([None] * 9))
        nl, np_, nb = cnt[0], cnt[1], cnt[2]
        out = dict(lx=np.zeros((nl, 3), np.float32), ln=np.zeros((nl, 3), np.float32),
                   px=np.zeros((np_, 3), np.float32), pn=np.zeros((np_, 3), np.float32),
                   ptype=np.zeros(np_, np.int32), ptag=np.zeros(np_, np.int32), bt=np.zeros(nb, np.int32),
                   bi=np.zeros(nb, np.int32), bj=np.zeros(nb, np.int32))
        area = self.L.ref_init_rbc(prefix.encode(), int(seed), float(rho), cnt,
                                   *[_p(out[k], _f) for k in ("lx", "ln", "px", "pn")],
                                   *[_p(out[k], _i32) for k in ("ptype", "ptag", "bt", "bi", "bj")])
        out["area"] = area
        return out

    def langevin_lipids(self, lx, ln, lv, kT, eta, dt, seed, nsteps, every):
        """verlet_langevin run of a lipid-only system with exact pairs; snapshots
        (x, n, v) every `every` steps, shape (nsteps // every, n, 3) each."""
        lx = _c(lx, np.float32); ln = _c(ln, np.float32); lv = _c(lv, np.float32)
        m = nsteps // every
        xs = np.zeros((m, len(lx), 3), np.float32); ns = np.zeros_like(xs); vs = np.zeros_like(xs)
        self.L.ref_langevin_lipids(len(lx), _p(lx, _f), _p(ln, _f), _p(lv, _f), kT, eta, dt, int(seed),
                                   int(nsteps), int(every), _p(xs, _f), _p(ns, _f), _p(vs, _f))
        return xs, ns, vs

    def _state_args(self, L, P, ptype):
        return ([len(L["x"])] + [_p(L[k], _f) for k in ("x", "v", "n", "o", "f", "t")] + [len(P["x"])] +
                [_p(P[k], _f) for k in ("x", "v", "n", "o", "f", "t")] + [_p(ptype, _i32)])

    def langevin(self, L, P, ptype, kT, eta, dt, seed=1):
        """L, P: dicts of float32 (n,3) arrays x v n o f t (modified in place)."""
        ptype = _c(ptype, np.int32)
        self.L.ref_langevin(*self._state_args(L, P, ptype), kT, eta, dt, seed)

    def nh(self, which, L, P, ptype, kT, dt, zeta, Q):
        ptype = _c(ptype, np.int32)
        return self.L.ref_nh(which, *self._state_args(L, P, ptype), kT, dt, zeta, Q)

    def temperature(self, lv, pv, ptype):
        lv = _c(lv, np.float32); pv = _c(pv, np.float32); pt = _c(ptype, np.int32)
        return self.L.ref_temperature(len(lv), _p(lv, _f), len(pv), _p(pv, _f), _p(pt, _i32))

    def morton(self, pts):
        pts = _c(pts, np.float32)
        keys = np.zeros(len(pts), np.uint32)
        self.L.ref_morton(len(pts), _p(pts, _f), _p(keys, _u32))
        return keys

    def reorder_morton(self, pts):
        pts = _c(pts, np.float32).copy()
        self.L.ref_reorder_morton(len(pts), _p(pts, _f))
        return pts

    def kdtree_nearest(self, pts, q):
        pts = _c(pts, np.float32); q = _c(q, np.float32)
        idx = np.zeros(len(q), np.int32)
        self.L.ref_kdtree_nearest(len(pts), _p(pts, _f), len(q), _p(q, _f), _p(idx, _i32))
        return idx

    def kdtree_within(self, pts, q, rmax, cap=512):
        pts = _c(pts, np.float32); q = _c(q, np.float32)
        cnt = np.zeros(len(q), np.int32); out = np.zeros((len(q), cap), np.int32)
        self.L.ref_kdtree_within(len(pts), _p(pts, _f), len(q), _p(q, _f), float(rmax), cap, _p(cnt, _i32),
                                 _p(out, _i32))
        return [np.sort(out[k, :cnt[k]]) for k in range(len(q))]

    def partition(self, x, centroids):
        x = _c(x, np.float32); c = _c(centroids, np.float32)
        aff = np.zeros(len(x), np.int32)
        self.L.ref_partition(len(x), _p(x, _f), len(c), _p(c, _f), _p(aff, _i32))
        return aff

    def save_frame(self, path, lx, lv, ln, laff, lbase, px, pv, pn, paff, ptype, ptag, nstep, dump_field,
                   append=False):
        """save_frame of containers built from the arrays (FP32)."""
        a = [_c(v, np.float32) for v in (lx, lv, ln, px, pv, pn)]
        la, pa = _c(laff, np.int32), _c(paff, np.int32)
        pt, pg = _c(ptype, np.int32), _c(ptag, np.int32)
        self.L.ref_save_frame(path.encode(), len(a[0]), _p(a[0], _f), _p(a[1], _f), _p(a[2], _f), _p(la, _i32),
                              int(lbase), len(a[3]), _p(a[3], _f), _p(a[4], _f), _p(a[5], _f), _p(pa, _i32),
                              _p(pt, _i32), _p(pg, _i32), int(nstep), int(dump_field), int(append))

    def save_topology(self, path, lx, lbase, px, ptype, ptag, bt, bi, bj):
        lx, px = _c(lx, np.float32), _c(px, np.float32)
        pt, pg = _c(ptype, np.int32), _c(ptag, np.int32)
        bt, bi, bj = _c(bt, np.int32), _c(bi, np.int32), _c(bj, np.int32)
        self.L.ref_save_topology(path.encode(), len(lx), _p(lx, _f), int(lbase), len(px), _p(px, _f),
                                 _p(pt, _i32), _p(pg, _i32), len(bt), _p(bt, _i32), _p(bi, _i32), _p(bj, _i32))

    def stray(self, x, centroids, tol):
        """delete_lipid on lipids x with fixed centroids: (cell of each lipid, survives)."""
        x = _c(x, np.float32); c = _c(centroids, np.float32)
        aff = np.zeros(len(x), np.int32); keep = np.zeros(len(x), np.uint8)
        self.L.ref_stray(len(x), _p(x, _f), len(c), _p(c, _f), float(tol), _p(aff, _i32), _p(keep, _u8))
        return aff, keep.astype(bool)

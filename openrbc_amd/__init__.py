"""openrbc_amd -- MI355X-native per-timestep MD loop of OpenRBC.

Python host layer over liborbc_hip.so (C ABI in include/orbc.h).  It mirrors
the reference's driver-level interface (src/openrbc.cpp): containers of lipids
and proteins, ``compute_pairwise_fused``, ``compute_bonded``,
``integrate(kernel)``, ``constrain_volume``, ``compute_temperature`` and the
Voronoi cells, with the same argument meaning; errors raise ``OrbcError``
instead of exiting the process (the reference's rt_assert calls exit(0),
src/util_misc.h:54-59).

The GPU library is required: nothing here falls back to a CPU path.  If
liborbc_hip.so is missing, ``load_library`` raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORBC_LIBRARY selects another build of the library (e.g. a tuning variant)
LIB_PATH = os.environ.get("ORBC_LIBRARY") or os.path.join(HERE, "_lib", "liborbc_hip.so")

# enums of include/orbc.h
LIPID, PROTEIN = 0, 1
FP64, MIXED = 0, 1
NEIGHBOR_EXACT, NEIGHBOR_VORONOI = 0, 1
CLEAR_FORCE, VERLET_LANGEVIN, NH_INITIAL_BOUNCE_CLEAR, NH_POST_TORQUE_FINAL, POST_TORQUE, BOUNCE_BACK = range(6)
RUN_LANGEVIN, RUN_NOSE_HOOVER = 0, 1
F_ZERO_FIRST, F_ENERGY = 1, 2

_KERNELS = {
    "clear_force": CLEAR_FORCE,
    "verlet_langevin": VERLET_LANGEVIN,
    "verlet_initial_bounce_clearforce_update": NH_INITIAL_BOUNCE_CLEAR,
    "post_toque_final_update": NH_POST_TORQUE_FINAL,
    "post_torque": POST_TORQUE,
    "bounce_back": BOUNCE_BACK,
}


class OrbcError(RuntimeError):
    """A liborbc_hip call returned a non-zero orbc_status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"orbc status {code}: {msg}")
        self.code = code


class Config(C.Structure):
    """orbc_config; defaults of RTParameter (src/runtime_parameter.h:38-80)."""

    _fields_ = [
        ("precision", C.c_int), ("kBT", C.c_double), ("eta", C.c_double), ("dt", C.c_double),
        ("rho", C.c_double), ("voronoi_cell_size", C.c_int), ("stray_tolerance", C.c_double),
        ("freq_voronoi", C.c_int), ("freq_sort_ctrd", C.c_int), ("freq_sort_bond", C.c_int),
        ("freq_cleanup", C.c_int), ("seed", C.c_ulonglong), ("Q", C.c_double), ("zeta", C.c_double),
        ("box_lo", C.c_double * 3), ("box_hi", C.c_double * 3), ("periodic", C.c_int * 3),
        ("skin", C.c_double), ("rebuild_period", C.c_int), ("neighbor_mode", C.c_int),
        ("thermostat", C.c_int), ("volume_target", C.c_double), ("volume_strength", C.c_double),
        ("device", C.c_int), ("block_size", C.c_int), ("block_halo", C.c_int), ("graphs", C.c_int),
    ]


class Tuning(C.Structure):
    """orbc_tuning: the knobs orbc_autotune chooses."""

    _fields_ = [
        ("skin", C.c_double), ("block_halo", C.c_int), ("centroid_cell", C.c_double), ("batch", C.c_int),
        ("ms_per_step", C.c_double), ("ms_per_step_start", C.c_double), ("trials", C.c_int),
        ("halo_capacity", C.c_int),
    ]


class Observables(C.Structure):
    _fields_ = [
        ("kinetic", C.c_double), ("temperature", C.c_double), ("potential", C.c_double),
        ("virial", C.c_double), ("zeta", C.c_double), ("n", C.c_longlong), ("step", C.c_longlong),
        ("rebuilds", C.c_longlong), ("pairs", C.c_longlong), ("anchor_h", C.c_double), ("skin", C.c_double),
        ("anchor_lo", C.c_double * 3), ("anchor_spacing", C.c_double * 3), ("anchor_origin", C.c_double * 3),
        ("anchor_bits", C.c_int * 3), ("n_lipid", C.c_longlong), ("n_protein", C.c_longlong),
        ("lipids_deleted", C.c_longlong), ("halo_blocks", C.c_longlong), ("halo_overflow", C.c_longlong),
        ("graph_captures", C.c_longlong), ("graph_replays", C.c_longlong),
    ]


class _System(C.Structure):
    _fields_ = [
        ("n_lipid", C.c_longlong), ("n_protein", C.c_longlong), ("n_bond", C.c_longlong),
        ("lipid_x", C.POINTER(C.c_double)), ("lipid_n", C.POINTER(C.c_double)),
        ("lipid_id", C.POINTER(C.c_uint)),
        ("protein_x", C.POINTER(C.c_double)), ("protein_n", C.POINTER(C.c_double)),
        ("protein_type", C.POINTER(C.c_int)), ("protein_id", C.POINTER(C.c_uint)),
        ("bond_type", C.POINTER(C.c_int)), ("bond_i", C.POINTER(C.c_uint)), ("bond_j", C.POINTER(C.c_uint)),
        ("area", C.c_double), ("clutter_capped", C.c_longlong),
    ]


_d = C.POINTER(C.c_double)
_f = C.POINTER(C.c_float)
_i = C.POINTER(C.c_int)
_u = C.POINTER(C.c_uint)
_vp = C.c_void_p

# (name, restype, argtypes) of every C-ABI entry point (include/orbc.h)
SIGNATURES = [
    ("orbc_config_default", None, [C.POINTER(Config)]),
    ("orbc_create", C.c_int, [C.POINTER(Config), C.POINTER(_vp)]),
    ("orbc_destroy", C.c_int, [_vp]),
    ("orbc_last_error", C.c_char_p, [_vp]),
    ("orbc_get_config", C.c_int, [_vp, C.POINTER(Config)]),
    ("orbc_set_state", C.c_int, [_vp, C.c_int, C.c_longlong, _d, _d, _d, _d, _i, _u]),
    ("orbc_set_bonds", C.c_int, [_vp, C.c_longlong, _i, _u, _u]),
    ("orbc_set_groups", C.c_int, [_vp, C.c_int, C.c_longlong, _i, C.c_int, _d]),
    ("orbc_get_state", C.c_int, [_vp, C.c_int, _d, _d, _d, _d]),
    ("orbc_get_forces", C.c_int, [_vp, C.c_int, _d, _d, _d, _d]),
    ("orbc_set_forces", C.c_int, [_vp, C.c_int, _d, _d]),
    ("orbc_build_neighbors", C.c_int, [_vp]),
    ("orbc_compute_pairwise", C.c_int, [_vp, C.c_uint]),
    ("orbc_compute_bonded", C.c_int, [_vp, C.c_uint]),
    ("orbc_integrate", C.c_int, [_vp, C.c_int]),
    ("orbc_constrain_volume", C.c_int, [_vp, C.c_double, C.c_double, _d]),
    ("orbc_compute_temperature", C.c_int, [_vp, _d]),
    ("orbc_run", C.c_int, [_vp, C.c_longlong]),
    ("orbc_minimize", C.c_int, [_vp, C.c_longlong, C.c_double, C.c_double]),
    ("orbc_get_observables", C.c_int, [_vp, C.POINTER(Observables)]),
    ("orbc_get_pairs", C.c_int, [_vp, C.c_longlong, _u, _u, C.POINTER(C.c_longlong)]),
    ("orbc_voronoi_init", C.c_int, [_vp, C.c_longlong, C.c_int]),
    ("orbc_voronoi_update", C.c_int, [_vp]),
    ("orbc_voronoi_set_centroids", C.c_int, [_vp, C.c_longlong, _f]),
    ("orbc_voronoi_get", C.c_int, [_vp, _f, _i, _i]),
    ("orbc_voronoi_count", C.c_int, [_vp, C.POINTER(C.c_longlong)]),
    ("orbc_delete_stray_lipids", C.c_int, [_vp, C.c_double, C.POINTER(C.c_longlong)]),
    ("orbc_morton_keys", C.c_int, [_vp, C.c_longlong, _f, C.c_double, _u]),
    ("orbc_sort_pairs", C.c_int, [_vp, C.c_longlong, _u, _u, C.c_int]),
    ("orbc_sort_pairs_noalloc", C.c_int, [_vp, C.c_longlong, _vp, _vp, C.c_int]),
    ("orbc_autotune", C.c_int, [_vp, C.c_int, C.c_int, C.c_char_p, C.POINTER(Tuning)]),
    ("orbc_set_tuning", C.c_int, [_vp, C.POINTER(Tuning)]),
    ("orbc_get_tuning", C.c_int, [_vp, C.POINTER(Tuning)]),
    ("orbc_nearest", C.c_int, [_vp, C.c_longlong, _f, C.c_longlong, _f, _i]),
    ("orbc_within", C.c_int, [_vp, C.c_longlong, _f, C.c_longlong, _f, C.c_float, C.c_int, _i, _i]),
    ("orbc_synchronize", C.c_int, [_vp]),
    ("orbc_save_frame", C.c_int, [_vp, C.c_char_p, C.c_int, C.c_int]),
    ("orbc_read_frame", C.c_int, [C.c_char_p, C.c_longlong, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                  _i, _u, _i, _f, _f, _f, _i, _f]),
    ("orbc_save_topology", C.c_int, [_vp, C.c_char_p]),
    ("orbc_save_restart", C.c_int, [_vp, C.c_char_p]),
    ("orbc_load_restart", C.c_int, [_vp, C.c_char_p]),
    ("orbc_set_timing", C.c_int, [_vp, C.c_int]),
    ("orbc_get_timing", C.c_int, [_vp, _d, C.POINTER(C.c_longlong), _d, C.POINTER(C.c_longlong)]),
    ("orbc_init_rbc", C.c_int, [_d, C.c_longlong, _i, C.c_longlong, _i, C.c_longlong, C.c_double, C.c_ulonglong,
                                C.POINTER(_System)]),
    ("orbc_system_free", None, [C.POINTER(_System)]),
]

_LIB = None


# DumpField bits of the trajectory format (src/runtime_parameter.h:27-33)
POSITION, ROTATION, VORONOI, VELOCITY, FORCE = 1, 2, 4, 8, 16


def read_frame(path: str, index: int = 0) -> dict:
    """read_frame (src/trajectory.h:107-150): frame `index` of a trajectory
    file as arrays (tag, type, and the fields present: x, n, aff, v, f)."""
    L = load_library()
    step, natom, fields = C.c_longlong(), C.c_longlong(), C.c_int()
    p = os.fsencode(path)
    rc = L.orbc_read_frame(p, int(index), C.byref(step), C.byref(natom), C.byref(fields), None, None, None, None,
                           None, None, None)
    if rc != 0:
        raise OrbcError(rc, f"cannot read frame {index} of {path}")
    m = natom.value
    tag = np.zeros(m, np.uint32); typ = np.zeros(m, np.int32); aff = np.zeros(m, np.int32)
    vec = {k: np.zeros((m, 3), np.float32) for k in ("x", "v", "n", "f")}
    rc = L.orbc_read_frame(p, int(index), C.byref(step), C.byref(natom), C.byref(fields), _ptr(tag, _u),
                           _ptr(typ, _i), _ptr(vec["x"], _f), _ptr(vec["v"], _f), _ptr(vec["n"], _f),
                           _ptr(aff, _i), _ptr(vec["f"], _f))
    if rc != 0:
        raise OrbcError(rc, f"cannot read frame {index} of {path}")
    out = dict(step=step.value, natom=m, fields=fields.value, tag=tag, type=typ)
    for bit, key in ((POSITION, "x"), (VELOCITY, "v"), (ROTATION, "n"), (FORCE, "f")):
        if fields.value & bit:
            out[key] = vec[key]
    if fields.value & VORONOI:
        out["aff"] = aff
    return out


def load_library(path: str = LIB_PATH):
    """Load liborbc_hip.so (raises if it was not built: no CPU fallback)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise OrbcError(-1, f"{path} is missing; build it with `python -m openrbc_amd.build`")
    L = C.CDLL(path)
    for name, res, args in SIGNATURES:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def default_config(**kw) -> Config:
    L = load_library()
    cfg = Config()
    L.orbc_config_default(C.byref(cfg))
    for k, v in kw.items():
        if k in ("box_lo", "box_hi", "periodic"):
            arr = getattr(cfg, k)
            for d in range(3):
                arr[d] = v[d]
        else:
            setattr(cfg, k, v)
    return cfg


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


def _f64(a, n=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.reshape(-1, 3) if n is None else a


@dataclass
class Container:
    """A LipidContainer / ProteContainer (src/container.h:65-157): FP64 AoS
    x, v, n (director), o (angular velocity); protein type and tag."""
    x: np.ndarray
    n: np.ndarray
    id: np.ndarray
    v: np.ndarray | None = None
    o: np.ndarray | None = None
    type: np.ndarray | None = None

    def __len__(self):
        return len(self.x)


class Simulation:
    """One orbc_ctx: the device particle store and the per-timestep loop."""

    def __init__(self, config: Config | None = None, **kw):
        self.L = load_library()
        self.cfg = config if config is not None else default_config(**kw)
        h = _vp()
        rc = self.L.orbc_create(C.byref(self.cfg), C.byref(h))
        if rc != 0:
            raise OrbcError(rc, "orbc_create failed (no GPU, or invalid configuration)")
        self.h = h
        self.n = {LIPID: 0, PROTEIN: 0}

    # -- housekeeping -------------------------------------------------------
    def _check(self, rc):
        if rc != 0:
            msg = self.L.orbc_last_error(self.h)
            raise OrbcError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            self.L.orbc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- state --------------------------------------------------------------
    def set_state(self, species: int, x, n, id, v=None, o=None, type=None):
        x = _f64(x); n = _f64(n); v = _f64(v); o = _f64(o)
        ids = np.ascontiguousarray(id, dtype=np.uint32)
        ty = None if type is None else np.ascontiguousarray(type, dtype=np.int32)
        cnt = len(ids)
        self._check(self.L.orbc_set_state(self.h, species, cnt, _ptr(x, _d), _ptr(v, _d), _ptr(n, _d), _ptr(o, _d),
                                          _ptr(ty, _i), _ptr(ids, _u)))
        self.n[species] = cnt

    def set_lipids(self, c: Container):
        self.set_state(LIPID, c.x, c.n, c.id, c.v, c.o)

    def set_proteins(self, c: Container):
        self.set_state(PROTEIN, c.x, c.n, c.id, c.v, c.o, c.type)

    def set_bonds(self, btype, id_i, id_j):
        t = np.ascontiguousarray(btype, dtype=np.int32)
        a = np.ascontiguousarray(id_i, dtype=np.uint32)
        b = np.ascontiguousarray(id_j, dtype=np.uint32)
        self._check(self.L.orbc_set_bonds(self.h, len(t), _ptr(t, _i), _ptr(a, _u), _ptr(b, _u)))

    def set_groups(self, species: int, group, kBT):
        """Ensemble members: group[k] of each particle of `species`, member temperatures kBT."""
        g = np.ascontiguousarray(group, dtype=np.int32)
        k = np.ascontiguousarray(kBT, dtype=np.float64)
        self._check(self.L.orbc_set_groups(self.h, int(species), len(g), _ptr(g, _i), len(k), _ptr(k, _d)))

    def get_state(self, species: int):
        m = self.n[species]
        out = {k: np.zeros((m, 3)) for k in ("x", "v", "n", "o")}
        self._check(self.L.orbc_get_state(self.h, species, *[_ptr(out[k], _d) for k in ("x", "v", "n", "o")]))
        return out

    def get_forces(self, species: int, energy: bool = False):
        m = self.n[species]
        f = np.zeros((m, 3)); t = np.zeros((m, 3))
        e = np.zeros(m) if energy else None
        w = np.zeros(m) if energy else None
        self._check(self.L.orbc_get_forces(self.h, species, _ptr(f, _d), _ptr(t, _d), _ptr(e, _d), _ptr(w, _d)))
        return (f, t, e, w) if energy else (f, t)

    def set_forces(self, species: int, f, t):
        f = _f64(f); t = _f64(t)
        self._check(self.L.orbc_set_forces(self.h, species, _ptr(f, _d), _ptr(t, _d)))

    # -- the reference's per-step calls ----------------------------------------
    def build_neighbors(self):
        self._check(self.L.orbc_build_neighbors(self.h))

    def compute_pairwise(self, zero_first=False, energy=False):
        self._check(self.L.orbc_compute_pairwise(self.h, (F_ZERO_FIRST if zero_first else 0) |
                                                 (F_ENERGY if energy else 0)))

    def compute_bonded(self, zero_first=False, energy=False):
        self._check(self.L.orbc_compute_bonded(self.h, (F_ZERO_FIRST if zero_first else 0) |
                                               (F_ENERGY if energy else 0)))

    def integrate(self, kernel):
        k = _KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
        self._check(self.L.orbc_integrate(self.h, k))

    def constrain_volume(self, target: float, strength: float) -> float:
        vol = C.c_double()
        self._check(self.L.orbc_constrain_volume(self.h, target, strength, C.byref(vol)))
        return vol.value

    def compute_temperature(self) -> float:
        T = C.c_double()
        self._check(self.L.orbc_compute_temperature(self.h, C.byref(T)))
        return T.value

    def run(self, nsteps: int):
        try:
            self._check(self.L.orbc_run(self.h, int(nsteps)))
        finally:
            self._sync_counts()

    def _sync_counts(self):
        """Species counts after the device may have removed stray lipids."""
        o = Observables()
        if self.L.orbc_get_observables(self.h, C.byref(o)) == 0:
            self.n[LIPID], self.n[PROTEIN] = o.n_lipid, o.n_protein

    def delete_stray_lipids(self, tolerance: float = 2.5) -> int:
        """delete_lipid (src/cleanup.h:29-91) with the current Voronoi cells;
        returns the number of lipids removed."""
        nd = C.c_longlong()
        try:
            self._check(self.L.orbc_delete_stray_lipids(self.h, float(tolerance), C.byref(nd)))
        finally:
            self._sync_counts()
        return nd.value

    def minimize(self, nsteps: int, dr_opt: float = 5e-2, dn_opt: float = 5e-2):
        self._check(self.L.orbc_minimize(self.h, int(nsteps), dr_opt, dn_opt))

    def observables(self) -> dict:
        o = Observables()
        self._check(self.L.orbc_get_observables(self.h, C.byref(o)))
        out = {}
        for k, _ in Observables._fields_:
            v = getattr(o, k)
            out[k] = list(v) if isinstance(v, C.Array) else v
        return out

    def pairs(self) -> np.ndarray:
        """Interacting unordered pairs (id_i < id_j) of the current list, sorted."""
        cnt = C.c_longlong()
        self._check(self.L.orbc_get_pairs(self.h, 0, None, None, C.byref(cnt)))
        m = cnt.value
        a = np.zeros(max(m, 1), np.uint32); b = np.zeros(max(m, 1), np.uint32)
        self._check(self.L.orbc_get_pairs(self.h, m, _ptr(a, _u), _ptr(b, _u), C.byref(cnt)))
        p = np.stack([a[:m], b[:m]], 1).astype(np.int64)
        return p[np.lexsort((p[:, 1], p[:, 0]))]

    def autotune(self, trial_steps: int = 40, table: str | None = None, warm_steps: int = -1) -> dict:
        """orbc_autotune: measure skin, block halos, centroid-grid edge and batch
        length on this system over runs of warm_steps + trial_steps steps and
        apply the fastest (state restored)."""
        t = Tuning()
        self._check(self.L.orbc_autotune(self.h, int(warm_steps), int(trial_steps),
                                         table.encode() if table else None, C.byref(t)))
        return {k: getattr(t, k) for k, _ in Tuning._fields_}

    def set_tuning(self, skin: float, block_halo: int, centroid_cell: float, batch: int, halo_capacity: int = 0):
        t = Tuning(skin=skin, block_halo=block_halo, centroid_cell=centroid_cell, batch=batch,
                   halo_capacity=halo_capacity)
        self._check(self.L.orbc_set_tuning(self.h, C.byref(t)))

    def tuning(self) -> dict:
        t = Tuning()
        self._check(self.L.orbc_get_tuning(self.h, C.byref(t)))
        return {k: getattr(t, k) for k in ("skin", "block_halo", "centroid_cell", "batch", "halo_capacity")}

    def synchronize(self):
        self._check(self.L.orbc_synchronize(self.h))

    # -- trajectory / topology files (src/trajectory.h, src/topology.h) ------------
    def save_frame(self, path: str, dump_field: int = POSITION | ROTATION | VORONOI, append: bool = True):
        """save_frame: one frame of the reference's binary trajectory format."""
        self._check(self.L.orbc_save_frame(self.h, os.fsencode(path), int(dump_field), int(append)))

    def save_topology(self, path: str):
        """save_topology: the reference's LAMMPS data file."""
        self._check(self.L.orbc_save_topology(self.h, os.fsencode(path)))

    def save_restart(self, path: str):
        """The complete FP64 state, for orbc_load_restart."""
        self._check(self.L.orbc_save_restart(self.h, os.fsencode(path)))

    def load_restart(self, path: str):
        """Replace this context's configuration and state by a restart file."""
        try:
            self._check(self.L.orbc_load_restart(self.h, os.fsencode(path)))
        finally:
            self._sync_counts()
        cfg = Config()
        if self.L.orbc_get_config(self.h, C.byref(cfg)) == 0:
            self.cfg = cfg

    def set_timing(self, on: bool = True):
        self._check(self.L.orbc_set_timing(self.h, int(on)))

    def timing(self) -> dict:
        a, c = C.c_double(), C.c_double()
        b, d = C.c_longlong(), C.c_longlong()
        self._check(self.L.orbc_get_timing(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return dict(step_kernel_ms=a.value, step_launches=b.value, rebuild_ms=c.value, rebuilds=d.value)

    # -- Voronoi cells ----------------------------------------------------------
    def voronoi_init(self, n_cells: int = 0, n_iter: int = 64):
        self._check(self.L.orbc_voronoi_init(self.h, int(n_cells), int(n_iter)))

    def voronoi_update(self):
        self._check(self.L.orbc_voronoi_update(self.h))

    def voronoi_set_centroids(self, c):
        c = np.ascontiguousarray(c, dtype=np.float32)
        self._check(self.L.orbc_voronoi_set_centroids(self.h, len(c), _ptr(c, _f)))

    def voronoi_get(self):
        m = C.c_longlong()
        self._check(self.L.orbc_voronoi_count(self.h, C.byref(m)))
        c = np.zeros((m.value, 3), np.float32)
        al = np.zeros(self.n[LIPID], np.int32); ap = np.zeros(self.n[PROTEIN], np.int32)
        self._check(self.L.orbc_voronoi_get(self.h, _ptr(c, _f), _ptr(al, _i), _ptr(ap, _i)))
        return c, al, ap

    # -- device kernels exposed for parity checks --------------------------------
    def morton_keys(self, pts, bsize=2000.0):
        p = np.ascontiguousarray(pts, dtype=np.float32)
        k = np.zeros(len(p), np.uint32)
        self._check(self.L.orbc_morton_keys(self.h, len(p), _ptr(p, _f), bsize, _ptr(k, _u)))
        return k

    def sort_pairs(self, keys, values, bits=32):
        k = np.ascontiguousarray(keys, dtype=np.uint32).copy()
        v = np.ascontiguousarray(values, dtype=np.uint32).copy()
        self._check(self.L.orbc_sort_pairs(self.h, len(k), _ptr(k, _u), _ptr(v, _u), int(bits)))
        return k, v

    def sort_pairs_device(self, keys_ptr: int, values_ptr: int, n: int, bits=32):
        """orbc_sort_pairs_noalloc on caller-owned device arrays (raw device
        addresses); enqueued on the context's stream."""
        self._check(self.L.orbc_sort_pairs_noalloc(self.h, int(n), C.c_void_p(keys_ptr), C.c_void_p(values_ptr),
                                                   int(bits)))

    def nearest(self, pts, q):
        p = np.ascontiguousarray(pts, dtype=np.float32); qq = np.ascontiguousarray(q, dtype=np.float32)
        idx = np.zeros(len(qq), np.int32)
        self._check(self.L.orbc_nearest(self.h, len(p), _ptr(p, _f), len(qq), _ptr(qq, _f), _ptr(idx, _i)))
        return idx

    def within(self, pts, q, rmax, cap=512):
        p = np.ascontiguousarray(pts, dtype=np.float32); qq = np.ascontiguousarray(q, dtype=np.float32)
        cnt = np.zeros(len(qq), np.int32); out = np.zeros((len(qq), cap), np.int32)
        self._check(self.L.orbc_within(self.h, len(p), _ptr(p, _f), len(qq), _ptr(qq, _f), float(rmax), cap,
                                       _ptr(cnt, _i), _ptr(out, _i)))
        return [np.sort(out[k, :cnt[k]]) for k in range(len(qq))]


# ---- reference-shaped free functions (src/openrbc.cpp call sites) -----------
def compute_pairwise_fused(sim: Simulation):
    """compute_pairwise_fused(voronoi, lipid, protein, cl, cp): f,t += pair terms."""
    sim.compute_pairwise()


def compute_bonded(sim: Simulation):
    """compute_bonded(protein): f += bond terms."""
    sim.compute_bonded()


def integrate(kernel, sim: Simulation):
    """integrate(kernel, lipid, protein) with a reference functor name."""
    sim.integrate(kernel)


def constrain_volume(sim: Simulation, target_volume: float, strength: float) -> float:
    return sim.constrain_volume(target_volume, strength)


def compute_temperature(sim: Simulation) -> float:
    return sim.compute_temperature()


from .systems import (evans_fung, icosphere, init_rbc, load_example_large, planar_membrane, planar_patch,  # noqa: E402
                      random_sphere, system_to_simulation, vesicle_ensemble, vesicle_ensemble_mesh)

__all__ = [
    "Config", "Container", "Simulation", "OrbcError", "default_config", "load_library",
    "compute_pairwise_fused", "compute_bonded", "integrate", "constrain_volume", "compute_temperature",
    "read_frame", "POSITION", "ROTATION", "VORONOI", "VELOCITY", "FORCE",
    "evans_fung", "icosphere", "init_rbc", "load_example_large", "planar_membrane", "planar_patch", "random_sphere", "system_to_simulation", "vesicle_ensemble", "vesicle_ensemble_mesh",
]

#!/usr/bin/env python3
"""Benchmark: particle-timesteps/s of the per-timestep MD loop on one MI355X.

Workload: the full-size RBC of OpenRBC's example-large mesh (Evans-Fung
biconcave, 9,128 vertices) populated by the init_rbc rules: 2,639,690 lipids +
565,816 proteins = 3,205,506 particles, 584,066 bonds.  Langevin at kT=0.22,
dt=0.01, eta=0.01, seed 7; Verlet list rebuilt on the skin criterion;
Voronoi cells (n_lipid / 14, 64 Lloyd steps) updated every 2 steps and the
volume constraint at the initial enclosed volume applied every step.  The
mixed-precision step (FP32 pair/bond forces, FP64 state and reductions) is the
headline ``value``; the plain FP64 drop-in runs the same loop beside it, with
the force error of mixed vs FP64 on the same configuration and the total-energy
drift of both runs.  ``--mesh evans-fung`` generates the biconcave surface
instead (frequency-34 geodesic triangulation, 11,562 vertices, 4,122,700
particles).

A "step" is one pass of the loop: the Voronoi update (every second step), the
volume push and one fused launch of forces + bonds + Langevin update over all
particles (rebuilds, when the skin criterion fires, are inside the timed
region).  Inputs are resident in HBM before timing starts.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # MI355X HBM3E, B/s
METRIC = "particle-timesteps/s on 1 MI355X, FP32-mixed vs FP64, at stated force error"
# what the mixed path computes in (k_force.hip)
MIXED_DTYPE = ("mixed: fp32 neighbour mirrors, selection, WCA/protein-protein/bond terms; anisotropic "
               "lipid pair term in fp64 on the exact fp32-mirror separation; fp64 sums, state, integrator, "
               "reductions")


def algorithmic_bytes(precision: str, n: int, e_pairs: int, n_bonds: int) -> float:
    """Bytes per step the loop must move at minimum (state in, state out,
    one read of each pair index and bond): per particle 288 B (mixed) / 336 B
    (FP64) + 8 B per unordered pair + 12 B per bond."""
    base = 288.0 if precision == "mixed" else 336.0
    return base * n + 8.0 * e_pairs + 12.0 * n_bonds


def build_system(mesh: str = "example-large", seed: int = 7):
    import openrbc_amd as orbc
    if mesh == "example-large":
        v, e, f = orbc.load_example_large()
    else:
        v, e, f = orbc.evans_fung(34)
    return orbc.init_rbc(v, e, f, rho=1.05, seed=seed)


def upload(sim, sysd, state=None, seed=7, kT=0.22):
    import openrbc_amd as orbc
    from openrbc_amd.systems import maxwell
    rng = np.random.default_rng(seed)
    if state is None:
        lv = maxwell(np.zeros(len(sysd["lipid_x"]), np.int32), kT, rng)
        pv = maxwell(sysd["protein_type"], kT, rng)
        state = dict(lx=sysd["lipid_x"], ln=sysd["lipid_n"], lv=lv, lo=None,
                     px=sysd["protein_x"], pn=sysd["protein_n"], pv=pv, po=None)
    sim.set_state(orbc.LIPID, state["lx"], state["ln"], sysd["lipid_id"], v=state["lv"], o=state["lo"])
    sim.set_state(orbc.PROTEIN, state["px"], state["pn"], sysd["protein_id"], v=state["pv"], o=state["po"],
                  type=sysd["protein_type"])
    sim.set_bonds(sysd["bond_type"], sysd["bond_i"], sysd["bond_j"])


def download(sim):
    import openrbc_amd as orbc
    L = sim.get_state(orbc.LIPID)
    P = sim.get_state(orbc.PROTEIN)
    return dict(lx=L["x"], ln=L["n"], lv=L["v"], lo=L["o"], px=P["x"], pn=P["n"], pv=P["v"], po=P["o"])


def forces(sim):
    import openrbc_amd as orbc
    sim.compute_pairwise(zero_first=True)
    sim.compute_bonded()
    lf, lt = sim.get_forces(orbc.LIPID)
    pf, pt = sim.get_forces(orbc.PROTEIN)
    return np.vstack([lf, pf]), np.vstack([lt, pt])


def force_error(sim32, sim64, ids):
    """eps_F = max_i |F32_i - F64_i| / rms_j |F64_j| (and the same for the
    director force t), plus the 99.9th percentile of the per-particle
    relative error |dF_i| / max(|F64_i|, median |F64|).  Pairs whose cutoff
    test comes out differently in FP32 and FP64 (r within rounding of a cutoff;
    the WCA force jumps by 2.2e-3 at the reference's cut 1.1225 > 2^(1/6)) are
    counted, and eps_F is also given without the particles of such pairs."""
    f64, t64 = forces(sim64)
    f32, t32 = forces(sim32)
    rms_f = np.sqrt(np.mean(np.sum(f64 * f64, 1)))
    rms_t = np.sqrt(np.mean(np.sum(t64 * t64, 1)))
    dF = np.linalg.norm(f32 - f64, axis=1)
    dT = np.linalg.norm(t32 - t64, axis=1)
    fn = np.linalg.norm(f64, axis=1)
    p32 = sim32.pairs(); p64 = sim64.pairs()
    k32 = p32[:, 0] * (1 << 32) + p32[:, 1]; k64 = p64[:, 0] * (1 << 32) + p64[:, 1]
    flip = np.setxor1d(k32, k64)
    flip_ids = np.unique(np.concatenate([flip >> 32, flip & 0xffffffff]))
    keep = ~np.isin(ids, flip_ids)
    return dict(eps_F=float(dF.max() / rms_f), eps_T=float(dT.max() / rms_t),
                p99_9_rel=float(np.percentile(dF / np.maximum(fn, np.median(fn)), 99.9)), rms_F=float(rms_f),
                cutoff_flips=int(len(flip)), eps_F_no_flip=float(dF[keep].max() / rms_f),
                eps_T_no_flip=float(dT[keep].max() / rms_t))


def timed_run(sim, steps: int, warmup: int):
    """Returns wall seconds of `steps` steps, kernel timing, the observables
    after the run, the rebuild count and the observables before it.

    The timed steps run as the library runs them (batches replayed from
    captured hipGraphs).  HIP events around each fused kernel need the steps
    launched one by one, so the kernel's duration comes from the `steps`
    steps that follow the timed region, launched with events on the same
    stream (same kernel, same arguments)."""
    sim.run(warmup)
    sim.synchronize()
    obs0 = sim.observables()
    t0 = time.perf_counter()
    sim.run(steps)
    sim.synchronize()
    t1 = time.perf_counter()
    obs = sim.observables()
    sim.set_timing(True)
    sim.run(steps)
    sim.synchronize()
    tm = sim.timing()
    sim.set_timing(False)
    return t1 - t0, tm, obs, obs["rebuilds"] - obs0["rebuilds"], obs0


def cpu_baseline(sysd, steps: int, n_lloyd: int, threads: int, state=None, timeout: int = 300):
    """The OpenRBC reference itself (oracle/_ref, compiled in place from its
    sources) on the same system -- the minimised configuration the GPU runs
    start from when `state` is given -- in a subprocess."""
    from oracle import REF_SO
    if not os.path.exists(REF_SO):
        return None
    import tempfile
    path = os.path.join(tempfile.gettempdir(), f"orbc_cpu_sample_{os.getpid()}.npz")
    st = state if state is not None else dict(lx=sysd["lipid_x"], ln=sysd["lipid_n"], px=sysd["protein_x"],
                                              pn=sysd["protein_n"])
    np.savez(path, lx=st["lx"].astype(np.float32), ln=st["ln"].astype(np.float32),
             px=st["px"].astype(np.float32), pn=st["pn"].astype(np.float32),
             pt=sysd["protein_type"], pid=sysd["protein_id"].astype(np.int32), bt=sysd["bond_type"],
             bi=sysd["bond_i"].astype(np.int32), bj=sysd["bond_j"].astype(np.int32))
    code = f"""
import sys, json, ctypes as C, numpy as np
sys.path.insert(0, {ROOT!r})
from oracle import REF_SO
L = C.CDLL(REF_SO)
z = np.load({path!r})
P = lambda a, t: np.ascontiguousarray(a).ctypes.data_as(C.POINTER(t))
L.ref_run.restype = C.c_double
nf = C.c_long()
args = [len(z['lx']), P(z['lx'], C.c_float), P(z['ln'], C.c_float), len(z['px']), P(z['px'], C.c_float),
        P(z['pn'], C.c_float), P(z['pt'], C.c_int), P(z['pid'], C.c_int), len(z['bt']), P(z['bt'], C.c_int),
        P(z['bi'], C.c_int), P(z['bj'], C.c_int), C.c_double(0.22), C.c_double(0.01), C.c_double(0.01),
        C.c_int(7), C.c_int({steps}), C.c_int({n_lloyd}), C.c_int({threads}), C.byref(nf)]
t = L.ref_run(*args)
print(json.dumps(dict(seconds=t, n_final=nf.value)))
"""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    try:
        out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=timeout, env=env)
        line = [l for l in out.stdout.splitlines() if l.startswith("{")][-1]
        res = json.loads(line)
    except Exception as ex:   # reported, never fatal
        return dict(error=str(ex)[:200])
    finally:
        try:
            os.remove(path)
        except OSError:
            pass
    n = len(sysd["lipid_x"]) + len(sysd["protein_x"])
    _, model, avail = host_cpus()
    return dict(value=n * steps / res["seconds"], unit="particle-timesteps/s", cores=threads, kind="reference",
                sample=f"OpenRBC reference main loop (fused pairwise + bonded + verlet_langevin + Voronoi "
                       f"updates every 2 steps, FP32 as shipped, -O3 -ffast-math) on the same {n}-particle "
                       f"system{' (the minimised configuration of the GPU runs)' if state is not None else ''}: "
                       f"{steps} steps after the reference's {n_lloyd}-iteration Lloyd init, {threads} OpenMP "
                       f"threads (OMP_PROC_BIND=close, OMP_PLACES=cores)",
                cpu_model=model, logical_cpus_visible=avail, seconds=res["seconds"])


def host_cpus():
    """(threads to use, CPU model, logical CPUs this process may run on).
    The threads are the host's CPU share as the environment states it
    (OMP_NUM_THREADS, which the GPU host sets to its per-GPU share), else
    every CPU of the affinity mask."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = int(env) if env.isdigit() and int(env) > 0 else avail
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return max(1, min(threads, avail)), model, avail


def traffic_profile(name: str, mesh: str):
    """HBM bytes per launch of the fused step kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary, used only if that summary was
    measured on the kernel sources of this build (their sha256 is recorded in
    it); otherwise None."""
    import hashlib
    tpath = os.path.join(ROOT, "profiles", f"traffic_k_step_langevin_{name}_{mesh}.json")
    if not os.path.exists(tpath):
        return None, None
    with open(tpath) as fh:
        prof = json.load(fh)
    h = hashlib.sha256()
    for src in prof.get("sources", []):
        with open(os.path.join(ROOT, src), "rb") as fh:
            h.update(fh.read())
    if not prof.get("sources") or h.hexdigest() != prof.get("sources_sha256"):
        return None, os.path.relpath(tpath, ROOT) + " (measured on other kernel sources: not used)"
    src = os.path.relpath(tpath, ROOT)
    if prof.get("measured_with"):
        src += f" (measured with {prof['measured_with']})"
    return prof.get("hbm_bytes_per_launch"), src


def side_workload(args):
    """The batched-ensemble and planar-membrane workloads: one JSON line with
    the same metric on that workload (mixed precision)."""
    import openrbc_amd as orbc
    t0 = time.perf_counter()
    if args.workload == "ensemble" and args.vesicles == "mesh":
        e = orbc.vesicle_ensemble_mesh(n_members=64, seed=0)
        sim = orbc.Simulation(precision=orbc.MIXED, seed=7, skin=args.skin)
        rng = np.random.default_rng(7)
        kl, kp = e["kT"][e["lipid_group"]], e["kT"][e["protein_group"]]
        lv = orbc.systems.maxwell(np.zeros(len(kl), np.int32), 1.0, rng) * np.sqrt(kl)[:, None]
        pv = orbc.systems.maxwell(e["protein_type"], 1.0, rng) * np.sqrt(kp)[:, None]
        sim.set_state(orbc.LIPID, e["lipid_x"], e["lipid_n"], e["lipid_id"], v=lv)
        sim.set_state(orbc.PROTEIN, e["protein_x"], e["protein_n"], e["protein_id"], v=pv, type=e["protein_type"])
        sim.set_bonds(e["bond_type"], e["bond_i"], e["bond_j"])
        sim.set_groups(orbc.LIPID, e["lipid_group"], e["kT"])
        sim.set_groups(orbc.PROTEIN, e["protein_group"], e["kT"])
        n = len(e["lipid_x"]) + len(e["protein_x"])
        desc = (f"batched ensemble: 64 spherical cells from geodesic meshes through the initialiser (frequency 1-9, "
                f"counts log-uniform in 5k-300k), {n} particles, {len(e['bond_type'])} bonds, member kBT "
                f"0.22 x U(0.9, 1.1), Langevin, one store and one fused launch per step; Voronoi cells "
                f"updated every 2 steps, stray-lipid cleanup every 60")
        steps = args.steps if args.steps != 1000 else 500
    elif args.workload == "ensemble":
        e = orbc.vesicle_ensemble(n_members=64, seed=0)
        sim = orbc.Simulation(precision=orbc.MIXED, seed=7, skin=args.skin)
        rng = np.random.default_rng(7)
        v = rng.normal(0, 1, e["lipid_x"].shape) * np.sqrt(e["kT"][e["group"]])[:, None]
        sim.set_state(orbc.LIPID, e["lipid_x"], e["lipid_n"], e["lipid_id"], v=v)
        sim.set_state(orbc.PROTEIN, np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0, np.uint32),
                      type=np.zeros(0, np.int32))
        sim.set_groups(orbc.LIPID, e["group"], e["kT"])
        n = len(e["lipid_x"])
        desc = (f"batched ensemble: 64 lipid vesicles from random points on spheres, {n} particles (5k-300k "
                f"each, log-uniform), member kBT 0.22 x U(0.9, 1.1), Langevin, one store and one fused launch per "
                f"step; Voronoi cells updated every 2 steps, stray-lipid cleanup every 60")
        steps = args.steps if args.steps != 1000 else 500
    else:
        sysd = orbc.planar_membrane(seed=1)
        lo, hi, per = sysd["box"]
        sim = orbc.Simulation(precision=orbc.MIXED, thermostat=orbc.RUN_NOSE_HOOVER, skin=args.skin,
                              box_lo=list(lo), box_hi=list(hi), periodic=[int(p) for p in per])
        orbc.system_to_simulation(sysd, sim)
        n = len(sysd["lipid_x"]) + len(sysd["protein_x"])
        desc = (f"periodic planar membrane: 4096 x 4096 lattice, {len(sysd['lipid_x'])} lipids + "
                f"{len(sysd['protein_x'])} proteins, {len(sysd['bond_type'])} bonds, Nose-Hoover")
        steps = args.steps if args.steps != 1000 else 200
    print(f"# {desc} (set-up {time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    sim.minimize(args.minimize if args.minimize != 1000 else 100)
    if args.workload == "ensemble":
        # Voronoi cells as in the reference loop: updated every freq_voronoi steps, and the
        # stray-lipid cleanup every freq_cleanup steps removes lipids ejected from the vesicles
        sim.voronoi_init(0, 4)
    tuning = None
    if args.tune:
        tuning = sim.autotune(trial_steps=args.tune_steps or min(steps, 100), warm_steps=min(args.warmup, 20),
                              table=args.tune_table)
        tuning["knobs"] = sim.tuning()
        print(f"# autotune: {tuning}", file=sys.stderr, flush=True)
    wall, tm, obs, rebuilds, obs0 = timed_run(sim, steps, min(args.warmup, 20))
    line = {"metric": METRIC, "value": n * steps / wall, "unit": "particle-timesteps/s", "n_gpus": 1,
            "steps": steps, "warmup": min(args.warmup, 20), "ms_per_step": 1e3 * wall / steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": MIXED_DTYPE,
            "data": "synthetic", "config": {"workload": desc, "n_particles": n, "rebuilds": rebuilds,
                                            "temperature": obs["temperature"],
                                            "lipids_deleted": obs["lipids_deleted"],
                                            # per fused Langevin launch (HIP events); None for the
                                            # Nose-Hoover loop, whose kernels are not timed one by one
                                            "step_kernel_ms": (tm["step_kernel_ms"] / tm["step_launches"]
                                                               if tm["step_launches"] else None),
                                            "block_halos": {"blocks": obs["halo_blocks"],
                                                            "overflow": obs["halo_overflow"]}},
            "tuning": tuning}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--minimize", type=int, default=1000)
    ap.add_argument("--skin", type=float, default=0.6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tune", type=int, choices=[0, 1], default=1,
                    help="1: orbc_autotune before the timed region (untimed set-up), trials of --warmup + "
                         "--tune-steps steps")
    ap.add_argument("--tune-steps", type=int, default=0,
                    help="timed steps per autotune trial (0: --steps, at most 100)")
    ap.add_argument("--tune-table", default=None, help="append the autotuner's trial table (JSON lines) here")
    ap.add_argument("--graphs", type=int, choices=[0, 1], default=1,
                    help="1: orbc_run replays captured hipGraphs of its batches (default); 0: kernel by kernel")
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--cpu-lloyd", type=int, default=64, help="Lloyd iterations of the reference's own "
                    "Voronoi init (src/openrbc.cpp:69-74 uses 64)")
    ap.add_argument("--precision", choices=["both", "mixed", "fp64"], default="both")
    ap.add_argument("--mesh", choices=["example-large", "evans-fung"], default="example-large")
    ap.add_argument("--volume-strength", type=float, default=0.05,
                    help="constrain_volume strength (0: no volume constraint in the loop)")
    ap.add_argument("--lloyd", type=int, default=64, help="Lloyd iterations of the Voronoi init")
    ap.add_argument("--vesicles", choices=["mesh", "random"], default="mesh",
                    help="ensemble members: mesh = geodesic spheres through the initialiser (proteins, spectrin, "
                         "bonds); random = lipid-only random points on spheres (init_random_sphere)")
    ap.add_argument("--workload", choices=["rbc", "ensemble", "planar"], default="rbc",
                    help="rbc: the full-size RBC, mixed and FP64 (the headline line); ensemble: 64 vesicles "
                         "of uneven size with their own kBT in one store (mixed, Langevin); planar: the "
                         "4096 x 4096 periodic membrane with its spectrin network (mixed, Nose-Hoover)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MiB",
                    help="only run 3 plain device-to-device copies of MiB each (the reference kernel that "
                         "scales rocprofv3's FETCH_SIZE / WRITE_SIZE to bytes) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        nb = args.copy_calibration << 20
        a = torch.ones(nb // 4, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": nb, "copy_bytes_written": nb, "copies": 3}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(how the path would shard over GPUs is described with the design notes).", file=sys.stderr)
        sys.exit(2)

    if args.workload != "rbc":
        return side_workload(args)

    import openrbc_amd as orbc
    t_build = time.perf_counter()
    sysd = build_system(args.mesh, seed=7)
    n = len(sysd["lipid_x"]) + len(sysd["protein_x"])
    nb = len(sysd["bond_type"])
    print(f"# system: {len(sysd['lipid_x'])} lipids + {len(sysd['protein_x'])} proteins = {n}, {nb} bonds "
          f"({time.perf_counter() - t_build:.1f} s)", file=sys.stderr, flush=True)

    # minimise once in FP64 (untimed; the reference always minimises first, src/openrbc.cpp:88-146)
    sim64 = orbc.Simulation(precision=orbc.FP64, seed=7, skin=args.skin)
    upload(sim64, sysd)
    sim64.minimize(args.minimize)
    state = download(sim64)
    vol = dict(volume_target=-1.0 if args.volume_strength != 0 else 0.0, volume_strength=args.volume_strength)
    vol["graphs"] = args.graphs
    sim32 = orbc.Simulation(precision=orbc.MIXED, seed=7, skin=args.skin, **vol)
    upload(sim32, sysd, state)
    sim64.close()
    sim64 = orbc.Simulation(precision=orbc.FP64, seed=7, skin=args.skin, **vol)
    upload(sim64, sysd, state)
    t_v = time.perf_counter()
    for sim in (sim32, sim64):
        sim.voronoi_init(0, args.lloyd)
    print(f"# Voronoi cells: {len(sim32.voronoi_get()[0])} ({args.lloyd} Lloyd steps, "
          f"{time.perf_counter() - t_v:.1f} s for both)", file=sys.stderr, flush=True)

    # force error of the mixed path against FP64 on the same configuration
    ids = np.concatenate([sysd['lipid_id'], sysd['protein_id']]).astype(np.int64)
    err_min = force_error(sim32, sim64, ids)
    e_pairs = int(len(sim32.pairs()))
    print(f"# minimised: {err_min} pairs={e_pairs}", file=sys.stderr, flush=True)

    res = {}
    tuning = {}
    for name, sim in (("mixed", sim32), ("fp64", sim64)):
        if args.precision not in ("both", name):
            continue
        if args.tune:
            # untimed set-up, like the minimisation: orbc_autotune on this system (state restored)
            t_t = time.perf_counter()
            # trials shaped like the timed run: warmup steps, then the timed steps
            tuning[name] = sim.autotune(trial_steps=args.tune_steps or min(args.steps, 100), warm_steps=args.warmup,
                                        table=args.tune_table)
            tuning[name]["knobs"] = sim.tuning()
            print(f"# {name} autotune ({time.perf_counter() - t_t:.1f} s): {tuning[name]}", file=sys.stderr,
                  flush=True)
        wall, tm, obs, rebuilds, obs0 = timed_run(sim, args.steps, args.warmup)
        e0 = obs0["kinetic"] + obs0["potential"]
        e1 = obs["kinetic"] + obs["potential"]
        kern_ms = tm["step_kernel_ms"] / max(1, tm["step_launches"])
        B = algorithmic_bytes(name, n, e_pairs, nb)
        res[name] = dict(value=n * args.steps / wall, ms_per_step=1e3 * wall / args.steps, kernel_ms=kern_ms,
                         kernel_launches=tm["step_launches"], rebuild_ms=tm["rebuild_ms"], rebuilds=rebuilds,
                         bytes_per_step=B, achieved_GBs=B / (kern_ms * 1e-3) / 1e9,
                         temperature=obs["temperature"], potential_per_particle=obs["potential"] / n,
                         energy_per_particle=dict(start=e0 / n, end=e1 / n, drift=(e1 - e0) / n),
                         halo_blocks=obs["halo_blocks"], halo_overflow=obs["halo_overflow"],
                         halo_capacity=sim.tuning()["halo_capacity"],
                         graph_replays=obs["graph_replays"] - obs0["graph_replays"],
                         graph_captures=obs["graph_captures"],
                         graph_captures_timed=obs["graph_captures"] - obs0["graph_captures"])
        print(f"# {name}: {res[name]['value']:.4g} particle-steps/s, {res[name]['ms_per_step']:.3f} ms/step, "
              f"kernel {kern_ms:.3f} ms, rebuilds {rebuilds} ({tm['rebuild_ms']:.1f} ms), "
              f"T={obs['temperature']:.4f}, E/N {e0 / n:.5f} -> {e1 / n:.5f}, "
              f"halo blocks {obs['halo_blocks']} (overflow {obs['halo_overflow']})", file=sys.stderr, flush=True)

    # error again on the configuration reached by the mixed dynamics
    err_dyn = None
    if "mixed" in res and "fp64" in res:
        upload(sim64, sysd, download(sim32))
        err_dyn = force_error(sim32, sim64, ids)
        print(f"# after dynamics: {err_dyn}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(sysd, args.cpu_steps, args.cpu_lloyd, host_cpus()[0], state=state)
        print(f"# cpu baseline: {cpu}", file=sys.stderr, flush=True)

    head = res.get("mixed") or res.get("fp64")
    hname = "mixed" if "mixed" in res else "fp64"
    traffic, traffic_src = traffic_profile(hname, args.mesh)
    if traffic is not None and (not head["halo_blocks"] or head["halo_capacity"] != 896):
        # the profile is of the default halo kernel (896 local indices)
        traffic, traffic_src = None, traffic_src + " (of the 896-index halo kernel; this run used another: not used)"
    line = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "particle-timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": MIXED_DTYPE if hname == "mixed" else "fp64",
        "data": f"synthetic ({'Evans-Fung biconcave mesh, geodesic nu=34' if args.mesh == 'evans-fung' else 'OpenRBC example-large mesh'} "
                "populated by init_rbc rules, seed 7; FP64-minimised; Maxwell velocities)",
        "config": {
            "workload": f"full-size biconcave RBC ({args.mesh} mesh): {len(sysd['lipid_x'])} lipids + "
                        f"{len(sysd['protein_x'])} proteins, {nb} bonds; Langevin kT=0.22 dt=0.01 eta=0.01; "
                        f"skin rebuild; Voronoi update every 2 steps; volume constraint at the initial enclosed "
                        f"volume (strength {args.volume_strength})",
            "n_particles": n, "n_bonds": nb, "pairs_within_cutoff": e_pairs, "skin": args.skin,
            "precision": hname, "mesh": args.mesh,
            "block_halos": {"blocks": head["halo_blocks"], "overflow": head["halo_overflow"],
                            "capacity": head["halo_capacity"]},
            "graphs": {"replays_in_timed_region": head["graph_replays"], "captured": head["graph_captures"],
                       "captured_in_timed_region": head["graph_captures_timed"]},
            "rebuilds_in_timed_region": head["rebuilds"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": head["achieved_GBs"],
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": head["achieved_GBs"] * 1e9 / HBM_PEAK,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel": (("k_step_langevin_halo" + ("_wide" if head["halo_capacity"] == 1024 else ""))
                       if head["halo_blocks"] else "k_step_langevin") + " (fused pair + bond + Langevin)",
            "kernel_ms": head["kernel_ms"],
            "kernel_timing": f"HIP events around each fused kernel on the context stream, {args.steps} steps "
                             f"launched one by one right after the timed region",
            "algorithmic_bytes_per_launch": head["bytes_per_step"],
        },
        "cpu_baseline": cpu,
        "tuning": tuning or None,
        "force_error": {"minimised": err_min, "after_dynamics": err_dyn},
    }
    if "mixed" in res and "fp64" in res:
        line["fp64"] = {k: res["fp64"][k] for k in ("value", "ms_per_step", "kernel_ms", "achieved_GBs", "rebuilds",
                                                    "energy_per_particle")}
        line["energy_per_particle"] = res["mixed"]["energy_per_particle"]
        line["fp64"]["roofline_frac"] = res["fp64"]["achieved_GBs"] * 1e9 / HBM_PEAK
        line["mixed_over_fp64"] = res["mixed"]["value"] / res["fp64"]["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

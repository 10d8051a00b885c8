/*
 * orbc.h -- C ABI of liborbc_hip.so, the MI355X-native per-timestep MD loop of
 * OpenRBC (protein-resolution red-blood-cell membranes).
 *
 * Plain C: pointers, sizes and status codes; no torch or HIP types.  Every
 * entry returns an orbc_status; orbc_last_error() holds the message.  Nothing
 * here exits the process (the reference's rt_assert calls exit(0),
 * src/util_misc.h:54-59).
 *
 * Reference interface each entry replaces is cited as OpenRBC path:line.
 *
 * Ownership: an orbc_ctx owns all device memory; host arrays passed in are
 * copied, host arrays passed out are filled; the caller owns both.  A context
 * is not re-entrant: one host thread per context.  Work is queued on the
 * context's HIP stream and synchronised inside every orbc_get_* call.
 *
 * Particle order: the device keeps particles in its own spatial (Morton) order
 * and permutes them at every neighbour rebuild, as the reference reorders its
 * containers (src/reorder.h:73-149).  orbc_get_* always returns each species
 * in the order in which orbc_set_state() received it.
 */
#ifndef ORBC_H_
#define ORBC_H_

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orbc_ctx orbc_ctx;

typedef enum orbc_status {
    ORBC_OK = 0,
    ORBC_ERR_INVALID = 1,   /* bad argument or call order */
    ORBC_ERR_HIP = 2,       /* HIP runtime error */
    ORBC_ERR_NAN = 3,       /* non-finite state detected */
    ORBC_ERR_OOM = 4,       /* device allocation failed */
    ORBC_ERR_STATE = 5      /* no state uploaded yet */
} orbc_status;

/* species: LipidContainer / ProteContainer, src/container.h:117-157 */
enum { ORBC_LIPID = 0, ORBC_PROTEIN = 1 };

/* precision of pair and bonded forces (state and reductions are FP64 in both) */
enum { ORBC_FP64 = 0, ORBC_MIXED = 1 };

/* neighbour selection */
enum {
    ORBC_NEIGHBOR_EXACT = 0,    /* every pair within the cutoff (Verlet list + skin) */
    ORBC_NEIGHBOR_VORONOI = 1   /* the reference's Voronoi-stencil pair set,
                                   src/compute_pairwise_fused.h:253-316 */
};

/* integrator kernels, named after the reference functors */
enum {
    ORBC_CLEAR_FORCE = 0,                /* src/integrate_nh.h:58-67 */
    ORBC_VERLET_LANGEVIN = 1,            /* src/integrate_langevin.h:99-149 */
    ORBC_NH_INITIAL_BOUNCE_CLEAR = 2,    /* src/integrate_nh.h:178-235 */
    ORBC_NH_POST_TORQUE_FINAL = 3,       /* src/integrate_nh.h:237-273 */
    ORBC_POST_TORQUE = 4,                /* src/integrate_nh.h:146-154 */
    ORBC_BOUNCE_BACK = 5                 /* src/integrate_nh.h:124-144 */
};

/* thermostat used by orbc_run */
enum { ORBC_RUN_LANGEVIN = 0, ORBC_RUN_NOSE_HOOVER = 1 };

/* orbc_compute_* flags */
#define ORBC_F_ZERO_FIRST 1u    /* clear f,t before accumulating */
#define ORBC_F_ENERGY 2u        /* also accumulate per-particle energy and virial */

/* Runtime parameters: mirrors RTParameter, src/runtime_parameter.h:38-80. */
typedef struct orbc_config {
    int precision;              /* ORBC_FP64 | ORBC_MIXED */
    double kBT;                 /* 0.22  */
    double eta;                 /* 0.01  */
    double dt;                  /* 0.01  */
    double rho;                 /* 1.05  (initialisers) */
    int voronoi_cell_size;      /* 14    lipids per Voronoi cell (> 0) */
    double stray_tolerance;     /* 2.5   */
    int freq_voronoi;           /* 2     */
    int freq_sort_ctrd;         /* 24    */
    int freq_sort_bond;         /* 120   */
    int freq_cleanup;           /* 60    */
    unsigned long long seed;    /* 0xBAD5EED */
    double Q;                   /* Nose-Hoover mass; 0 => 0.01 * N at the first Nose-Hoover step of the
                                   uploaded system, then kept (src/runtime_parameter.h:76) */
    double zeta;                /* Nose-Hoover friction, 0 */
    double box_lo[3], box_hi[3];/* bounce-back box, +-1000 (src/runtime_parameter.h:111-115) */
    int periodic[3];            /* 0: bounce/open, 1: periodic over [box_lo, box_hi) (minimum image;
                                   positions wrapped at neighbour rebuilds; no reference counterpart) */
    double skin;                /* Verlet skin, 0.4 */
    int rebuild_period;         /* 0: rebuild on the skin criterion; k>0: every k steps */
    int neighbor_mode;          /* ORBC_NEIGHBOR_* */
    int thermostat;             /* ORBC_RUN_* */
    double volume_target;       /* volume constraint inside orbc_run: > 0 target volume, < 0 the
                                   volume enclosed at the first orbc_run step, 0 off (needs
                                   orbc_voronoi_init; cells updated every freq_voronoi steps) */
    double volume_strength;     /* constrain_volume strength */
    int device;                 /* HIP device ordinal */
    int block_size;             /* workgroup size of the force and step kernels: 0 (default) or 256, the
                                   size they are built for; any other value is rejected */
    int block_halo;             /* fused steps (Langevin, Nose-Hoover second half): stage each particle
                                   block's neighbours in LDS (FP32 mirrors, or FP64 positions and
                                   directors) and walk 16-bit local indices
                                   (built at every neighbour rebuild).  1 (default): while a list lasts
                                   4 steps or more on average; 2: always; 0: never (every neighbour
                                   gathered from the 32-bit list).  Same forces either way. */
    int graphs;                 /* orbc_run's Langevin batches: 1 (default) replay captured hipGraphs of
                                   the whole timestep (Voronoi / volume work, the fused step kernel),
                                   steps numbered on the device so one graph serves every batch of its
                                   length and parity; 0: launched kernel by kernel.  Same results. */
} orbc_config;

typedef struct orbc_observables {
    double kinetic;     /* sum 0.5 m v^2 (FP64) */
    double temperature; /* sum m v^2 / 3N, src/compute_temperature.h:23-29 */
    double potential;   /* closed-form pair + bond energy (absent in the reference) */
    double virial;      /* sum over pairs and bonds of dx . f */
    double zeta;        /* Nose-Hoover friction */
    long long n;        /* particles */
    long long step;     /* completed steps */
    long long rebuilds; /* neighbour rebuilds so far */
    long long pairs;    /* directed neighbour-list entries at the last rebuild */
    double anchor_h;    /* base (power-of-two) spacing of the FP32 anchor grid (mixed geometry) */
    double skin;        /* Verlet skin in use */
    /* the anchor grid per axis: anchor A sits at anchor_lo + anchor_spacing * (A + anchor_origin);
       anchor_bits is the packed width (anchor differences wrap by it on periodic axes) */
    double anchor_lo[3], anchor_spacing[3], anchor_origin[3];
    int anchor_bits[3];
    long long n_lipid, n_protein;   /* species counts (lipids shrink with the stray-lipid cleanup) */
    long long lipids_deleted;       /* lipids removed by the cleanup so far */
    long long halo_blocks;          /* particle blocks of the fused steps with an LDS halo at the last
                                       rebuild (0 when no halos were built) */
    long long halo_overflow;        /* blocks whose neighbours did not fit (32-bit gathers) */
    long long graph_captures;       /* batch graphs captured so far (orbc_run, cfg.graphs) */
    long long graph_replays;        /* batch graph launches so far */
} orbc_observables;

void orbc_config_default(orbc_config *cfg);
int orbc_create(const orbc_config *cfg, orbc_ctx **ctx);
int orbc_destroy(orbc_ctx *ctx);
const char *orbc_last_error(const orbc_ctx *ctx);
int orbc_get_config(const orbc_ctx *ctx, orbc_config *cfg);

/* Container upload (LipidContainer/ProteContainer fields x,v,n,o; protein
 * type and tag; src/container.h:65-157).  AoS xyz, FP64.  `type` is ignored
 * for lipids (always 0, src/container.h:128-130); `id` is the reference tag
 * (lipids: base + i, src/container.h:122-126).  v, o may be NULL (zero).
 * A re-upload with the same count and the same tags (and protein types) in
 * the same order updates the state in place; one that changes the count,
 * permutes the tags or changes a type rebuilds the device store. */
int orbc_set_state(orbc_ctx *ctx, int species, long long n, const double *x, const double *v,
                   const double *nrm, const double *o, const int *type, const unsigned *id);
/* ProteContainer::bonds (src/container.h:30-34,138): {type, tag_i, tag_j} */
int orbc_set_bonds(orbc_ctx *ctx, long long nb, const int *type, const unsigned *id_i, const unsigned *id_j);
/* Ensemble members.  No reference counterpart: this is the batching of many
 * independent systems (e.g. vesicles of uneven size) into one store and one
 * launch per step.  Particle k of `species` (upload order) belongs to member
 * group[k] in [0, n_groups); member g runs Langevin at kBT[g] (the noise
 * amplitude of src/integrate_langevin.h:262-266 with that kBT).  Members must
 * be placed apart in space (no pair between members is then in range); their
 * noise streams are independent because Philox is keyed by particle id.
 * Call after orbc_set_state of the species; n_groups is the same for both
 * species.  orbc_run with the Nose-Hoover thermostat rejects members. */
int orbc_set_groups(orbc_ctx *ctx, int species, long long n, const int *group, int n_groups, const double *kBT);
int orbc_get_state(orbc_ctx *ctx, int species, double *x, double *v, double *nrm, double *o);
/* Container::f and ::t (t = director force -dU/dn until a post-torque step
 * converts it).  e, w: per-particle energy and virial from the last
 * ORBC_F_ENERGY evaluation (may be NULL). */
int orbc_get_forces(orbc_ctx *ctx, int species, double *f, double *t, double *e, double *w);
int orbc_set_forces(orbc_ctx *ctx, int species, const double *f, const double *t);

/* VoronoiDiagram::update + VCellList::update (src/voronoi.h:78,154; driver
 * src/openrbc.cpp:200-205): rebuild the spatial order and neighbour list now. */
int orbc_build_neighbors(orbc_ctx *ctx);
/* compute_pairwise_fused (src/compute_pairwise_fused.h:238): f,t += pair terms */
int orbc_compute_pairwise(orbc_ctx *ctx, unsigned flags);
/* compute_bonded (src/compute_bonded.h:90): f += bond terms */
int orbc_compute_bonded(orbc_ctx *ctx, unsigned flags);
/* integrate(kernel, lipid, protein) (src/integrate_nh.h:29-55) */
int orbc_integrate(orbc_ctx *ctx, int kernel);
/* constrain_volume (src/constrain_volume.h:26-83), fixed semantics: cell
 * normals recomputed each call, mass of each particle's own type.  Needs
 * Voronoi cells (orbc_voronoi_init).  f += push; *volume = cone-sum volume. */
int orbc_constrain_volume(orbc_ctx *ctx, double target, double strength, double *volume);
/* compute_temperature (src/compute_temperature.h:23-29), FP64 reduction */
int orbc_compute_temperature(orbc_ctx *ctx, double *T);
/* The main loop (src/openrbc.cpp:189-256): nsteps timesteps with the
 * configured thermostat, rebuild policy and volume constraint. */
int orbc_run(orbc_ctx *ctx, long long nsteps);
/* Energy minimisation (src/openrbc.cpp:88-146): nsteps capped
 * steepest-descent moves. */
int orbc_minimize(orbc_ctx *ctx, long long nsteps, double dr_opt, double dn_opt);
int orbc_get_observables(orbc_ctx *ctx, orbc_observables *obs);

/* Interacting pairs of the current neighbour list, evaluated with the force
 * kernel's own selection test (1e-5 < r^2 < rc^2).  Unordered, id_i < id_j,
 * unsorted.  *count = total; at most cap written. */
int orbc_get_pairs(orbc_ctx *ctx, long long cap, unsigned *id_i, unsigned *id_j, long long *count);

/* Voronoi cells (src/voronoi.h:33-141): n_cells centroids, Lloyd init with
 * n_iter iterations (src/voronoi.h:53-75; n_cells <= 0 uses N_lipid/14). */
int orbc_voronoi_init(orbc_ctx *ctx, long long n_cells, int n_iter);
/* VoronoiDiagram::update + partition (src/voronoi.h:77-86,179-237) */
int orbc_voronoi_update(orbc_ctx *ctx);
int orbc_voronoi_set_centroids(orbc_ctx *ctx, long long n_cells, const float *c);
/* centroids (n_cells*3) and per-species affiliation in upload order */
int orbc_voronoi_get(orbc_ctx *ctx, float *centroids, int *aff_lipid, int *aff_protein);
int orbc_voronoi_count(const orbc_ctx *ctx, long long *n_cells);
/* delete_lipid (src/cleanup.h:29-91): in every Voronoi cell, removes the
 * lipids whose squared distance to the centroid is not below tolerance^2
 * times the cell's median one (element size/2 of the sorted distances), using
 * the current cells.  orbc_run does the same every freq_cleanup steps on
 * Voronoi-update steps (src/openrbc.cpp:200-202) while it keeps cells (volume
 * constraint on, or freq_cleanup > 0 with cells initialised).  The lipid count
 * shrinks (orbc_observables.n_lipid); orbc_get_state returns the remaining
 * lipids in upload order.  *n_deleted = lipids removed by this call. */
int orbc_delete_stray_lipids(orbc_ctx *ctx, double tolerance, long long *n_deleted);

/* Device kernels behind reorder_morton / KDTree, exposed for parity checks:
 * morton_encode (src/reorder_morton.h:37-42) with bsize, on the GPU. */
int orbc_morton_keys(orbc_ctx *ctx, long long n, const float *pts, double bsize, unsigned *keys);
/* Stable LSD radix sort of (key, value) pairs on the GPU (bits <= 32). */
int orbc_sort_pairs(orbc_ctx *ctx, long long n, unsigned *keys, unsigned *values, int bits);
/* No counterpart in the reference: orbc_sort_pairs for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  keys and values are DEVICE arrays of n
 * entries, sorted in place (stable, by the low `bits` bits of the key, the
 * same order as orbc_sort_pairs).  Everything else the sort needs lives in
 * a fixed per-device scratch of the library (about 9 MiB, held by every GPU
 * the library runs on until the process ends); the call allocates nothing
 * and frees nothing.  n <= ORBC_NOALLOC_SORT_MAX, else ORBC_ERR_INVALID.
 * Asynchronous on the context's stream (orbc_synchronize before reading the
 * results).  Calls that overlap on one GPU, from any context, race on that
 * GPU's scratch: serialise them. */
#define ORBC_NOALLOC_SORT_MAX (1LL << 20)
int orbc_sort_pairs_noalloc(orbc_ctx *ctx, long long n, unsigned *keys, unsigned *values, int bits);
/* KDTree::find_nearest / find_within replacements (src/kdtree.h:206-285):
 * nearest centroid per query and the centroid stencil within rmax. */
int orbc_nearest(orbc_ctx *ctx, long long npts, const float *pts, long long nq, const float *q, int *idx);
int orbc_within(orbc_ctx *ctx, long long npts, const float *pts, long long nq, const float *q, float rmax,
                int cap, int *counts, int *out);

/* Trajectory and topology files (src/trajectory.h:26-150, src/topology.h:23-52).
 * orbc_save_frame writes (append != 0: appends) one frame in the reference's
 * binary layout: "FRAMEBEG" step (int32), "NATOM" count (uint64),
 * "IDENTITY" tag and type (int32 each), then the sections dump_field selects
 * with the reference's DumpField bits (1 POSITION, 2 ROTATION, 4 VORONOI,
 * 8 VELOCITY, 16 FORCE; FP32 vectors, int32 cells), "FRAMEEND".  Lipids
 * first, then proteins, each in upload order.
 * orbc_read_frame reads frame `index` (0-based): *step, *natom, *fields (the
 * DumpField bits present); array arguments may be NULL (call once with NULLs
 * to size them; tag/type hold natom entries, vectors 3 natom).
 * orbc_save_topology writes the reference's LAMMPS data file (proteins, then
 * lipids; bonds by tag). */
int orbc_save_frame(orbc_ctx *ctx, const char *path, int dump_field, int append);
int orbc_read_frame(const char *path, long long index, long long *step, long long *natom, int *fields,
                    unsigned *tag, int *type, float *x, float *v, float *n, int *aff, float *f);
int orbc_save_topology(orbc_ctx *ctx, const char *path);
/* Restart files (no reference counterpart; row f2's resume): the complete
 * FP64 state (configuration with the current Nose-Hoover friction and
 * resolved volume target, step counter, both species with x, v, n, o and the
 * force accumulators, ensemble members, bonds, Voronoi centroids and cells).
 * orbc_load_restart replaces the context's configuration (its device must
 * match) and state; a run continued from it follows the uninterrupted run
 * (same noise: it is keyed by particle id and step). */
int orbc_save_restart(orbc_ctx *ctx, const char *path);
int orbc_load_restart(orbc_ctx *ctx, const char *path);

int orbc_synchronize(orbc_ctx *ctx);

/* Kernel timing of orbc_run with HIP events on the context's stream:
 * accumulated time of the fused step kernel and of neighbour rebuilds
 * (measurement hook; the reference's Timers, src/timer.h:52-69). */
/* Run-time autotuner of orbc_run (no reference counterpart).  Measures the
 * knobs that decide the main loop's speed on the uploaded system itself: the
 * Verlet skin, block halos of the fused step (block_halo 0/1/2), the edge of
 * the centroid grid behind the nearest-centroid partition (with Voronoi
 * cells), and the steps launched per host synchronisation.  Every trial
 * restores the saved state, changes one knob and times orbc_run over
 * trial_steps steps (wall clock) after warm_steps untimed ones (< 0:
 * trial_steps / 4) -- the shape of the run the caller is about to do; knobs
 * are tuned one after another in that order.  The state (positions,
 * velocities, directors, step counter, cells, Nose-Hoover friction) is
 * restored at the end and the best configuration applied; the stray-lipid
 * cleanup does not run inside the trials.  With table_path, one JSON line of
 * the trials and the choice is appended to that file. */
typedef struct orbc_tuning {
    double skin;
    int block_halo;
    double centroid_cell;        /* edge of the centroid grid (default 8) */
    int batch;                   /* steps launched per host synchronisation, 1..32 (default 32) */
    double ms_per_step;          /* measured, chosen configuration (orbc_autotune) */
    double ms_per_step_start;    /* measured, configuration before tuning (orbc_autotune) */
    int trials;
    int halo_capacity;           /* local indices per block halo: 896 (5 blocks per CU) or 1024 (4 per CU,
                                    fewer overflowing blocks); 0 in orbc_set_tuning: unchanged */
} orbc_tuning;
int orbc_autotune(orbc_ctx *ctx, int warm_steps, int trial_steps, const char *table_path, orbc_tuning *out);
/* apply a configuration (e.g. a line of a stored table); takes effect at the next rebuild */
int orbc_set_tuning(orbc_ctx *ctx, const orbc_tuning *t);
int orbc_get_tuning(const orbc_ctx *ctx, orbc_tuning *t);

int orbc_set_timing(orbc_ctx *ctx, int on);
int orbc_get_timing(orbc_ctx *ctx, double *step_kernel_ms, long long *step_launches, double *rebuild_ms,
                    long long *rebuilds);

/* ---- initialiser: init_rbc (src/init_rbc.h:30-337) on the host ----------
 * Builds glycophorin, actin, band-3, spectrin and lipids from a triangle mesh
 * (vert: nv*3, edge: ne*2, face: nf*3, 0-based, the example-large text
 * format) at lipid density rho.  Arrays are malloc'd; free with
 * orbc_system_free.  Ids follow the reference tags: proteins 1..np, lipids
 * np+1..; bonds by id. */
typedef struct orbc_system {
    long long n_lipid, n_protein, n_bond;
    double *lipid_x, *lipid_n;        /* 3*n_lipid */
    unsigned *lipid_id;
    double *protein_x, *protein_n;    /* 3*n_protein */
    int *protein_type;
    unsigned *protein_id;
    int *bond_type;
    unsigned *bond_i, *bond_j;
    double area;                      /* total mesh area (reduced units) */
    long long clutter_capped;         /* lipids accepted after the retry cap */
} orbc_system;

int orbc_init_rbc(const double *vert, long long nv, const int *edge, long long ne, const int *face, long long nf,
                  double rho, unsigned long long seed, orbc_system *out);
void orbc_system_free(orbc_system *sys);

#ifdef __cplusplus
}
#endif
#endif

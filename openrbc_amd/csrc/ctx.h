// ctx.h -- the orbc_ctx structure and host helpers shared by api.cpp and
// voronoi.cpp.  Internal; the public ABI is include/orbc.h.
#pragma once

#include <algorithm>
#include <list>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/orbc.h"
#include "orbc_internal.h"

using orbc::Anchor;
using orbc::Box;
using orbc::FFTab;
using orbc::Grid;
using orbc::HipError;
using orbc::PQ;
using orbc::Gate;

struct DevArena {
    std::vector<void *> ptrs;
    template <class T>
    T *alloc(size_t n) {
        void *p = nullptr;
        if (n == 0) n = 1;
        hipError_t e = hipMalloc(&p, n * sizeof(T));
        if (e != hipSuccess) throw HipError(e, "hipMalloc", __FILE__, __LINE__);
        ptrs.push_back(p);
        return static_cast<T *>(p);
    }
    void release(void *p) {
        if (!p) return;
        auto it = std::find(ptrs.begin(), ptrs.end(), p);
        if (it != ptrs.end()) { (void)hipFree(p); ptrs.erase(it); }
    }
    void free_all() {
        for (void *p : ptrs) (void)hipFree(p);
        ptrs.clear();
    }
};

struct Species {
    long long n = 0;
    std::vector<double> x, v, nrm, o;
    std::vector<int> type;
    std::vector<uint32_t> id;
    std::vector<uint16_t> group;     // ensemble member per particle (empty: member 0)
};

struct orbc_ctx {
    orbc_config cfg;
    std::string err;
    hipStream_t stream = nullptr;
    DevArena mem;
    FFTab<double> ffd;
    FFTab<float> fff;

    Species sp[2];
    std::vector<int> b_type;
    std::vector<uint32_t> b_i, b_j;
    bool bonds_dirty = false;

    // device store
    bool live = false;
    long long n = 0, nl = 0, np = 0, stride = 0;
    long long pslot0 = 0;         // first protein slot (upload order: lipids, then proteins)
    long long last_cleanup = -1;  // step of the last stray-lipid cleanup inside orbc_run
    long long deleted = 0;        // lipids removed by the cleanup so far
    bool ft_scratch = false;      // the force accumulators are scratch (inside orbc_run's Langevin loop)
    double *x[2] = {nullptr, nullptr}, *nrm[2] = {nullptr, nullptr};
    int cur = 0;
    double *v = nullptr, *o = nullptr, *v_alt = nullptr, *o_alt = nullptr;
    double *f64 = nullptr, *t64 = nullptr, *f64_alt = nullptr, *t64_alt = nullptr;
    float *f32 = nullptr, *t32 = nullptr, *f32_alt = nullptr, *t32_alt = nullptr;
    double *e = nullptr, *w = nullptr;
    double *fe = nullptr, *te = nullptr;     // scratch force buffers for energy evaluation
    float *fe32 = nullptr, *te32 = nullptr;
    double *f_ext = nullptr;                 // volume-constraint force for orbc_run
    PQ *pq[2] = {nullptr, nullptr};
    float *rb = nullptr;
    uint32_t *anc = nullptr;
    uint8_t *type = nullptr, *type_alt = nullptr;
    uint32_t *id = nullptr, *id_alt = nullptr, *slot = nullptr, *slot_alt = nullptr, *inv = nullptr;
    uint32_t *keys = nullptr, *vals = nullptr, *keys_alt = nullptr, *vals_alt = nullptr;
    void *sort_temp = nullptr;
    uint32_t *cstart = nullptr, *cend = nullptr;
    long long ncell_cap = 0;
    uint32_t *nbr = nullptr, *nnbr = nullptr;
    int kmax = 0;
    // block halos of the mixed step (launch_halo), valid after a rebuild when halo_ok
    uint32_t *halo_ext = nullptr, *halo_cnt = nullptr, *nbr16 = nullptr;
    uint8_t *hord = nullptr;      // lane order per halo block (halo_nblk_cap * BLOCK)
    int halo_cap = orbc::HALO_CAP;      // local indices of the next halo build (autotuner knob: HALO_CAP, HALO_CAP_WIDE)
    int halo_cap_built = orbc::HALO_CAP;   // ... of the halos built at the last rebuild
    long long halo_nblk_cap = 0, nbr16_cap = 0;
    bool halo_ok = false;
    uint32_t *bnd = nullptr;
    uint8_t *nbnd = nullptr;
    int kbmax = 0;
    uint32_t *bstart = nullptr, *bpartner = nullptr;
    uint8_t *btype_d = nullptr;
    long long nbonds = 0;
    unsigned *flags = nullptr;   // [0] disp_max bits [1] bad [2] max_count [4..6] step displacement slots
    int *iflags = nullptr;       // [0] skip [1] done
    double *red = nullptr;       // reduction partials
    double *scal = nullptr;      // [0] zeta [1..7] sums
    double *bbox = nullptr;      // 6 + 6*256
    double *stage = nullptr;     // 3n doubles, upload-order staging
    float *stagef = nullptr;
    void *scratch = nullptr;     // this device's fixed sort scratch (orbc_sort_pairs_noalloc), set at create
    // captured batches of orbc_run's Langevin loop (cfg.graphs): a graph of
    // `len` steps numbered from the device counter *step_dev, found by the
    // signature of everything else its kernels bake in
    struct BatchGraph {
        std::vector<unsigned char> sig;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        unsigned long long last_use = 0;
    };
    std::list<BatchGraph> graphs;     // a list: segments of one batch keep pointers to their entries
    bool graphs_ok = true;            // capture works here (cleared at the first failure)
    unsigned long long graph_clock = 0;
    long long graph_captures = 0, graph_replays = 0;
    unsigned long long *step_dev = nullptr;
    float *bud2 = nullptr;            // squared skin budget of isolated particles (k_budget), 0: skin/2
    // device copies of the anchor grid and the centroid grid, rewritten at
    // every neighbour rebuild; kernels of captured batches read them there
    Anchor *anc_dev = nullptr;
    Grid *cg_dev = nullptr;
    bool capturing = false;           // a batch is being captured (voro_args points kernels at cg_dev)
    unsigned *pinned = nullptr;  // host-pinned flags
    double *pinned_d = nullptr;

    Grid grid;
    Anchor ag;
    Box box;
    bool need_rebuild = true;
    bool moved = false;          // positions changed by a standalone kernel since the last check
    long long step = 0, rebuilds = 0, last_pairs = 0, since_build = 0;
    bool nh_started = false;
    double nh_q = 0.0;           // Nose-Hoover mass resolved at its first use (cfg.Q == 0: 0.01 N then)
    unsigned types_present = 0;  // bit t set if some particle has type t
    int batch = 32;              // fused steps launched per host synchronisation
    double cg_edge = 8.0;        // edge of the centroid grid (nearest-centroid partition)
    bool tuning = false;         // inside orbc_autotune's trials: no stray-lipid cleanup
    bool nh_per_step = false;    // Nose-Hoover run loop one step at a time (ORBC_NH_PER_STEP=1, tests)
    // kernel timing (orbc_set_timing)
    bool timing = false;
    double t_step_ms = 0, t_rebuild_ms = 0;
    long long n_step_launches = 0, n_timed_rebuilds = 0;
    std::vector<hipEvent_t> ev;

    // Voronoi (k_voronoi / voronoi.cpp)
    struct Voro {
        long long ncell = 0;
        int32_t *aff = nullptr, *aff_alt = nullptr;
        float *c = nullptr;           // centroids
        unsigned long long *sums = nullptr;   // 4 * ncell fixed-point sums
        double *partial = nullptr;            // reduction partials over cells
        double *normals = nullptr;    // 3 * ncell
        uint32_t *ck = nullptr, *cv = nullptr, *ck2 = nullptr, *cv2 = nullptr;
        uint32_t *cg_start = nullptr, *cg_end = nullptr, *cg_idx = nullptr;
        long long cg_cap = 0;
        void *temp = nullptr;
        Grid cg;
        long long updates = 0;
        bool aff_valid = false;            // aff holds an affiliation (warm start of the next partition)
        float *cref = nullptr;             // centroids as binned in the centroid grid
        unsigned *cdisp = nullptr;         // max squared drift since (float bits)
        Gate gate;                         // skip gate of the batched run (VoroArgs::gate)
        // stray-lipid cleanup scratch: member ranges per cell key, delete flags
        uint32_t *st_start = nullptr, *st_end = nullptr, *del = nullptr;
        uint32_t *csize = nullptr;         // lipid count per cell + "current" word (VoroArgs::csize)
    } voro;
    double list_life = 0.0;       // steps a neighbour list lasts on average (batch sizing)
    // ensemble members (orbc_set_groups): per-particle member and per-member kBT
    std::vector<double> group_kbt;
    uint16_t *grp = nullptr, *grp_alt = nullptr;
    double *sig_group = nullptr;
    bool far_rings = false;       // a present type searches past the first ring of cells
    int rings[6] = {1, 1, 1, 1, 1, 1};   // grid rings each type searches
    int anchor_bits[3] = {0, 0, 0};      // anchor bits per axis (wrap width on periodic axes)
    // volume constraint inside orbc_run (cfg.volume_target != 0)
    double vol_target = 0.0;
    bool vol_target_set = false;

    bool mixed() const { return cfg.precision == ORBC_MIXED; }
};


namespace orbc {
int fail(orbc_ctx *c, int code, const std::string &msg);
void sync(orbc_ctx *c);
unsigned read_flag(orbc_ctx *c, int k);
void zero_flag(orbc_ctx *c, int k);
void ensure_live(orbc_ctx *c);
void ensure_list(orbc_ctx *c);
void rebuild(orbc_ctx *c);
void voronoi_free(orbc_ctx *c);
// destroys the captured batch graphs (they hold the store's device pointers)
inline void free_graphs(orbc_ctx *c) {
    for (auto &g : c->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    c->graphs.clear();
}
void voronoi_rebuild_index(orbc_ctx *c);
void voronoi_update_dev(orbc_ctx *c);
// stray-lipid cleanup (src/cleanup.h:29-91) with the current cells; removes
// the strays from the store and rebuilds.  Returns the number removed.
long long delete_stray_lipids(orbc_ctx *c, double tolerance);
void compact_store(orbc_ctx *c, const uint32_t *del, long long ndel);
// restart files (io.cpp): replace the configuration; restore cells by upload order
int orbc_set_config_internal(orbc_ctx *c, const orbc_config &cfg);
void voronoi_set_affiliation(orbc_ctx *c, const int *aff_lipid, const int *aff_protein);
double volume_push(orbc_ctx *c, double target, double strength, double *f_ext, bool to_f, bool read_back);
void volume_eval(orbc_ctx *c, double target, double strength);
}  // namespace orbc

template <class F>
int orbc_guarded(orbc_ctx *c, F &&f) {
    if (!c) return ORBC_ERR_INVALID;
    try {
        return f();
    } catch (const orbc::HipError &h) {
        char buf[512];
        std::snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s at %s:%d", (int)h.err, hipGetErrorString(h.err),
                      h.expr, h.file, h.line);
        return orbc::fail(c, h.err == hipErrorOutOfMemory ? ORBC_ERR_OOM : ORBC_ERR_HIP, buf);
    } catch (const std::exception &ex) {
        return orbc::fail(c, ORBC_ERR_INVALID, ex.what());
    }
}

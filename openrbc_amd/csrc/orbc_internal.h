// orbc_internal.h -- shared definitions of the gfx950 kernels and the host
// side of liborbc_hip.so.  Not part of the public ABI (include/orbc.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace orbc {

constexpr int NT = 6;     // particle types, src/forcefield_canonical.h:31-37
constexpr int NBT = 4;    // bond types, src/forcefield_canonical.h:101-109
constexpr int BLOCK = 256;
constexpr uint32_t BOND_INDEX_MASK = 0x3fffffffu;

// ---------------------------------------------------------------------------
// Force-field constants, src/forcefield_canonical.h:24-28,40-109.
// The lipid-lipid values are compile-time constants so the dominant pair type
// needs no table lookups; the tables are filled from the same expressions.
constexpr double kCutLL = 2.6, kReqLL = 1.1225, kEpsLL = 1.2, kAlphaLL = 1.55;
constexpr double ipow_c(double b, int e) { return e == 0 ? 1.0 : b * ipow_c(b, e - 1); }
constexpr double kRepLL = kEpsLL / ipow_c(kCutLL - kReqLL, 8);
constexpr double kAttLL = -2.0 * kEpsLL / ipow_c(kCutLL - kReqLL, 4);
constexpr double kCut2LL = kCutLL * kCutLL;

template <typename R>
struct FFTab {
    R cut2_poly[36], cut_poly[36], rep[36], att[36], alpha[36];
    R cut2_lj[36], lj1[36], lj2[36];
    int aniso[36];
    R mass[NT], inv_mass[NT];
    R bond_r0[NBT], bond_K[NBT];
};

void ff_fill(FFTab<double> &d, FFTab<float> &f);          // host, ff.cpp
double ff_max_cut(int a, int b);                            // rc(a,b), 0 if none

// ---------------------------------------------------------------------------
// Geometry of the FP32 mirrors: position = h * A + r, A an integer anchor per
// axis fixed at the last neighbour build, r the FP32 offset.  A - A0 is packed
// into 32 bits (bits[0] | bits[1] << shy | bits[2] << shz).
// Packed 32-bit integer anchor of each particle; the FP32 mirror stores the
// offset r = x - anchor position.  Per axis d the position of anchor A is
// lo[d] + hd[d] * (A + A0[d]).  Open axes: lo = 0 and hd a power of two, so
// hf * dA is exact in FP32.  Periodic axes: A0 = 0, lo = box origin,
// hd = L / 2^bits, and anchor differences wrap to the nearest image by sign
// extension from `bits` (wrap = 32 - bits; 0 on open axes).
struct Anchor {
    double hd[3];    // anchor spacing
    float hf[3];
    double lo[3];
    long long A0[3]; // anchor origin (open axes)
    unsigned shy, shz, mx, my, mz;
    unsigned wrap[3];
    double L[3], Linv[3];   // periodic axes: box length (FP64 minimum image); 0 on open axes
};

struct AnchorI {
    int x, y, z;
};
__device__ inline AnchorI anchor_unpack(uint32_t w, const Anchor &a) {
    return AnchorI{(int)(w & a.mx), (int)((w >> a.shy) & a.my), (int)(w >> a.shz)};
}
// anchor difference to the nearest image (identity on open axes)
__device__ inline int anchor_wrap(int dA, unsigned s) { return (int)((unsigned)dA << s) >> s; }
// FP32 distance between two anchors along one axis
__device__ inline float anchor_dx(const Anchor &a, int axis, int dA) {
    return a.hf[axis] * (float)anchor_wrap(dA, a.wrap[axis]);
}
// r + (distance between the anchors) with one rounding: on open axes the
// product is exact (power-of-two spacing), on periodic axes it is not
__device__ inline float anchor_add(const Anchor &a, int axis, float r, int dA) {
    return __fmaf_rn(a.hf[axis], (float)anchor_wrap(dA, a.wrap[axis]), r);
}
// the same separation in FP64 without rounding: the offsets differ exactly in
// FP64, and h * dA is exact on open axes (on periodic axes it carries the
// rounding of h = L / 2^bits only)
__device__ inline double anchor_add_exact(const Anchor &a, int axis, float ri, float rj, int dA) {
    return __fma_rn(a.hd[axis], (double)anchor_wrap(dA, a.wrap[axis]), (double)ri - (double)rj);
}
// FP64 position of anchor A along one axis
__device__ inline double anchor_pos(const Anchor &a, int axis, long long A) {
    return a.lo[axis] + a.hd[axis] * (double)(A + a.A0[axis]);
}
// FP64 minimum image on the periodic axes
__device__ inline void min_image(const Anchor &a, double &dx, double &dy, double &dz) {
    if (a.L[0] > 0) dx -= a.L[0] * rint(dx * a.Linv[0]);
    if (a.L[1] > 0) dy -= a.L[1] * rint(dy * a.Linv[1]);
    if (a.L[2] > 0) dz -= a.L[2] * rint(dz * a.Linv[2]);
}

// Uniform binning grid of the neighbour build (Morton-ordered cells).  On a
// periodic axis the cells tile [lo, lo + L) exactly and neighbour cells wrap.
struct Grid {
    double lo[3];
    double cs;       // nominal cell edge
    double csd[3];   // cell edge per axis (>= cs; L / dim on periodic axes)
    int dim[3];
    int bits[3];     // Morton bits per axis
    int per[3];      // periodic axis
    long long ncell;
};

// Periodic box (PBC on an axis when per[d] != 0), used by the kernels.
struct Box {
    double lo[3], hi[3], L[3];
    int per[3];
};

// Skip gate for work launched ahead inside a batched run (orbc_run).  The
// work of step k is enqueued before step k's fused kernel; when *stop is set or
// step k-1 raised its skin flag (*slot), step k will be skipped, and every
// kernel of its work returns at once.  A default Gate is always open.
// Skip gate of the work queued before a batched step (see StepArgs::dslot):
// closed when the run has stopped or when the step's input positions violate
// the skin (slot (step + 2) % 3).  The step is *step0 + k when step0 is set
// (steps numbered on the device, so a captured graph serves any start step),
// else k.
struct Gate {
    const int *stop = nullptr;
    const unsigned *slots = nullptr;
    const unsigned long long *step0 = nullptr;
    unsigned long long k = 0;
};
__device__ inline bool gate_closed(const Gate &g) {
    if (!g.stop) return false;
    const unsigned long long st = (g.step0 ? *g.step0 : 0ull) + g.k;
    return *g.stop != 0 || g.slots[(st + 2ull) % 3ull] != 0u;
}

// FP32 mirror of one particle for the mixed path, one 32-byte record so a
// neighbour gather touches one cache line: x = position offset from the
// anchor (xyz) + packed anchor (w, bits); n = director (xyz) + type (w, bits).
struct PQ {
    float4 x, n;
};

// Block halos of the mixed step (k_halo, k_step_langevin): particle block b
// (the BLOCK particles from b * BLOCK in store order) stages its own mirrors
// and those of at most HALO_EXT distinct outside neighbours in LDS, and walks
// its list as 16-bit entries (10-bit index into that table, 3-bit type).  A block with more outside
// neighbours is marked HALO_OVERFLOW and gathers from the 32-bit list.
#ifndef ORBC_HALO_CAP
#define ORBC_HALO_CAP 896
#endif
constexpr int HALO_CAP = ORBC_HALO_CAP;        // local indices: BLOCK own + HALO_EXT outside (default)
constexpr int HALO_EXT = HALO_CAP - BLOCK;
// the wide variant (the autotuner's halo capacity knob): 1024 local indices,
// the most the 10-bit index holds, at one block fewer per CU
constexpr int HALO_CAP_WIDE = 1024;
constexpr int HALO_EXT_MAX = HALO_CAP_WIDE - BLOCK;   // ext entries kept per block (array stride)
constexpr uint32_t HALO_OVERFLOW = 0xffffffffu;
// entries per group of the LDS walk (rows are padded to whole groups)
#ifndef ORBC_HALO_GROUP
#define ORBC_HALO_GROUP 4
#endif
constexpr int HALO_GROUP = ORBC_HALO_GROUP;
static_assert(HALO_GROUP == 4 || HALO_GROUP == 8, "halo group of 4 or 8 entries");
// 32-bit words of 16-bit entries in a row of nn entries (whole groups)
__host__ __device__ inline int halo_row_words(int nn) { return (HALO_GROUP / 2) * ((nn + HALO_GROUP - 1) / HALO_GROUP); }
inline long long halo_words(int kmax) { return halo_row_words(kmax); }

// Everything a step kernel reads or writes, by value in kernel arguments.
struct StepArgs {
    long long n;
    long long stride;          // ELL column stride (>= n)
    // FP64 state (x, n read from *_in and written to *_out; v, o in place)
    const double *x_in, *n_in;
    double *x_out, *n_out, *v, *o;
    // FP32 mirrors (mixed precision)
    const PQ *pq_in;
    PQ *pq_out;
    const uint8_t *type;
    const uint32_t *id;
    const uint32_t *anc;       // packed anchors (FP64 path)
    const float *rb;           // offsets at the last build (skin check)
    const uint32_t *nbr, *nnbr;
    // block halos (mixed path; null: every block gathers from nbr)
    const uint32_t *halo_ext, *halo_cnt, *nbr16;
    const uint8_t *hord;     // lane order of each halo block (k_halo): lane r takes particle blk * 256 + hord[.. + r]
    int halo_cap;            // local indices of the halos (HALO_CAP or HALO_CAP_WIDE): the step-kernel variant
    const uint32_t *bnd;
    const uint8_t *nbnd;
    int has_bonds;
    const double *f_ext;       // extra per-particle force, or null
    // volume push fused into the step (constrain_volume): f += *push * mass_type * normal[aff]
    const double *push;        // device scalar strength (V0 - V) / V0, or null
    const double *cell_normal; // 3 per Voronoi cell
    const int32_t *aff;        // Voronoi cell of each particle
    Anchor anc_grid;
    double dt;
    double gamma[NT], sigma[NT];
    // ensemble members with their own temperature: sigma of particle i is
    // sig_group[grp[i] * NT + type]; null: sigma[type]
    const uint16_t *grp;
    const double *sig_group;
    unsigned long long seed, step;
    // step0 set: the step is *step0 + step (device-side numbering of a
    // captured batch); null: the step is `step`
    const unsigned long long *step0;
    // the anchor grid as the step kernels read it: a device copy of anc_grid,
    // rewritten at every neighbour rebuild (so captured batches stay valid)
    const Anchor *anc_dev;
    // skin criterion on the device: step k reads slot (k+2)%3 (displacement of
    // its input positions), writes slot k%3 (of its output) and clears slot
    // (k+1)%3 for step k+1.  A step whose input violates the skin, or that
    // finds *stop set, does nothing and sets *stop, so the host can launch a
    // batch of steps and rebuild once after it.
    unsigned *dslot;           // 3 flags: some particle moved more than skin/2 since build
    float half_skin2;          // (skin/2)^2
    const float *bud2;         // squared skin budget of isolated particles (k_budget; 0: skin/2), or null
    int *stop;                 // sticky: a step was skipped
    int *done;                 // steps executed (incremented by block 0)
    int nh_gate;               // Nose-Hoover second half in a batch: skip on *stop or dslot[step % 3]
    unsigned *bad;             // nonzero if a non-finite value appeared
    int swizzle;               // XCD-aware block remap
};

// Per-particle accumulators for the standalone force / bond / integrate API.
struct ForceArgs {
    long long n, stride;
    const double *x, *nrm;
    const PQ *pq;
    const uint8_t *type;
    const uint32_t *id;
    const uint32_t *nbr, *nnbr;
    const uint32_t *bnd;
    const uint8_t *nbnd;
    Anchor anc_grid;
    float *f32, *t32;
    double *f64, *t64;
    double *e, *w;
    int swizzle;
};

// ---------------------------------------------------------------------------
// launchers (each .hip file owns its kernels)
void upload_ff(const FFTab<double> &d, const FFTab<float> &f, hipStream_t s);

// k_sort.hip
size_t scan_temp_bytes(long long n);
void exclusive_scan_u32(const uint32_t *in, uint32_t *out, long long n, void *temp, hipStream_t s, Gate gt = {});
size_t radix_temp_bytes(long long n);
// allocation-free sort (k_sort.hip): the device address of this device's
// fixed scratch (looked up once per device, with it current) and the sort of
// caller-owned device arrays through that scratch
constexpr long long NOALLOC_SORT_MAX = 1LL << 20;
void *scratch_address(int device);
void sort_pairs_in_scratch(uint32_t *keys, uint32_t *vals, long long n, int bits, void *scr, hipStream_t s);
void radix_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t *keys_alt, uint32_t *vals_alt, long long n,
                      int bits, void *temp, hipStream_t s, bool *result_in_alt, Gate gt = {});

// k_build.hip
void launch_bbox(const double *x, long long n, double *out6, hipStream_t s);
// periodic axes: x -> lo + (x - lo) - L floor((x - lo) / L)
void launch_wrap(double *x, long long n, const double lo[3], const double L[3], const int per[3], hipStream_t s);
void launch_keys(const double *x, long long n, Grid g, uint32_t *keys, uint32_t *vals, hipStream_t s);
struct PermuteArgs {
    long long n;
    const uint32_t *perm;
    const double *src3[6]; double *dst3[6]; int n3;     // AoS double3 arrays
    const float *srcf3[2]; float *dstf3[2]; int nf3;    // AoS float3 arrays
    const uint8_t *type_src; uint8_t *type_dst;
    const uint32_t *id_src; uint32_t *id_dst;
    const uint32_t *slot_src; uint32_t *slot_dst;
    const int32_t *aff_src; int32_t *aff_dst;
    const uint16_t *grp_src; uint16_t *grp_dst;         // ensemble member (or null)
};
void launch_permute(const PermuteArgs &a, hipStream_t s);
// removal of the particles with del[i] != 0: perm[0..n - ndel) lists the kept
// particles in store order; lipid slots (< pslot0) are renumbered without
// gaps (flag, slot_pre, keep, pos: n-entry scratch)
void launch_compact(long long n, const uint32_t *del, uint32_t *slot, uint32_t pslot0, uint32_t *flag,
                    uint32_t *slot_pre, uint32_t *keep, uint32_t *pos, uint32_t *perm, void *scan_temp,
                    hipStream_t s);
void launch_inverse(const uint32_t *slot, uint32_t *inv, long long n, hipStream_t s);
void launch_refresh(long long n, const double *x, const double *nrm, const uint8_t *type, Anchor a,
                    PQ *pq, float *rb, uint32_t *anc, hipStream_t s);
void launch_cell_ranges(const double *x, long long n, Grid g, uint32_t *cstart, uint32_t *cend, hipStream_t s);
struct NlistArgs {
    long long n, stride;
    const double *x;
    const PQ *pq;            // FP32 mirrors with anchors (refreshed just before)
    Anchor anc;
    const uint8_t *type;
    Grid g;
    Box box;
    const uint32_t *cstart, *cend;
    uint32_t *nbr, *nnbr;
    int kmax;
    unsigned *max_count;
    double skin;
    const int32_t *aff;      // Voronoi filter (null in exact mode)
    const float *centroids;
};
long long launch_cell_list(const double *x, long long n, Grid g, uint32_t *flag, uint32_t *pos, uint32_t *occ,
                           void *scan_temp, hipStream_t s);
// far: some type reaches past the first ring of cells (outer shells appended by a second kernel,
// over the compacted list of those particles; FarScratch: n-sized flag, pos, list, a device count
// and the scan's temporary)
struct FarScratch {
    uint32_t *flag, *pos, *list, *count;
    void *scan_temp;
};
void launch_nlist(const NlistArgs &a, const uint32_t *occ, long long nocc, bool far, const FarScratch &fs,
                  hipStream_t s);
void set_list_cutoffs(const float *cut2_list, const int *rings, const float *pair_cut, const int *budget_rings,
                      hipStream_t s);
// skin budget of the particles without list entries (k_build.hip, k_budget)
// (fs: scratch as for the outer-shell pass of launch_nlist)
void launch_budget(long long n, const double *x, const uint8_t *type, const uint32_t *nnbr, const Grid &g,
                   const uint32_t *cstart, const uint32_t *cend, float half_skin, float *bud2, const FarScratch &fs,
                   hipStream_t s);
// bud2: skin budgets of the last build (nullptr: skin/2 for every particle)
void launch_disp(long long n, const double *x, const uint32_t *anc, const float *rb, Anchor g, const float *bud2,
                 const uint32_t *nnbr, unsigned *disp_max, hipStream_t s);
void launch_bondlist(long long n, long long stride, const uint32_t *slot, const uint32_t *inv,
                     const uint32_t *bstart, const uint32_t *bpartner, const uint8_t *btype,
                     uint32_t *bnd, uint8_t *nbnd, hipStream_t s);
// block halos from the 32-bit list: ext (HALO_EXT per block), cnt (per block),
// nbr16 (halo_words(kmax) words per particle, column stride `stride`)
void launch_halo(long long n, long long stride, const uint32_t *nbr, const uint32_t *nnbr, const uint8_t *type,
                 uint32_t *ext, uint32_t *cnt, uint32_t *nbr16, uint8_t *hord, int cap, hipStream_t s);

// k_force.hip
void launch_forces(const ForceArgs &a, bool mixed, bool energy, bool pair, bool bond, hipStream_t s);
void launch_step_langevin(const StepArgs &a, bool mixed, hipStream_t s);
// forces + the NH second half in one pass (k_forces + post_toque_final_update);
// f, n x t go to f32/t32 (mixed) or f64/t64, KE partials per particle block
void launch_step_nh_second(const StepArgs &a, bool mixed, float *f32, float *t32, double *f64, double *t64,
                           double *ke_partial, const double *zeta, hipStream_t s);
long long step_blocks(long long n);
void launch_pairs_export(const ForceArgs &a, bool mixed, unsigned *ids_i, unsigned *ids_j,
                         unsigned long long *count, long long cap, hipStream_t s);

// k_integrate.hip
struct IntegArgs {
    long long n;
    double *x, *v, *nrm, *o;
    float *f32, *t32;
    double *f64, *t64;
    const uint8_t *type;
    const uint32_t *id;
    const uint32_t *anc;
    PQ *pq;            // mirrors refreshed in place when nonnull
    Anchor anc_grid;
    double dt;
    double gamma[NT], sigma[NT];
    const uint16_t *grp;      // ensemble member of each particle, or null (see StepArgs)
    const double *sig_group;
    unsigned long long seed, step;
    double box_lo[3], box_hi[3];
    const double *zeta;       // device scalar
    double *ke_partial;       // per block
    double dr_opt, dn_opt;
    // NH first half: when disp_max is set, the largest squared displacement of
    // the new positions from the build-time offsets rb (as k_disp computes it)
    // is folded into *disp_max (float bits, atomicMax)
    const float *rb;
    unsigned *disp_max;
    // batched Nose-Hoover run (orbc_run): stop set -> the kernel does nothing;
    // skin_slot set -> the first half ors 1 into it where a particle moved
    // past its budget (skin/2, or bud2 for particles without list entries) and
    // clears skin_clear for the step after
    const int *stop;
    unsigned *skin_slot, *skin_clear;
    const float *bud2;
    const uint32_t *nnbr;
    float half_skin2;
};
void launch_integrate(int kernel, const IntegArgs &a, bool mixed, hipStream_t s);
void launch_minimize(const IntegArgs &a, bool mixed, hipStream_t s);
void launch_reduce_ke(const double *v, const uint8_t *type, long long n, double *partial, hipStream_t s);
void launch_sum_partials(const double *partial, long long nblocks, double *out, hipStream_t s);
void launch_zeta_update(const double *partial, long long nblocks, double *zeta, double dt, double Q,
                        double nkT15, hipStream_t s, const int *stop = nullptr);
void launch_sum_array(const double *a, long long n, double *partial, hipStream_t s);
long long reduce_blocks(long long n);
void launch_check_finite(const double *x, long long n3, unsigned *bad, hipStream_t s);
void launch_scatter_slots(long long n, const uint32_t *slot, const double *src3, double *dst3, hipStream_t s);
void launch_scatter_slots_f(long long n, const uint32_t *slot, const float *src3, double *dst3, hipStream_t s);
void launch_scatter_slots1(long long n, const uint32_t *slot, const double *src, double *dst, hipStream_t s);
void launch_gather_slots(long long n, const uint32_t *slot, const double *src3, double *dst3, hipStream_t s);
void launch_gather_slots_f(long long n, const uint32_t *slot, const double *src3, float *dst3, hipStream_t s);
// single device counters of the batched run: *p = v, *p += v
void launch_set_u64(unsigned long long *p, unsigned long long v, hipStream_t s);
void launch_add_u64(unsigned long long *p, unsigned long long v, hipStream_t s);
void launch_fill_u32(uint32_t *p, uint32_t v, long long n, hipStream_t s, Gate gt = {});
void launch_gather_u16(long long n, const uint32_t *slot, const uint16_t *src, uint16_t *dst, hipStream_t s);

// k_voronoi.hip
struct VoroArgs {
    long long n;            // particles
    long long ncell;
    const double *x;
    const double *nrm;
    const uint8_t *type;
    int32_t *aff;
    float *centroids;
    double mass[NT];
    // centroid grid (for nearest / within queries)
    Grid cg;
    const uint32_t *cg_start, *cg_end, *cg_idx;
    Gate gate;               // inside a batched run: the step this work precedes may be skipped
    const Grid *cg_dev;      // the centroid grid as the kernels read it: a device copy of cg, rewritten
                             // whenever the grid is rebuilt (so captured batches stay valid)
    // The centroid grid is built at neighbour rebuilds only: cref holds the
    // centroids as binned, *cdisp the largest squared drift of a centroid since
    // (as float bits).  Queries widen their reach by that drift, so they stay
    // exact.  cdisp null: the index is current.
    float *cref;
    unsigned *cdisp;
    // lipid count per cell of the current affiliation, written by
    // k_volume_finish; csize[ncell] != 0 says they are current (a partition
    // clears it).  A block of k_cell_sums holding every member of a cell
    // stores that cell's sums instead of adding them atomically.  Null: none.
    uint32_t *csize;
};
void launch_morton_ref(const float *pts, long long n, double bsize, uint32_t *keys, hipStream_t s);
void launch_centroid_keys(const float *c, long long ncell, Grid g, uint32_t *keys, uint32_t *vals, hipStream_t s,
                          Gate gt = {});
void launch_cgrid_ranges(const uint32_t *sorted_keys, long long ncell, uint32_t *start, uint32_t *end,
                         hipStream_t s, Gate gt = {});
// stray-lipid cleanup (src/cleanup.h:29-91)
void launch_stray_keys(const VoroArgs &a, uint32_t *keys, uint32_t *vals, double *d2, hipStream_t s);
void launch_stray_mark(long long ncell, const uint32_t *start, const uint32_t *end, const uint32_t *svals,
                       const double *d2, double tol2, uint32_t *del, unsigned *count, hipStream_t s);
void launch_nearest(const VoroArgs &a, const float *q, long long nq, const double *qd, int32_t *out, bool warm,
                    hipStream_t s);
void launch_within(const VoroArgs &a, const float *q, long long nq, float rmax, int cap, int *counts,
                   int *out, hipStream_t s);
void launch_centroid_sums(const VoroArgs &a, unsigned long long *sums, hipStream_t s);
void launch_centroid_finish(const VoroArgs &a, unsigned long long *sums, hipStream_t s);
void launch_centroid_center(const VoroArgs &a, double *partial, hipStream_t s);
void launch_volume(const VoroArgs &a, unsigned long long *sums, double *normals, double *partial, double *scal,
                   double target, double strength, hipStream_t s);
void launch_volume_push(const VoroArgs &a, const double *normals, const double *fpush, double *f_ext,
                        double *f64, float *f32, hipStream_t s);

}  // namespace orbc

#define ORBC_HIP_CHECK(expr)                                                     \
    do {                                                                         \
        hipError_t e_ = (expr);                                                  \
        if (e_ != hipSuccess) throw orbc::HipError(e_, #expr, __FILE__, __LINE__); \
    } while (0)

namespace orbc {
struct HipError {
    hipError_t err;
    const char *expr, *file;
    int line;
    HipError(hipError_t e, const char *x, const char *f, int l) : err(e), expr(x), file(f), line(l) {}
};
}  // namespace orbc

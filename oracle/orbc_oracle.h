/*
 * orbc_oracle.h -- CPU restatement of OpenRBC's per-timestep MD loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (openrbc_amd, liborbc_hip.so) never links or calls it.
 *
 * Every function states the reference file:line it restates (paths relative
 * to the OpenRBC source tree).  Arithmetic is FP64 unless a function name
 * says otherwise; the *_mixed functions restate the GPU's FP32
 * anchor-relative geometry bit for bit so that pair sets can be compared
 * exactly.
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   orc_pair_*            : closed-form energy, finite-difference gradient,
 *                           known answers (F=0 at req), Newton antisymmetry,
 *                           reference golden vectors (tests/golden).
 *   orc_pairs_*           : brute force O(N^2) vs cell list, both ways.
 *   orc_bonded            : Hooke closed form, FD gradient, golden vectors.
 *   orc_langevin          : single-particle closed form, golden vectors at kT=0.
 *   orc_nh_*              : golden vectors, closed form.
 *   orc_philox4x32_10     : Random123 known-answer vectors + rocRAND.
 *   orc_morton_ref        : brute-force bit interleave, golden vectors.
 *   orc_partition/stencil : brute force by definition, golden vectors.
 *   orc_volume            : sphere closed form.
 *   orc_stray_delete      : hand-computed medians/thresholds (even and odd
 *                           cells, ties at the threshold), reference golden
 *                           vectors from delete_lipid itself.
 */
#ifndef ORBC_ORACLE_H_
#define ORBC_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- force field (src/forcefield_canonical.h:30-156) -------------------- */
typedef struct {
    /* 6x6 tables indexed [a*6+b]; symmetric */
    double cut_poly[36], cut2_poly[36], rep[36], att[36], alpha[36];
    int    aniso[36];                   /* 1: anisotropic 4-8 (a lipid is involved) */
    double cut2_lj[36], lj1[36], lj2[36];
    double mass[6], radius[6];
    double bond_r0[4], bond_K[4];
} orc_ff;

const orc_ff *orc_ff_get(void);

/* one directed pair: owner i, neighbour j, dx = x_i - x_j.
 * Adds the force on i to fi, the director force (-dU/dn_i) to ti, returns
 * 1 if the pair interacts.  u = pair energy (full), w = dx . f_i.
 * src/pairwise_kernel.h:30-68, src/pairwise_kernel_fused.h:22-97,
 * selection src/compute_pairwise_fused.h:210,235,273-276,302-305,330-333 */
int orc_pair(int ti, int tj, const double dx[3], const double ni[3], const double nj[3],
             double fi[3], double tq[3], double *u, double *w);

/* ---- periodic box ---------------------------------------------------------
 * The reference has no periodic boundaries (its box only bounces particles
 * back, src/integrate_nh.h:124-144).  The planar-patch workload needs them;
 * this restates the build's definition: on a periodic axis d every pair
 * vector takes the minimum image dx - L*rint(dx/L), and positions are wrapped
 * into [lo, lo+L) by x -> lo + (x-lo) - L*floor((x-lo)/L).  All FP64
 * functions below honour the box set here (default: no periodic axis). */
void orc_set_box(const double lo[3], const double L[3], const int per[3]);
void orc_wrap(int64_t n, double *x);

/* ---- pair sets ----------------------------------------------------------- */
/* Unordered interacting pairs (i<j), FP64 selection 1e-5 < r2 < rc2(ti,tj).
 * Brute force O(N^2).  Returns count; writes up to cap pairs into out[2*k]. */
int64_t orc_pairs_brute_fp64(int64_t n, const double *x, const int32_t *type,
                             int64_t cap, int64_t *out);
/* Same set through a uniform cell list (O(N)). */
int64_t orc_pairs_cell_fp64(int64_t n, const double *x, const int32_t *type,
                            int64_t cap, int64_t *out);
/* Mixed geometry: positions as FP32 offsets from an integer anchor grid of
 * spacing h (power of two): A = floor(x/h), r32 = (float)(x - h*A),
 * dx = (r32_i - r32_j) + h*(A_i - A_j) in FP32, r2 = fma(dz,dz,fma(dy,dy,dx*dx)).
 * Anchors are taken from xa (positions at the last neighbour build). */
/* Mixed selection with the general anchor scheme of the device, per axis d:
 * A = floor((xa - lo_d) / h_d) - A0_d, clamped to [0, 2^bits_d - 1] on
 * periodic axes, r = (float)(x - (lo_d + h_d (A + A0_d))), and
 * dx = fma((float)h_d, (float)dA, fl(r_i - r_j)) (one rounding) where on a periodic axis
 * dA = A_i - A_j sign-extended from bits_d (wrap to the nearest image).
 * Open axes: lo_d = 0, h_d a power of two, bits_d = 0 (no wrap). */
int64_t orc_pairs_brute_mixed_anchor(int64_t n, const double *x, const double *xa, const int32_t *type,
                                     const double lo[3], const double h[3], const double A0[3],
                                     const int bits[3], int64_t cap, int64_t *out);
int64_t orc_pairs_brute_mixed(int64_t n, const double *x, const double *xa, const int32_t *type,
                              double h, int64_t cap, int64_t *out);

/* ---- forces -------------------------------------------------------------- */
/* f,t accumulate (caller clears), e/w per particle (may be NULL).
 * Pairs: unordered list (i<j).  Newton's third law applied per pair. */
void orc_forces_pairs(int64_t n, const double *x, const double *nrm, const int32_t *type,
                      int64_t npairs, const int64_t *pairs,
                      double *f, double *t, double *e, double *w);
/* Bonds by index: bond k joins bi[k] and bj[k] with type bt[k].
 * src/compute_bonded.h:89-146 (F = K(1 - r0/r)(x_j - x_i) on i). */
void orc_bonded(int64_t nb, const int32_t *bt, const int64_t *bi, const int64_t *bj,
                const double *x, double *f, double *e, double *w);

/* ---- integrators (FP64) -------------------------------------------------- */
/* Philox-4x32-10 (Salmon et al., SC'11). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Uniform [-1,1) noise for particle id at step, as the GPU draws it. */
void orc_langevin_noise(uint64_t seed, uint32_t id, uint64_t step, double r[3]);
/* src/integrate_langevin.h:99-149 (verlet_langevin); noise from Philox. */
void orc_langevin(int64_t n, double *x, double *v, double *nrm, double *o,
                  double *f, double *t, const int32_t *type, const uint32_t *id,
                  double kT, double eta, double dt, uint64_t seed, uint64_t step);
/* src/integrate_nh.h:178-235 (verlet_initial_bounce_clearforce_update).
 * Returns kinetic energy sum 0.5 m v^2 after the update. */
double orc_nh_first(int64_t n, double *x, double *v, double *nrm, double *o,
                    double *f, double *t, const int32_t *type,
                    double dt, double zeta, const double box_lo[3], const double box_hi[3]);
/* src/integrate_nh.h:237-273 (post_toque_final_update).  Returns KE. */
double orc_nh_second(int64_t n, const double *x, double *v, const double *nrm, double *o,
                     const double *f, double *t, const int32_t *type, double dt, double zeta);
/* zeta update of the destructors src/integrate_nh.h:182-185,241-244 */
double orc_nh_zeta(double zeta, double dt, double Q, double ke, int64_t n, double kT);
/* src/compute_temperature.h:23-29 */
double orc_temperature(int64_t n, const double *v, const int32_t *type);
/* src/openrbc.cpp:114-131: capped steepest-descent move (after post_torque). */
void orc_minimize_move(int64_t n, double *x, double *nrm, const double *f, const double *t,
                       const int32_t *type, double dr_opt, double dn_opt, double dt);

/* ---- Voronoi cells ------------------------------------------------------- */
/* Morton key of the reference, src/reorder_morton.h:25-42 (11 usable bits). */
uint32_t orc_morton_ref(float x, float y, float z, double bsize);
/* Nearest centroid per point by brute force over FP32 squared distances;
 * ties go to the lowest cell index.  src/voronoi.h:179-237 */
void orc_partition(int64_t n, const double *x, int64_t ncell, const float *c, int32_t *aff);
/* Mean of member positions per cell, FP64 sum, rounded to FP32.
 * src/voronoi.h:123-140 */
void orc_centroids(int64_t n, const double *x, const int32_t *aff, int64_t ncell, float *c);
/* Cells within rmax of cell `cell` (strict, FP32 normsq), ascending.
 * src/voronoi.h:105-117, src/kdtree.h:263-285 */
int64_t orc_stencil(int64_t ncell, const float *c, int64_t cell, float rmax, int64_t cap, int64_t *out);
/* Unordered pairs of the reference's Voronoi-stencil sweep:
 * src/compute_pairwise_fused.h:253-316 (stencil radii 9/8/6 for pp/pl/ll). */
int64_t orc_pairs_voronoi_fp64(int64_t n, const double *x, const int32_t *type,
                               const int32_t *aff, int64_t ncell, const float *c,
                               int64_t cap, int64_t *out);
/* delete_lipid, src/cleanup.h:29-60: per cell the squared distances of its
 * lipids to the centroid (FP64 against the FP32 centroid) are sorted and a
 * lipid is deleted (del = 1) unless d2 < tol^2 * dr2[size / 2].  Proteins and
 * lipids with aff < 0 are kept.  Returns the number deleted. */
int64_t orc_stray_delete(int64_t n, const double *x, const int32_t *type, const int32_t *aff, int64_t ncell,
                         const float *c, double tol, uint8_t *del);
/* Cone-sum volume + push, fixed mode (normals reset, type[j]).
 * src/constrain_volume.h:26-83.  Adds to f, returns volume. */
double orc_constrain_volume(int64_t n, const double *nrm, const int32_t *type,
                            const int32_t *aff, int64_t ncell, const float *c,
                            double target, double strength, double *f);

#ifdef __cplusplus
}
#endif
#endif

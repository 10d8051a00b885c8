// k_force.hip -- pair and bond forces, and the fused timestep kernel, gfx950.
//
// One lane owns one particle i (the reference's owner-computes rule across
// thread chunks, src/compute_pairwise_fused.h:249-316, applied to every
// pair): it walks its full neighbour list (ELL, column-major so that the k-th
// entries of a wavefront's 64 particles are one coalesced 256-byte load),
// accumulates f_i and the director force t_i in registers and writes them
// once.  No atomics, fixed summation order: results are deterministic.
//
// Pair terms (src/pairwise_kernel.h:30-68, src/pairwise_kernel_fused.h:22-97):
//   anisotropic 4-8 polynomial when a lipid is involved, isotropic repulsive
//   polynomial between proteins, else WCA Lennard-Jones; selection
//   1e-5 < r^2 < rc^2 with the polynomial tested first
//   (src/compute_pairwise_fused.h:210,273-276,302-305).
// Bonds (src/compute_bonded.h:108-135): F_i = K (1 - r0/r) (x_j - x_i), each
//   end accumulated by its owner from a per-particle bond list.
//
// Precision: MIXED evaluates every pair and bond in FP32 from FP32 mirrors
// whose positions are offsets from an integer anchor grid, so dx is exact to
// FP32 rounding of O(1) numbers; FP64 evaluates everything in FP64.  The
// integrator update and all reductions are FP64 in both.
#include <type_traits>

#include "orbc_internal.h"

namespace orbc {

void upload_mass(const double *mass, hipStream_t s);   // k_integrate.hip

namespace {

__constant__ FFTab<float> c_fff;
__constant__ FFTab<double> c_ffd;

template <typename R> __device__ inline const FFTab<R> &ff();
template <> __device__ inline const FFTab<float> &ff<float>() { return c_fff; }
template <> __device__ inline const FFTab<double> &ff<double>() { return c_ffd; }

template <typename R>
struct Acc {
    R fx = 0, fy = 0, fz = 0, tx = 0, ty = 0, tz = 0, e = 0, w = 0;
};

__device__ inline float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ inline double rsq(double x) { return rsqrt(x); }

// exact-rounding helpers: the selection test must round exactly as the oracle
__device__ inline float sub_rn(float a, float b) { return __fsub_rn(a, b); }
__device__ inline double sub_rn(double a, double b) { return __dsub_rn(a, b); }
__device__ inline float add_rn(float a, float b) { return __fadd_rn(a, b); }
__device__ inline double add_rn(double a, double b) { return __dadd_rn(a, b); }
__device__ inline float r2_rn(float x, float y, float z) { return __fmaf_rn(z, z, __fmaf_rn(y, y, __fmul_rn(x, x))); }
__device__ inline double r2_rn(double x, double y, double z) { return __fma_rn(z, z, __fma_rn(y, y, __dmul_rn(x, x))); }

// Force-field table staged in LDS by every force kernel: all pair types go
// through the same code path (lipid-lipid and lipid-protein are both the
// anisotropic form with different parameters), so a wave does not diverge on
// the pair type except for the rare LJ / protein-protein terms.
// The FP64 copies cut_d / rep_d / att_d serve the radial scalars of the
// mixed anisotropic term (see aniso_term).
template <typename R>
struct FFS {
    R cut2p[36], cutp[36], rep[36], att[36], alpha[36], cut2lj[36], lj1[36], lj2[36];
    int aniso[36];
    R K[NBT], r0[NBT];
    double cut_d[36], rep_d[36], att_d[36];
};

template <typename R>
__device__ inline void load_ffs(FFS<R> *s) {
    const FFTab<R> &t = ff<R>();
    for (int k = threadIdx.x; k < 36; k += blockDim.x) {
        s->cut2p[k] = t.cut2_poly[k]; s->cutp[k] = t.cut_poly[k]; s->rep[k] = t.rep[k]; s->att[k] = t.att[k];
        s->alpha[k] = t.alpha[k]; s->cut2lj[k] = t.cut2_lj[k]; s->lj1[k] = t.lj1[k]; s->lj2[k] = t.lj2[k];
        s->aniso[k] = t.aniso[k];
        s->cut_d[k] = c_ffd.cut_poly[k]; s->rep_d[k] = c_ffd.rep[k]; s->att_d[k] = c_ffd.att[k];
    }
    if (threadIdx.x < NBT) { s->K[threadIdx.x] = t.bond_K[threadIdx.x]; s->r0[threadIdx.x] = t.bond_r0[threadIdx.x]; }
    __syncthreads();
}

// |d| from r2 = |d|^2: rinv by the hardware reciprocal square root, then one
// Newton step on r so that cut - r (raised to the 7th power in the repulsion)
// carries the rounding of r2 only.
__device__ inline void radius(float r2, float &r, float &rinv) {
    rinv = rsq(r2);
    r = r2 * rinv;
    r = __fmaf_rn(0.5f * __fmaf_rn(-r, r, r2), rinv, r);
}
__device__ inline void radius(double r2, double &r, double &rinv) {
    rinv = rsq(r2);
    r = r2 * rinv;
}

// Anisotropic polynomial term of one directed pair inside its cutoff
// (src/pairwise_kernel.h:37-62): force on i, director force on i, energy and
// virial halves.
//
// Mixed-path arithmetic (ORBC_PAIR_F64): the radial factor carries
// (cut - r)^7 and A = 1 + alpha (a - 1) with alpha up to 5, so FP32 rounding
// of r, cut or a is amplified into relative errors near 1e-6 on the stiff
// lipid-protein terms.
//   0: all FP32;
//   1: r, s = cut - r, A and the radial scalars in FP64, vectors in FP32;
//   2: the whole term in FP64 from the exact separation of the FP32 mirrors
//      (X = (r_i - r_j) + h dA, formed in FP64 without rounding) and the
//      FP32 directors, summed in FP64 (FP64 and FP32 FMA issue at the same
//      rate on this GPU).  Rounding dx to FP32 would be the largest error.
// The selection test stays FP32 in every mode.
#ifndef ORBC_PAIR_F64
#define ORBC_PAIR_F64 2
#endif
// mode 2: the term in FP64 from the exact separation X, Y, Z (cut, rep, att
// the FP64 constants, alpha the FP32 one, r2 the FP32 squared distance of the
// selection test)
template <bool ENERGY, typename AR>
__device__ __forceinline__ void aniso_fp64(double cut, double rep, double att, float alpha_f, float r2, double X,
                                           double Y, double Z, float nix, float niy, float niz, float njx, float njy,
                                           float njz, Acc<AR> &acc) {
    const double d2 = __fma_rn(Z, Z, __fma_rn(Y, Y, X * X));
    double y = (double)rsq(r2);                               // 1/r to FP32 accuracy
    y = y * __fma_rn(-0.5 * d2, y * y, 1.5);                  // one Newton step: FP64 accuracy
    const double rd = d2 * y;
    const double ex = X * y, ey = Y * y, ez = Z * y;
    const double ax = nix, ay = niy, az = niz, bx = njx, by = njy, bz = njz;
    const double ninj = __fma_rn(az, bz, __fma_rn(ay, by, ax * bx));
    const double nie = __fma_rn(az, ez, __fma_rn(ay, ey, ax * ex));
    const double nje = __fma_rn(bz, ez, __fma_rn(by, ey, bx * ex));
    const double a = __fma_rn(-nie, nje, ninj);
    const double pix = __fma_rn(-nie, ex, ax), piy = __fma_rn(-nie, ey, ay), piz = __fma_rn(-nie, ez, az);
    const double pjx = __fma_rn(-nje, ex, bx), pjy = __fma_rn(-nje, ey, by), pjz = __fma_rn(-nje, ez, bz);
    const double s = cut - rd;
    const double s2 = s * s, s3 = s2 * s, s4 = s2 * s2;
    const double alpha = alpha_f;
    const double A = __fma_rn(alpha, a - 1.0, 1.0);
    const double aua = alpha * (att * s4);
    const double auar = aua * y;
    const double fra = s3 * __fma_rn(4.0 * A, att, 8.0 * rep * s4);
    const double fx = __fma_rn(fra, ex, auar * __fma_rn(nje, pix, nie * pjx));
    const double fy = __fma_rn(fra, ey, auar * __fma_rn(nje, piy, nie * pjy));
    const double fz = __fma_rn(fra, ez, auar * __fma_rn(nje, piz, nie * pjz));
    acc.tx -= (AR)(aua * pjx);
    acc.ty -= (AR)(aua * pjy);
    acc.tz -= (AR)(aua * pjz);
    acc.fx += (AR)fx; acc.fy += (AR)fy; acc.fz += (AR)fz;
    if (ENERGY) {
        acc.e += (AR)(0.5 * s4 * __fma_rn(A, att, rep * s4));
        acc.w += (AR)(0.5 * __fma_rn(Z, fz, __fma_rn(Y, fy, X * fx)));
    }
}

template <typename R, bool ENERGY, typename AR>
__device__ __forceinline__ void aniso_term(const FFS<R> &t, int k, R r2, R dx, R dy, R dz, double X, double Y,
                                           double Z, R nix, R niy, R niz, R njx, R njy, R njz, Acc<AR> &acc) {
    if constexpr (std::is_same<R, float>::value && ORBC_PAIR_F64 >= 2) {
        aniso_fp64<ENERGY>(t.cut_d[k], t.rep_d[k], t.att_d[k], t.alpha[k], r2, X, Y, Z, nix, niy, niz, njx, njy,
                           njz, acc);
        return;
    }
    {
        R r, rinv;
        radius(r2, r, rinv);
        const R alpha = t.alpha[k];
        const R ex = dx * rinv, ey = dy * rinv, ez = dz * rinv;
        const R ninj = nix * njx + niy * njy + niz * njz;
        const R nie = nix * ex + niy * ey + niz * ez;
        const R nje = njx * ex + njy * ey + njz * ez;
        const R a = ninj - nie * nje;
        const R pix = nix - nie * ex, piy = niy - nie * ey, piz = niz - nie * ez;
        const R pjx = njx - nje * ex, pjy = njy - nje * ey, pjz = njz - nje * ez;
        R fra, aua, auar, en = R(0);
        if constexpr (std::is_same<R, float>::value && ORBC_PAIR_F64 == 1) {
            const double X = dx, Y = dy, Z = dz;
            const double d2 = __fma_rn(Z, Z, __fma_rn(Y, Y, X * X));
            double y = (double)rinv;
            y = y * __fma_rn(-0.5 * d2, y * y, 1.5);
            const double rd = d2 * y;
            const double s = t.cut_d[k] - rd;
            const double s2 = s * s, s3 = s2 * s, s4 = s2 * s2;
            const double A = __fma_rn((double)alpha, (double)a - 1.0, 1.0);
            const double rep = t.rep_d[k], att = t.att_d[k];
            const double aua_d = (double)alpha * (att * s4);
            fra = (R)(s3 * __fma_rn(4.0 * A, att, 8.0 * rep * s4));
            aua = (R)aua_d;
            auar = (R)(aua_d * y);
            if (ENERGY) en = (R)(0.5 * s4 * __fma_rn(A, att, rep * s4));
        } else {
            const R s = t.cutp[k] - r;
            const R s2 = s * s, s3 = s2 * s, s4 = s2 * s2;
            const R A = R(1) + alpha * (a - R(1));
            const R rep = t.rep[k], att = t.att[k];
            aua = alpha * (att * s4);
            auar = aua * rinv;
            fra = s3 * (R(8) * rep * s4 + R(4) * A * att);
            if (ENERGY) en = R(0.5) * s4 * (rep * s4 + A * att);
        }
        const R fx = fra * ex + auar * (nje * pix + nie * pjx);
        const R fy = fra * ey + auar * (nje * piy + nie * pjy);
        const R fz = fra * ez + auar * (nje * piz + nie * pjz);
        acc.tx -= (AR)(aua * pjx);
        acc.ty -= (AR)(aua * pjy);
        acc.tz -= (AR)(aua * pjz);
        acc.fx += (AR)fx; acc.fy += (AR)fy; acc.fz += (AR)fz;
        if (ENERGY) {
            acc.e += (AR)en;
            acc.w += (AR)(R(0.5) * (dx * fx + dy * fy + dz * fz));
        }
    }
}

// One directed pair (owner i, neighbour j, d = x_i - x_j).  Returns true if
// the pair interacts (same selection test as the oracle, bit for bit).
// X, Y, Z: the separation in FP64 (the mixed path's exact mirror separation;
// equal to dx, dy, dz in the FP64 path).
// Repulsive protein-protein polynomial, src/pairwise_kernel_fused.h:79-97, and
// WCA Lennard-Jones, :63-77, for the packed-record path (pair_rec); the same
// arithmetic as the branches of pair_ij, the forces rounded before they are
// summed.
template <typename R, bool ENERGY, typename AR>
__device__ __forceinline__ void pp_term(R cutp, R rep, R r2, R dx, R dy, R dz, Acc<AR> &acc) {
    R r, rinv;
    radius(r2, r, rinv);
    const R ex = dx * rinv, ey = dy * rinv, ez = dz * rinv;
    const R s = cutp - r;
    const R s3 = s * s * s, s4 = (s * s) * (s * s);
    const R fr = R(8) * rep * s3 * s4;
    const R fx = fr * ex, fy = fr * ey, fz = fr * ez;
    if (ENERGY) acc.e += (AR)(R(0.5) * rep * s4 * s4);
    // rounded products, then the sums (no fused multiply-add across them)
    acc.fx = add_rn(acc.fx, (AR)fx); acc.fy = add_rn(acc.fy, (AR)fy); acc.fz = add_rn(acc.fz, (AR)fz);
    if (ENERGY) acc.w += (AR)(R(0.5) * (dx * fx + dy * fy + dz * fz));
}

template <typename R, bool ENERGY, typename AR>
__device__ __forceinline__ void lj_term(R lj1, R lj2, R r2, R dx, R dy, R dz, Acc<AR> &acc) {
    const R r2inv = R(1) / r2;
    const R r6inv = r2inv * r2inv * r2inv;
    const R fpair = r6inv * (lj1 * r6inv - lj2) * r2inv;
    const R fx = dx * fpair, fy = dy * fpair, fz = dz * fpair;
    if (ENERGY) acc.e += (AR)(R(0.5) * r6inv * (lj1 * (R(1) / R(12)) * r6inv - lj2 * (R(1) / R(6))));
    acc.fx = add_rn(acc.fx, (AR)fx); acc.fy = add_rn(acc.fy, (AR)fy); acc.fz = add_rn(acc.fz, (AR)fz);
    if (ENERGY) acc.w += (AR)(R(0.5) * (dx * fx + dy * fy + dz * fz));
}

template <typename R, bool ENERGY, typename AR>
__device__ __forceinline__ bool pair_ij(const FFS<R> &t, int ti, int tj, R dx, R dy, R dz, double X, double Y,
                                        double Z, R r2, R nix, R niy, R niz, R njx, R njy, R njz, Acc<AR> &acc) {
    if (!(r2 > R(1e-5))) return false;
    const int k = ti * NT + tj;
    R fx, fy, fz;
    if (r2 < t.cut2p[k]) {
        if (t.aniso[k]) {
            aniso_term<R, ENERGY>(t, k, r2, dx, dy, dz, X, Y, Z, nix, niy, niz, njx, njy, njz, acc);
            return true;
        }
        // src/pairwise_kernel_fused.h:79-97 (pp_term)
        R r, rinv;
        radius(r2, r, rinv);
        const R ex = dx * rinv, ey = dy * rinv, ez = dz * rinv;
        const R s = t.cutp[k] - r;
        const R s3 = s * s * s, s4 = (s * s) * (s * s);
        const R rep = t.rep[k];
        const R fr = R(8) * rep * s3 * s4;
        fx = fr * ex; fy = fr * ey; fz = fr * ez;
        if (ENERGY) acc.e += (AR)(R(0.5) * rep * s4 * s4);
    } else {
        if (!(r2 < t.cut2lj[k])) return false;
        // src/pairwise_kernel_fused.h:63-77 (lj_term)
        const R r2inv = R(1) / r2;
        const R r6inv = r2inv * r2inv * r2inv;
        const R fpair = r6inv * (t.lj1[k] * r6inv - t.lj2[k]) * r2inv;
        fx = dx * fpair; fy = dy * fpair; fz = dz * fpair;
        if (ENERGY) acc.e += (AR)(R(0.5) * r6inv * (t.lj1[k] * (R(1) / R(12)) * r6inv - t.lj2[k] * (R(1) / R(6))));
    }
    acc.fx += (AR)fx; acc.fy += (AR)fy; acc.fz += (AR)fz;
    if (ENERGY) acc.w += (AR)(R(0.5) * (dx * fx + dy * fy + dz * fz));
    return true;
}

// Pair-type record of the block-halo step: everything one type pair needs in
// three 16-byte LDS words read together (FP32 values are the FP64 ones
// rounded, as in the FP32 tables, so the terms are those of pair_ij).
struct PairRec {
    float4 sel;        // cut2 of the polynomial, cut2 of the WCA, alpha, anisotropic flag (int bits)
    double2 poly;      // cut, rep (FP64)
    double att;        // FP64
    float lj1, lj2;
};
static_assert(sizeof(PairRec) == 48, "three 16-byte LDS words");

__device__ inline void load_pair_recs(PairRec *s) {
    for (int k = threadIdx.x; k < 36; k += blockDim.x) {
        PairRec r;
        r.sel = make_float4(c_fff.cut2_poly[k], c_fff.cut2_lj[k], c_fff.alpha[k], __int_as_float(c_fff.aniso[k]));
        r.poly = make_double2(c_ffd.cut_poly[k], c_ffd.rep[k]);
        r.att = c_ffd.att[k];
        r.lj1 = c_fff.lj1[k]; r.lj2 = c_fff.lj2[k];
        s[k] = r;
    }
}

// pair_mixed (mode 2) with the type pair's constants from a PairRec
__device__ __forceinline__ void pair_rec(const Anchor &ag, const AnchorI &ai, float4 xi, float4 ni, float4 xj,
                                         float4 nj, const PairRec &P, Acc<double> &acc) {
    const AnchorI aj = anchor_unpack(__float_as_uint(xj.w), ag);
    const float dx = anchor_add(ag, 0, sub_rn(xi.x, xj.x), ai.x - aj.x);
    const float dy = anchor_add(ag, 1, sub_rn(xi.y, xj.y), ai.y - aj.y);
    const float dz = anchor_add(ag, 2, sub_rn(xi.z, xj.z), ai.z - aj.z);
    const float r2 = r2_rn(dx, dy, dz);
    if (!(r2 > 1e-5f)) return;
    if (r2 < P.sel.x) {
        if (__float_as_int(P.sel.w)) {
            const double X = anchor_add_exact(ag, 0, xi.x, xj.x, ai.x - aj.x);
            const double Y = anchor_add_exact(ag, 1, xi.y, xj.y, ai.y - aj.y);
            const double Z = anchor_add_exact(ag, 2, xi.z, xj.z, ai.z - aj.z);
            aniso_fp64<false>(P.poly.x, P.poly.y, P.att, P.sel.z, r2, X, Y, Z, ni.x, ni.y, ni.z, nj.x, nj.y, nj.z,
                              acc);
        } else {
            pp_term<float, false>((float)P.poly.x, (float)P.poly.y, r2, dx, dy, dz, acc);
        }
        return;
    }
    if (!(r2 < P.sel.y)) return;
    lj_term<float, false>(P.lj1, P.lj2, r2, dx, dy, dz, acc);
}

template <typename R, bool ENERGY, typename TAB>
__device__ __forceinline__ void bond_ij(const TAB &t, int bt, R dx, R dy, R dz, Acc<R> &acc) {
    const R r2 = dx * dx + dy * dy + dz * dz;
    R r, rinv;
    radius(r2, r, rinv);
    const R K = t.K[bt], r0 = t.r0[bt];
    const R c = K * ((r - r0) * rinv);
    acc.fx += c * dx; acc.fy += c * dy; acc.fz += c * dz;
    if (ENERGY) {
        const R dr = r - r0;
        acc.e += R(0.25) * K * dr * dr;
        acc.w += R(-0.5) * c * r2;
    }
}

__device__ inline long long block_index(int swizzle) {
    const unsigned b = blockIdx.x, nb = gridDim.x;
    if (!swizzle) return b;
    // consecutive particle tiles on one XCD (blocks are dealt round-robin over
    // the 8 XCDs), so neighbour gathers hit that XCD's L2
    const unsigned q = nb >> 3, r = nb & 7u, xcd = b & 7u, k = b >> 3;
    return (xcd < r) ? (long long)xcd * (q + 1) + k : (long long)r * (q + 1) + (long long)(xcd - r) * q + k;
}

__device__ inline AnchorI unpack(uint32_t w, const Anchor &a) { return anchor_unpack(w, a); }

template <typename R>
__device__ inline void add_to(Acc<double> &o, const Acc<R> &a) {
    o.fx += (double)a.fx; o.fy += (double)a.fy; o.fz += (double)a.fz;
    o.tx += (double)a.tx; o.ty += (double)a.ty; o.tz += (double)a.tz;
    o.e += (double)a.e; o.w += (double)a.w;
}

// Tuning knobs (compile-time): neighbours gathered per group in the mixed
// and FP64 walks (memory-level parallelism per wave; measured on the full
// RBC: mixed 3 with the pipelined gathers below, 0.574 ms against 0.594 for 2
// and 0.613 for 4 unpipelined) and, if nonzero, the waves per SIMD the fused
// step kernel's register budget is capped for.
#ifndef ORBC_STEP_UNROLL
#define ORBC_STEP_UNROLL 3
#endif
#ifndef ORBC_FP64_UNROLL
#define ORBC_FP64_UNROLL 4
#endif
#ifndef ORBC_STEP_WAVES
#define ORBC_STEP_WAVES 4
#endif
constexpr int UNROLL = ORBC_STEP_UNROLL;
constexpr int UNROLL64 = ORBC_FP64_UNROLL;
#if ORBC_STEP_WAVES > 0
#define ORBC_STEP_ATTR __attribute__((amdgpu_waves_per_eu(ORBC_STEP_WAVES)))
#else
#define ORBC_STEP_ATTR
#endif
// waves per SIMD of the block-halo step kernel: its LDS (HALO_CAP mirrors)
// admits 5 blocks of 256 per CU at HALO_CAP = 896, and 96 registers hold it
// without spills
#ifndef ORBC_HALO_WAVES
#define ORBC_HALO_WAVES 5
#endif
#if ORBC_HALO_WAVES > 0
#define ORBC_HALO_ATTR __attribute__((amdgpu_waves_per_eu(ORBC_HALO_WAVES)))
#else
#define ORBC_HALO_ATTR
#endif

// ---- per-particle force accumulation ---------------------------------------
// The list is walked in groups of UNROLL entries, loading the next group's
// indices together with this group's gathers; a group is padded with the
// particle itself (r2 = 0, no contribution).

// Mixed: pair terms from the FP32 mirrors, summed in FP64 (ORBC_PAIR_F64 >= 2)
// or in FP32 over each group and in FP64 across groups.
using PairAcc = std::conditional<(ORBC_PAIR_F64 >= 2), double, float>::type;
// one gathered neighbour against owner i (mixed path)
template <bool ENERGY, typename AR>
__device__ __forceinline__ void pair_mixed(const Anchor &ag, const AnchorI &ai, float4 xi, float4 ni, int ti,
                                           float4 xj, float4 nj, const FFS<float> &t, Acc<AR> &acc) {
    const AnchorI aj = unpack(__float_as_uint(xj.w), ag);
    const float dx = anchor_add(ag, 0, sub_rn(xi.x, xj.x), ai.x - aj.x);
    const float dy = anchor_add(ag, 1, sub_rn(xi.y, xj.y), ai.y - aj.y);
    const float dz = anchor_add(ag, 2, sub_rn(xi.z, xj.z), ai.z - aj.z);
    const float r2 = r2_rn(dx, dy, dz);
    double X = 0.0, Y = 0.0, Z = 0.0;
    if constexpr (ORBC_PAIR_F64 >= 2) {
        X = anchor_add_exact(ag, 0, xi.x, xj.x, ai.x - aj.x);
        Y = anchor_add_exact(ag, 1, xi.y, xj.y, ai.y - aj.y);
        Z = anchor_add_exact(ag, 2, xi.z, xj.z, ai.z - aj.z);
    }
    pair_ij<float, ENERGY>(t, ti, __float_as_int(nj.w), dx, dy, dz, X, Y, Z, r2, ni.x, ni.y, ni.z, nj.x, nj.y, nj.z,
                           acc);
}

// ORBC_PIPE: the mirrors of the next group are gathered while this group is
// evaluated (two groups of UNROLL records in registers), instead of
// gathering each group right before its use.
#ifndef ORBC_PIPE
#define ORBC_PIPE 1
#endif
template <bool ENERGY, int U = UNROLL, bool PIPE = (ORBC_PIPE != 0)>
__device__ __forceinline__ void range_mixed(long long i, int kb, int ke, long long stride, const PQ *pq,
                                            const uint32_t *nb, const Anchor &ag, const AnchorI &ai, float4 xi,
                                            float4 ni, int ti, const FFS<float> &t, Acc<double> &out) {
    constexpr int UNROLL = U;
    uint32_t jn[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) jn[u] = (kb + u < ke) ? nb[(long long)(kb + u) * stride] : (uint32_t)i;
    if constexpr (PIPE) {
        float4 xa[UNROLL], na[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) { xa[u] = pq[jn[u]].x; na[u] = pq[jn[u]].n; }
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const int kn = kb + UNROLL + u;
            jn[u] = (kn < ke) ? nb[(long long)kn * stride] : (uint32_t)i;
        }
        for (int k0 = kb; k0 < ke; k0 += UNROLL) {
            float4 xb[UNROLL], nbv[UNROLL];
            const bool more = k0 + UNROLL < ke;
            if (more) {
#pragma unroll
                for (int u = 0; u < UNROLL; ++u) { xb[u] = pq[jn[u]].x; nbv[u] = pq[jn[u]].n; }
#pragma unroll
                for (int u = 0; u < UNROLL; ++u) {
                    const int kn = k0 + 2 * UNROLL + u;
                    jn[u] = (kn < ke) ? nb[(long long)kn * stride] : (uint32_t)i;
                }
            }
            Acc<float> acc32;
            Acc<PairAcc> &acc = *[&]() {
                if constexpr (ORBC_PAIR_F64 >= 2) return &out; else return &acc32;
            }();
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) pair_mixed<ENERGY>(ag, ai, xi, ni, ti, xa[u], na[u], t, acc);
            if constexpr (ORBC_PAIR_F64 < 2) add_to(out, acc32);
            if (more) {
#pragma unroll
                for (int u = 0; u < UNROLL; ++u) { xa[u] = xb[u]; na[u] = nbv[u]; }
            }
        }
    } else {
        for (int k0 = kb; k0 < ke; k0 += UNROLL) {
            uint32_t jj[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
                jj[u] = jn[u];
                const int kn = k0 + UNROLL + u;
                jn[u] = (kn < ke) ? nb[(long long)kn * stride] : (uint32_t)i;
            }
            float4 xj[UNROLL], nj[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) { xj[u] = pq[jj[u]].x; nj[u] = pq[jj[u]].n; }
            Acc<float> acc32;
            // FP64 terms go straight into the FP64 sums
            Acc<PairAcc> &acc = *[&]() {
                if constexpr (ORBC_PAIR_F64 >= 2) return &out; else return &acc32;
            }();
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) pair_mixed<ENERGY>(ag, ai, xi, ni, ti, xj[u], nj[u], t, acc);
            if constexpr (ORBC_PAIR_F64 < 2) add_to(out, acc32);
        }
    }
}

// Mixed walk over a block halo in LDS (k_halo): 16-bit entries, two per
// word (local index in bits 0-9, the neighbour's type in bits 10-12), rows
// padded to whole groups of HALO_GROUP entries with the particle itself.
// Words are loaded two ahead; each entry's mirrors are read from LDS one
// entry ahead of its use, and its pair-type record as soon as its word is in.
// The terms and their order are those of range_mixed over the 32-bit list.
__device__ __forceinline__ void range_halo(int nn, long long stride, const uint32_t *w16, const float4 *sx,
                                           const float4 *sn, const PairRec *spp, int crow, const Anchor &ag,
                                           const AnchorI &ai, float4 xi, float4 ni, Acc<double> &out) {
    const int nw = halo_row_words(nn);
    if (nw == 0) return;
    uint32_t w = w16[0];
    uint32_t wn = (nw > 1) ? w16[stride] : 0u;
    float4 xa = sx[w & 0x3ffu], na = sn[w & 0x3ffu];
    for (int m = 0; m < nw; ++m) {
        const uint32_t wc = w;
        w = wn;
        if (m + 2 < nw) wn = w16[(long long)(m + 2) * stride];
        const uint32_t eb = wc >> 16;
        const float4 xb = sx[eb & 0x3ffu], nb = sn[eb & 0x3ffu];
        const PairRec pa = spp[crow + (int)((wc >> 10) & 7u)];
        pair_rec(ag, ai, xi, ni, xa, na, pa, out);
        if (m + 1 < nw) { xa = sx[w & 0x3ffu]; na = sn[w & 0x3ffu]; }
        const PairRec pb = spp[crow + (int)(eb >> 10)];
        pair_rec(ag, ai, xi, ni, xb, nb, pb, out);
    }
}

// compute_bonded from the FP32 mirrors (f only); TAB holds K[] and r0[]
template <bool ENERGY, typename TAB>
__device__ __forceinline__ void bonds_mixed(long long i, long long stride, const PQ *pq, const uint32_t *bnd,
                                            const uint8_t *nbnd, const Anchor &ag, const AnchorI &ai, float4 xi,
                                            const TAB &t, Acc<double> &out) {
    const int nb = nbnd[i];
    Acc<float> acc;
    for (int k = 0; k < nb; ++k) {
        const uint32_t e = bnd[(long long)k * stride + i];
        const uint32_t j = e & BOND_INDEX_MASK;
        const float4 xj = pq[j].x;
        const AnchorI aj = unpack(__float_as_uint(xj.w), ag);
        const float dx = anchor_add(ag, 0, sub_rn(xj.x, xi.x), aj.x - ai.x);
        const float dy = anchor_add(ag, 1, sub_rn(xj.y, xi.y), aj.y - ai.y);
        const float dz = anchor_add(ag, 2, sub_rn(xj.z, xi.z), aj.z - ai.z);
        bond_ij<float, ENERGY>(t, (int)(e >> 30), dx, dy, dz, acc);
    }
    add_to(out, acc);
}

template <bool ENERGY>
__device__ __forceinline__ void accumulate_mixed(long long i, long long stride, const PQ *pq,
                                                 const uint32_t *nbr, const uint32_t *nnbr,
                                                 const uint32_t *bnd, const uint8_t *nbnd, const Anchor &ag, float4 xi,
                                                 float4 ni, const FFS<float> &t, Acc<double> &out, bool pair,
                                                 bool bond) {
    const int ti = __float_as_int(ni.w);
    const AnchorI ai = unpack(__float_as_uint(xi.w), ag);
    if (pair) {
        range_mixed<ENERGY>(i, 0, (int)nnbr[i], stride, pq, nbr + i, ag, ai, xi, ni, ti, t, out);
    }
    if (bond && nbnd) bonds_mixed<ENERGY>(i, stride, pq, bnd, nbnd, ag, ai, xi, t, out);
}

// FP64: one gathered neighbour against owner i (minimum image on periodic axes)
template <bool ENERGY>
__device__ __forceinline__ void pair_fp64(const Anchor &ag, bool periodic, double xi, double yi, double zi, double nix,
                                          double niy, double niz, int ti, const double (&xj)[3], const double (&nj)[3],
                                          int tj, const FFS<double> &t, Acc<double> &acc) {
    double dx = sub_rn(xi, xj[0]);
    double dy = sub_rn(yi, xj[1]);
    double dz = sub_rn(zi, xj[2]);
    if (periodic) min_image(ag, dx, dy, dz);
    const double r2 = r2_rn(dx, dy, dz);
    pair_ij<double, ENERGY>(t, ti, tj, dx, dy, dz, dx, dy, dz, r2, nix, niy, niz, nj[0], nj[1], nj[2], acc);
}

#ifndef ORBC_PIPE64
#define ORBC_PIPE64 0
#endif
template <bool ENERGY>
__device__ __forceinline__ void range_fp64(long long i, int kb, int ke, long long stride, const double *x,
                                           const double *nrm, const uint8_t *type, const uint32_t *nb,
                                           const Anchor &ag, bool periodic, double xi,
                                           double yi, double zi, double nix, double niy, double niz, int ti,
                                           const FFS<double> &t, Acc<double> &acc) {
    long long jn[UNROLL64];
#pragma unroll
    for (int u = 0; u < UNROLL64; ++u) jn[u] = (kb + u < ke) ? (long long)nb[(long long)(kb + u) * stride] : i;
    if constexpr (ORBC_PIPE64) {
        // gathers of the next group in flight while this group is evaluated
        double xa[UNROLL64][3], na[UNROLL64][3];
        int ta[UNROLL64];
#pragma unroll
        for (int u = 0; u < UNROLL64; ++u) {
            for (int d = 0; d < 3; ++d) { xa[u][d] = x[3 * jn[u] + d]; na[u][d] = nrm[3 * jn[u] + d]; }
            ta[u] = type[jn[u]];
        }
#pragma unroll
        for (int u = 0; u < UNROLL64; ++u) {
            const int kn = kb + UNROLL64 + u;
            jn[u] = (kn < ke) ? (long long)nb[(long long)kn * stride] : i;
        }
        for (int k0 = kb; k0 < ke; k0 += UNROLL64) {
            double xb[UNROLL64][3], nbv[UNROLL64][3];
            int tb[UNROLL64];
            const bool more = k0 + UNROLL64 < ke;
            if (more) {
#pragma unroll
                for (int u = 0; u < UNROLL64; ++u) {
                    for (int d = 0; d < 3; ++d) { xb[u][d] = x[3 * jn[u] + d]; nbv[u][d] = nrm[3 * jn[u] + d]; }
                    tb[u] = type[jn[u]];
                }
#pragma unroll
                for (int u = 0; u < UNROLL64; ++u) {
                    const int kn = k0 + 2 * UNROLL64 + u;
                    jn[u] = (kn < ke) ? (long long)nb[(long long)kn * stride] : i;
                }
            }
#pragma unroll
            for (int u = 0; u < UNROLL64; ++u)
                pair_fp64<ENERGY>(ag, periodic, xi, yi, zi, nix, niy, niz, ti, xa[u], na[u], ta[u], t, acc);
            if (more) {
#pragma unroll
                for (int u = 0; u < UNROLL64; ++u) {
                    for (int d = 0; d < 3; ++d) { xa[u][d] = xb[u][d]; na[u][d] = nbv[u][d]; }
                    ta[u] = tb[u];
                }
            }
        }
    } else {
        for (int k0 = kb; k0 < ke; k0 += UNROLL64) {
            long long jj[UNROLL64];
#pragma unroll
            for (int u = 0; u < UNROLL64; ++u) {
                jj[u] = jn[u];
                const int kn = k0 + UNROLL64 + u;
                jn[u] = (kn < ke) ? (long long)nb[(long long)kn * stride] : i;
            }
            double xj[UNROLL64][3], nj[UNROLL64][3];
            int tj[UNROLL64];
#pragma unroll
            for (int u = 0; u < UNROLL64; ++u) {
                for (int d = 0; d < 3; ++d) { xj[u][d] = x[3 * jj[u] + d]; nj[u][d] = nrm[3 * jj[u] + d]; }
                tj[u] = type[jj[u]];
            }
#pragma unroll
            for (int u = 0; u < UNROLL64; ++u)
                pair_fp64<ENERGY>(ag, periodic, xi, yi, zi, nix, niy, niz, ti, xj[u], nj[u], tj[u], t, acc);
        }
    }
}

template <bool ENERGY>
__device__ __forceinline__ void accumulate_fp64(long long i, long long stride, const double *x, const double *nrm,
                                                const uint8_t *type, const uint32_t *nbr, const uint32_t *nnbr,
                                                const uint32_t *bnd, const uint8_t *nbnd,
                                                const Anchor &ag, const FFS<double> &t, Acc<double> &acc, bool pair,
                                                bool bond) {
    const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
    if (pair) {
        const double nix = nrm[3 * i], niy = nrm[3 * i + 1], niz = nrm[3 * i + 2];
        const int ti = type[i];
        const bool periodic = ag.L[0] > 0 || ag.L[1] > 0 || ag.L[2] > 0;
        range_fp64<ENERGY>(i, 0, (int)nnbr[i], stride, x, nrm, type, nbr + i, ag, periodic, xi, yi, zi, nix, niy, niz,
                           ti, t, acc);
    }
    if (bond && nbnd) {
        const int nb = nbnd[i];
        for (int k = 0; k < nb; ++k) {
            const uint32_t e = bnd[(long long)k * stride + i];
            const long long j = e & BOND_INDEX_MASK;
            double dx = x[3 * j] - xi, dy = x[3 * j + 1] - yi, dz = x[3 * j + 2] - zi;
            min_image(ag, dx, dy, dz);
            bond_ij<double, ENERGY>(t, (int)(e >> 30), dx, dy, dz, acc);
        }
    }
}

// ---- standalone force evaluation (compute_pairwise_fused / compute_bonded) --
template <bool MIXED, bool ENERGY>
__global__ __launch_bounds__(BLOCK) void k_forces(ForceArgs a, int pair, int bond) {
    __shared__ FFS<typename std::conditional<MIXED, float, double>::type> t;
    load_ffs(&t);
    const long long i = block_index(a.swizzle) * BLOCK + threadIdx.x;
    if (i >= a.n) return;
    Acc<double> acc;
    if constexpr (MIXED) {
        accumulate_mixed<ENERGY>(i, a.stride, a.pq, a.nbr, a.nnbr, a.bnd, a.nbnd, a.anc_grid, a.pq[i].x,
                                 a.pq[i].n, t, acc, pair, bond);
        a.f32[3 * i] += (float)acc.fx; a.f32[3 * i + 1] += (float)acc.fy; a.f32[3 * i + 2] += (float)acc.fz;
        a.t32[3 * i] += (float)acc.tx; a.t32[3 * i + 1] += (float)acc.ty; a.t32[3 * i + 2] += (float)acc.tz;
    } else {
        accumulate_fp64<ENERGY>(i, a.stride, a.x, a.nrm, a.type, a.nbr, a.nnbr, a.bnd, a.nbnd,
                                a.anc_grid, t, acc, pair, bond);
        a.f64[3 * i] += acc.fx; a.f64[3 * i + 1] += acc.fy; a.f64[3 * i + 2] += acc.fz;
        a.t64[3 * i] += acc.tx; a.t64[3 * i + 1] += acc.ty; a.t64[3 * i + 2] += acc.tz;
    }
    if (ENERGY) { a.e[i] += acc.e; a.w[i] += acc.w; }
}

template <bool MIXED>
__global__ __launch_bounds__(BLOCK) void k_pairs_export(ForceArgs a, unsigned *ids_i, unsigned *ids_j,
                                                        unsigned long long *count, long long cap) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t idi = a.id[i];
    const int cnt = (int)a.nnbr[i];
    for (int k = 0; k < cnt; ++k) {
        const uint32_t j = a.nbr[(long long)k * a.stride + i];
        const uint32_t idj = a.id[j];
        if (!(idi < idj)) continue;
        bool hit;
        if (MIXED) {
            const float4 xi = a.pq[i].x, xj = a.pq[j].x;
            const AnchorI ai = unpack(__float_as_uint(xi.w), a.anc_grid), aj = unpack(__float_as_uint(xj.w), a.anc_grid);
            const float dx = anchor_add(a.anc_grid, 0, sub_rn(xi.x, xj.x), ai.x - aj.x);
            const float dy = anchor_add(a.anc_grid, 1, sub_rn(xi.y, xj.y), ai.y - aj.y);
            const float dz = anchor_add(a.anc_grid, 2, sub_rn(xi.z, xj.z), ai.z - aj.z);
            const float r2 = r2_rn(dx, dy, dz);
            const int kk = __float_as_int(a.pq[i].n.w) * NT + __float_as_int(a.pq[j].n.w);
            hit = r2 > 1e-5f && (r2 < c_fff.cut2_poly[kk] || r2 < c_fff.cut2_lj[kk]);
        } else {
            double dx = sub_rn(a.x[3 * i], a.x[3 * (long long)j]);
            double dy = sub_rn(a.x[3 * i + 1], a.x[3 * (long long)j + 1]);
            double dz = sub_rn(a.x[3 * i + 2], a.x[3 * (long long)j + 2]);
            min_image(a.anc_grid, dx, dy, dz);
            const double r2 = r2_rn(dx, dy, dz);
            const int kk = a.type[i] * NT + a.type[j];
            hit = r2 > 1e-5 && (r2 < c_ffd.cut2_poly[kk] || r2 < c_ffd.cut2_lj[kk]);
        }
        if (hit) {
            const unsigned long long slot = atomicAdd(count, 1ull);
            if ((long long)slot < cap) { ids_i[slot] = idi; ids_j[slot] = idj; }
        }

    }
}

// ---- Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) -------------------
__device__ inline void philox4x32_10(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3, uint32_t k0,
                                     uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
}

// ---- fused timestep: forces + bonds + verlet_langevin ---------------------
// src/openrbc.cpp:219-233 (compute_pairwise_fused, compute_bonded,
// integrate(verlet_langevin)) in one pass; the Langevin update restates
// src/integrate_langevin.h:275-297 in FP64 with counter-based noise.
// HALO (mixed only): the block's neighbours are staged in LDS from its halo
// (k_halo) and the list is walked as 16-bit local indices; a block marked
// HALO_OVERFLOW gathers from the 32-bit list, one neighbour at a time.
// ---- block halos in the fused step kernels ----------------------------------
// LDS of a block-halo step kernel: the halo's mirrors (own particles at
// [0, BLOCK), outside neighbours after them), one record per type pair, the
// bond constants.
template <int CAP>
struct HaloLDS {
    float4 x[CAP], n[CAP];
    PairRec pp[36];
    float K[NBT], r0[NBT];
};

// Every thread of the block: the constants and, unless the block's halo
// overflowed (*hcnt = HALO_OVERFLOW), the halo mirrors; returns the mirror of
// particle i (blk: this block's particle block).
template <int CAP>
__device__ __forceinline__ PQ halo_stage(const StepArgs &a, long long blk, bool act, long long i, HaloLDS<CAP> &L,
                                         uint32_t &hcnt) {
    load_pair_recs(L.pp);
    if (threadIdx.x < NBT) { L.K[threadIdx.x] = c_fff.bond_K[threadIdx.x]; L.r0[threadIdx.x] = c_fff.bond_r0[threadIdx.x]; }
    hcnt = a.halo_cnt[blk];
    PQ qi{};
    if (act) qi = a.pq_in[i];
    if (hcnt != HALO_OVERFLOW) {
        L.x[threadIdx.x] = qi.x; L.n[threadIdx.x] = qi.n;
        const uint32_t *ext = a.halo_ext + blk * HALO_EXT_MAX;
        for (uint32_t q = threadIdx.x; q < hcnt; q += BLOCK) {
            const PQ p = a.pq_in[ext[q]];
            L.x[BLOCK + q] = p.x; L.n[BLOCK + q] = p.n;
        }
    }
    __syncthreads();
    return qi;
}

// pair and bond forces of active particle i from the staged halo (its row at
// column `col` of the 16-bit list), or, for an overflowed block, gathered one
// neighbour at a time
template <int CAP>
__device__ __forceinline__ void halo_accumulate(const StepArgs &a, const Anchor &ag, long long i, long long col,
                                                const PQ &qi, uint32_t hcnt, const HaloLDS<CAP> &L,
                                                Acc<double> &acc) {
    const int ti = __float_as_int(qi.n.w);
    const AnchorI ai = unpack(__float_as_uint(qi.x.w), ag);
    const int nn = (int)a.nnbr[i];
    const int crow = ti * NT;
    if (hcnt != HALO_OVERFLOW) {
        range_halo(nn, a.stride, a.nbr16 + col, L.x, L.n, L.pp, crow, ag, ai, qi.x, qi.n, acc);
    } else {
        for (int k = 0; k < nn; ++k) {
            const PQ q = a.pq_in[a.nbr[(long long)k * a.stride + i]];
            pair_rec(ag, ai, qi.x, qi.n, q.x, q.n, L.pp[crow + __float_as_int(q.n.w)], acc);
        }
    }
    if (a.has_bonds) bonds_mixed<false>(i, a.stride, a.pq_in, a.bnd, a.nbnd, ag, ai, qi.x, L, acc);
}

// FP64 block halo: positions and directors of the halo particles (48 bytes,
// three 16-byte LDS words each); the pair terms go through pair_fp64 with the
// FP64 tables, in the order of range_fp64, so the forces are those of the
// gather walk.  The FP64 terms are written as plain expressions whose
// multiply-adds the compiler fuses; a walk that interleaves the loads of the
// next entry with a pair's arithmetic (measured) changes which ones it fuses,
// so this walk stays as simple as range_fp64's.
struct alignas(16) XN64 {
    double x[3], n[3];
};
template <int CAP>
struct HaloLDS64 {
    XN64 r[CAP];
};

template <int CAP>
__device__ __forceinline__ void halo_stage64(const StepArgs &a, long long blk, bool act, long long i,
                                             HaloLDS64<CAP> &L, uint32_t &hcnt) {
    hcnt = a.halo_cnt[blk];
    if (hcnt != HALO_OVERFLOW) {
        XN64 q{};
        if (act)
            for (int d = 0; d < 3; ++d) { q.x[d] = a.x_in[3 * i + d]; q.n[d] = a.n_in[3 * i + d]; }
        L.r[threadIdx.x] = q;
        const uint32_t *ext = a.halo_ext + blk * HALO_EXT_MAX;
        for (uint32_t e = threadIdx.x; e < hcnt; e += BLOCK) {
            const long long j = ext[e];
            XN64 p;
            for (int d = 0; d < 3; ++d) { p.x[d] = a.x_in[3 * j + d]; p.n[d] = a.n_in[3 * j + d]; }
            L.r[BLOCK + e] = p;
        }
    }
    __syncthreads();
}

template <int CAP>
__device__ __forceinline__ void halo_accumulate64(const StepArgs &a, const Anchor &ag, long long i, long long col,
                                                  uint32_t hcnt, const HaloLDS64<CAP> &L, const FFS<double> &t,
                                                  Acc<double> &acc) {
    const double xi = a.x_in[3 * i], yi = a.x_in[3 * i + 1], zi = a.x_in[3 * i + 2];
    const double nix = a.n_in[3 * i], niy = a.n_in[3 * i + 1], niz = a.n_in[3 * i + 2];
    const int ti = a.type[i];
    const bool periodic = ag.L[0] > 0 || ag.L[1] > 0 || ag.L[2] > 0;
    const int nn = (int)a.nnbr[i];
    if (hcnt != HALO_OVERFLOW) {
        const uint32_t *w16 = a.nbr16 + col;
        const int nw = halo_row_words(nn);
        for (int m = 0; m < nw; ++m) {
            const uint32_t w = w16[(long long)m * a.stride];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint32_t e = h ? (w >> 16) : (w & 0xffffu);
                const XN64 &q = L.r[e & 0x3ffu];
                pair_fp64<false>(ag, periodic, xi, yi, zi, nix, niy, niz, ti, q.x, q.n, (int)(e >> 10), t, acc);
            }
        }
    } else {
        range_fp64<false>(i, 0, nn, a.stride, a.x_in, a.n_in, a.type, a.nbr + i, ag, periodic, xi, yi, zi, nix, niy,
                          niz, ti, t, acc);
    }
    if (a.has_bonds && a.nbnd) {
        const int nb = a.nbnd[i];
        for (int k = 0; k < nb; ++k) {
            const uint32_t e = a.bnd[(long long)k * a.stride + i];
            const long long j = e & BOND_INDEX_MASK;
            double dx = a.x_in[3 * j] - xi, dy = a.x_in[3 * j + 1] - yi, dz = a.x_in[3 * j + 2] - zi;
            min_image(ag, dx, dy, dz);
            bond_ij<double, false>(t, (int)(e >> 30), dx, dy, dz, acc);
        }
    }
}

template <bool MIXED, bool HALO, int CAP = HALO_CAP>
__device__ __forceinline__ void step_langevin(const StepArgs &a) {
    // the anchor grid from its device copy (rewritten at every neighbour
    // rebuild, so a captured batch stays valid), a uniform load at the top
    const Anchor ag = *a.anc_dev;
    const unsigned long long step = (a.step0 ? *a.step0 : 0ull) + a.step;
    const int k3 = (int)(step % 3ull);
    // skin flag of this step's input positions; no block of this launch writes it
    if (*a.stop || a.dslot[(k3 + 2) % 3] != 0u) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *a.stop = 1;
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.dslot[(k3 + 1) % 3] = 0u;
        atomicAdd(a.done, 1);
    }
    // force-field constants: the FP32/FP64 tables, or the block halo with
    // its constants (each instantiation keeps only the LDS it references)
    __shared__ FFS<typename std::conditional<MIXED, float, double>::type> tab;
    __shared__ HaloLDS<CAP> hl;
    __shared__ HaloLDS64<CAP> hl64;
    const long long blk = block_index(a.swizzle);
    // this lane's position in the block (its column of the 16-bit list),
    // recomputed where needed rather than kept live
#define ORBC_LANE_COL (blk * BLOCK + (long long)threadIdx.x)
    uint32_t hcnt = HALO_OVERFLOW;
    PQ qi{};
    if constexpr (HALO && MIXED) {
        qi = halo_stage(a, blk, ORBC_LANE_COL < a.n, ORBC_LANE_COL, hl, hcnt);
    } else {
        load_ffs(&tab);
        if constexpr (HALO) halo_stage64(a, blk, ORBC_LANE_COL < a.n, ORBC_LANE_COL, hl64, hcnt);
    }
    // the particle of this lane: the block's lane order (k_halo) when its
    // neighbours are staged, else the lane's own position
    int l = (int)threadIdx.x;
    if constexpr (HALO) {
        if (hcnt != HALO_OVERFLOW) {
            l = a.hord[ORBC_LANE_COL];
            if constexpr (MIXED) { qi.x = hl.x[l]; qi.n = hl.n[l]; }
        }
    }
    const long long i = blk * BLOCK + l;
    const bool act = i < a.n;
    float d2 = 0.f;
    bool bad = false;
    if (act) {
        double fx, fy, fz, tx, ty, tz;
        int ti;
        uint32_t wi;
        Acc<double> acc;
        if constexpr (HALO && MIXED) {
            ti = __float_as_int(qi.n.w);
            wi = __float_as_uint(qi.x.w);
            halo_accumulate(a, ag, i, ORBC_LANE_COL, qi, hcnt, hl, acc);
        } else if constexpr (HALO) {
            halo_accumulate64(a, ag, i, ORBC_LANE_COL, hcnt, hl64, tab, acc);
            ti = a.type[i];
            wi = a.anc[i];
        } else if constexpr (MIXED) {
            const PQ qi = a.pq_in[i];
            const float4 xi4 = qi.x, ni4 = qi.n;
            accumulate_mixed<false>(i, a.stride, a.pq_in, a.nbr, a.nnbr, a.bnd, a.nbnd, ag, xi4,
                                    ni4, tab, acc, true, a.has_bonds != 0);
            ti = __float_as_int(ni4.w);
            wi = __float_as_uint(xi4.w);
        } else {
            accumulate_fp64<false>(i, a.stride, a.x_in, a.n_in, a.type, a.nbr, a.nnbr, a.bnd, a.nbnd,
                                   ag, tab, acc, true, a.has_bonds != 0);
            ti = a.type[i];
            wi = a.anc[i];
        }
        fx = acc.fx; fy = acc.fy; fz = acc.fz; tx = acc.tx; ty = acc.ty; tz = acc.tz;
        if (a.f_ext) { fx += a.f_ext[3 * i]; fy += a.f_ext[3 * i + 1]; fz += a.f_ext[3 * i + 2]; }
        if (a.push) {
            // constrain_volume push along the particle's cell normal (src/constrain_volume.h:67-75)
            const int kc = a.aff[i];
            if (kc >= 0) {
                const double sp = (*a.push) * c_ffd.mass[ti];
                fx += sp * a.cell_normal[3 * kc]; fy += sp * a.cell_normal[3 * kc + 1]; fz += sp * a.cell_normal[3 * kc + 2];
            }
        }
        // rotation: t = n x t; o += dt t; n = normalize(n + (o x n) dt)
        const double nx = a.n_in[3 * i], ny = a.n_in[3 * i + 1], nz = a.n_in[3 * i + 2];
        double ox = a.o[3 * i], oy = a.o[3 * i + 1], oz = a.o[3 * i + 2];
        const double dt = a.dt;
        const double qx = ny * tz - nz * ty, qy = nz * tx - nx * tz, qz = nx * ty - ny * tx;
        ox += dt * qx; oy += dt * qy; oz += dt * qz;
        double gx = nx + (oy * nz - oz * ny) * dt;
        double gy = ny + (oz * nx - ox * nz) * dt;
        double gz = nz + (ox * ny - oy * nx) * dt;
        const double gs = rsqrt(gx * gx + gy * gy + gz * gz);
        gx *= gs; gy *= gs; gz *= gs;
        // translation with uniform [-1,1) noise (src/integrate_langevin.h:282-296)
        uint32_t c0 = a.id[i], c1 = (uint32_t)step, c2 = (uint32_t)(step >> 32), c3 = 0u;
        philox4x32_10(c0, c1, c2, c3, (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
        const double u = 4.656612873077392578125e-10;
        const double rx = (double)c0 * u - 1.0, ry = (double)c1 * u - 1.0, rz = (double)c2 * u - 1.0;
        double vx = a.v[3 * i], vy = a.v[3 * i + 1], vz = a.v[3 * i + 2];
        const double gam = a.gamma[ti], sig = a.sig_group ? a.sig_group[a.grp[i] * NT + ti] : a.sigma[ti];
        fx -= gam * vx + sig * rx;
        fy -= gam * vy + sig * ry;
        fz -= gam * vz + sig * rz;
        const double dtm = dt * c_ffd.inv_mass[ti];
        vx += fx * dtm; vy += fy * dtm; vz += fz * dtm;
        const double px = a.x_in[3 * i] + vx * dt, py = a.x_in[3 * i + 1] + vy * dt, pz = a.x_in[3 * i + 2] + vz * dt;
        a.x_out[3 * i] = px; a.x_out[3 * i + 1] = py; a.x_out[3 * i + 2] = pz;
        a.n_out[3 * i] = gx; a.n_out[3 * i + 1] = gy; a.n_out[3 * i + 2] = gz;
        a.v[3 * i] = vx; a.v[3 * i + 1] = vy; a.v[3 * i + 2] = vz;
        a.o[3 * i] = ox; a.o[3 * i + 1] = oy; a.o[3 * i + 2] = oz;
        // FP32 mirror with the build-time anchor, and the skin check
        const AnchorI ai = unpack(wi, ag);
        const float r0 = __double2float_rn(px - anchor_pos(ag, 0, ai.x));
        const float r1 = __double2float_rn(py - anchor_pos(ag, 1, ai.y));
        const float r2 = __double2float_rn(pz - anchor_pos(ag, 2, ai.z));
        if (MIXED) {
            PQ q;
            q.x = make_float4(r0, r1, r2, __uint_as_float(wi));
            q.n = make_float4(__double2float_rn(gx), __double2float_rn(gy), __double2float_rn(gz), __int_as_float(ti));
            a.pq_out[i] = q;
        }
        const float e0 = r0 - a.rb[3 * i], e1 = r1 - a.rb[3 * i + 1], e2 = r2 - a.rb[3 * i + 2];
        d2 = e0 * e0 + e1 * e1 + e2 * e2;
        bad = !(isfinite(px) && isfinite(py) && isfinite(pz) && isfinite(gx));
        if (bad) d2 = 3.0e38f;
    }
    // one atomic per wave, and only when some lane moved past its budget:
    // skin/2, or for a particle without list entries the budget k_budget gave it
    bool past = d2 > a.half_skin2;
    if (past && a.bud2 && !bad && a.nnbr[i] == 0) past = d2 > a.bud2[i];
    const unsigned long long over = __ballot(past);
    if (over && __lane_id() == (int)__ffsll((long long)over) - 1) atomicOr(a.dslot + k3, 1u);
    if (bad) atomicOr(a.bad, 1u);
}

template <bool MIXED>
__global__ __launch_bounds__(BLOCK) ORBC_STEP_ATTR void k_step_langevin(StepArgs a) {
    step_langevin<MIXED, false>(a);
}
__global__ __launch_bounds__(BLOCK) ORBC_HALO_ATTR void k_step_langevin_halo(StepArgs a) {
    step_langevin<true, true>(a);
}
// FP64 halo: 48-byte records, 3 blocks of 256 per CU's LDS
__global__ __launch_bounds__(BLOCK) void k_step_langevin64_halo(StepArgs a) { step_langevin<false, true>(a); }
// wide halos (HALO_CAP_WIDE local indices): the mixed kernel at 4 blocks of
// 256 per CU (128 registers), the FP64 one still at 3
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_step_langevin_halo_wide(StepArgs a) {
    step_langevin<true, true, HALO_CAP_WIDE>(a);
}
__global__ __launch_bounds__(BLOCK) void k_step_langevin64_halo_wide(StepArgs a) {
    step_langevin<false, true, HALO_CAP_WIDE>(a);
}

// ---- fused forces + post_toque_final_update (NH second half) --------------
// src/integrate_nh.h:237-273 after compute_pairwise_fused/compute_bonded, in
// the arithmetic of the unfused pair k_forces + k_nh_second: the mixed sums
// are rounded to FP32 as the FP32 force arrays hold them; t = n x t;
// v += dt/2 (f/m - zeta v); o += dt/2 t; f and n x t stored (FP32 arrays in
// mixed precision) for the next first half; KE partial per particle block.
template <bool MIXED, bool HALO, int CAP = HALO_CAP>
__device__ __forceinline__ void step_nh_second(const StepArgs &a, float *f32, float *t32, double *f64, double *t64,
                                               double *ke_partial, const double *zeta_p) {
    if (a.nh_gate) {
        // batched run: the first half of this step found a particle past its
        // budget (slot step % 3), or an earlier step stopped: nothing to do
        const int k3 = (int)(a.step % 3ull);
        if (*a.stop || a.dslot[k3] != 0u) {
            if (blockIdx.x == 0 && threadIdx.x == 0) *a.stop = 1;
            return;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(a.done, 1);
    }
    __shared__ FFS<typename std::conditional<MIXED, float, double>::type> tab;
    __shared__ HaloLDS<CAP> hl;
    __shared__ HaloLDS64<CAP> hl64;
    const long long blk = block_index(a.swizzle);
    uint32_t hcnt = HALO_OVERFLOW;
    PQ qh{};
    if constexpr (HALO && MIXED) {
        qh = halo_stage(a, blk, ORBC_LANE_COL < a.n, ORBC_LANE_COL, hl, hcnt);
    } else {
        load_ffs(&tab);
        if constexpr (HALO) halo_stage64(a, blk, ORBC_LANE_COL < a.n, ORBC_LANE_COL, hl64, hcnt);
    }
    // the particle of this lane (see step_langevin)
    int l = (int)threadIdx.x;
    if constexpr (HALO) {
        if (hcnt != HALO_OVERFLOW) {
            l = a.hord[ORBC_LANE_COL];
            if constexpr (MIXED) { qh.x = hl.x[l]; qh.n = hl.n[l]; }
        }
    }
    const long long i = blk * BLOCK + l;
    const bool act = i < a.n;
    double ke = 0.0;
    if (act) {
        double f[3], t[3];
        int ti;
        Acc<double> acc;
        if constexpr (HALO && MIXED) {
            halo_accumulate(a, a.anc_grid, i, ORBC_LANE_COL, qh, hcnt, hl, acc);
            ti = __float_as_int(qh.n.w);
        } else if constexpr (HALO) {
            halo_accumulate64(a, a.anc_grid, i, ORBC_LANE_COL, hcnt, hl64, tab, acc);
            ti = a.type[i];
        } else if constexpr (MIXED) {
            const PQ qi = a.pq_in[i];
            accumulate_mixed<false>(i, a.stride, a.pq_in, a.nbr, a.nnbr, a.bnd, a.nbnd, a.anc_grid,
                                    qi.x, qi.n, tab, acc, true, a.has_bonds != 0);
            ti = __float_as_int(qi.n.w);
        } else {
            accumulate_fp64<false>(i, a.stride, a.x_in, a.n_in, a.type, a.nbr, a.nnbr, a.bnd, a.nbnd,
                                   a.anc_grid, tab, acc, true, a.has_bonds != 0);
            ti = a.type[i];
        }
        if constexpr (MIXED) {
            f[0] = (float)acc.fx; f[1] = (float)acc.fy; f[2] = (float)acc.fz;
            t[0] = (float)acc.tx; t[1] = (float)acc.ty; t[2] = (float)acc.tz;
        } else {
            f[0] = acc.fx; f[1] = acc.fy; f[2] = acc.fz;
            t[0] = acc.tx; t[1] = acc.ty; t[2] = acc.tz;
        }
        const double m = c_ffd.mass[ti];
        const double zeta = *zeta_p;
        double v[3], n[3], q[3];
        for (int d = 0; d < 3; ++d) { v[d] = a.v[3 * i + d]; n[d] = a.n_in[3 * i + d]; }
        q[0] = n[1] * t[2] - n[2] * t[1];
        q[1] = n[2] * t[0] - n[0] * t[2];
        q[2] = n[0] * t[1] - n[1] * t[0];
        for (int d = 0; d < 3; ++d) {
            v[d] += 0.5 * a.dt * (f[d] / m - zeta * v[d]);
            a.o[3 * i + d] += 0.5 * a.dt * q[d];
            a.v[3 * i + d] = v[d];
        }
        if constexpr (MIXED) {
            for (int d = 0; d < 3; ++d) { f32[3 * i + d] = (float)f[d]; t32[3 * i + d] = (float)q[d]; }
        } else {
            for (int d = 0; d < 3; ++d) { f64[3 * i + d] = f[d]; t64[3 * i + d] = q[d]; }
        }
        ke = 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    if constexpr (HALO) {
        // back to particle order, the order of the unfused reduction
        if (hcnt != HALO_OVERFLOW) {          // block-uniform
            __shared__ double s_ke[BLOCK];
            s_ke[i - blk * BLOCK] = ke;
            __syncthreads();
            ke = s_ke[threadIdx.x];
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) ke += __shfl_xor(ke, off, 64);
    __shared__ double ws[BLOCK / 64];
    if (__lane_id() == 0) ws[threadIdx.x >> 6] = ke;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int w = 0; w < BLOCK / 64; ++w) s += ws[w];
        ke_partial[blk] = s;          // by particle block, the order of the unfused reduction
    }
}

template <bool MIXED>
__global__ __launch_bounds__(BLOCK) void k_step_nh_second(StepArgs a, float *f32, float *t32, double *f64,
                                                          double *t64, double *ke_partial, const double *zeta_p) {
    step_nh_second<MIXED, false>(a, f32, t32, f64, t64, ke_partial, zeta_p);
}
#ifndef ORBC_NH_HALO_WAVES
#define ORBC_NH_HALO_WAVES ORBC_HALO_WAVES
#endif
#if ORBC_NH_HALO_WAVES > 0
#define ORBC_NH_HALO_ATTR __attribute__((amdgpu_waves_per_eu(ORBC_NH_HALO_WAVES)))
#else
#define ORBC_NH_HALO_ATTR
#endif
__global__ __launch_bounds__(BLOCK) ORBC_NH_HALO_ATTR void k_step_nh_second_halo(StepArgs a, float *f32, float *t32,
                                                                              double *f64, double *t64,
                                                                              double *ke_partial,
                                                                              const double *zeta_p) {
    step_nh_second<true, true>(a, f32, t32, f64, t64, ke_partial, zeta_p);
}
__global__ __launch_bounds__(BLOCK) void k_step_nh_second64_halo(StepArgs a, float *f32, float *t32, double *f64,
                                                                 double *t64, double *ke_partial,
                                                                 const double *zeta_p) {
    step_nh_second<false, true>(a, f32, t32, f64, t64, ke_partial, zeta_p);
}
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_step_nh_second_halo_wide(
    StepArgs a, float *f32, float *t32, double *f64, double *t64, double *ke_partial, const double *zeta_p) {
    step_nh_second<true, true, HALO_CAP_WIDE>(a, f32, t32, f64, t64, ke_partial, zeta_p);
}
__global__ __launch_bounds__(BLOCK) void k_step_nh_second64_halo_wide(StepArgs a, float *f32, float *t32,
                                                                      double *f64, double *t64, double *ke_partial,
                                                                      const double *zeta_p) {
    step_nh_second<false, true, HALO_CAP_WIDE>(a, f32, t32, f64, t64, ke_partial, zeta_p);
}

inline unsigned nblk(long long n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

}  // namespace

void upload_ff(const FFTab<double> &d, const FFTab<float> &f, hipStream_t s) {
    ORBC_HIP_CHECK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_ffd), &d, sizeof(d), 0, hipMemcpyHostToDevice, s));
    ORBC_HIP_CHECK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_fff), &f, sizeof(f), 0, hipMemcpyHostToDevice, s));
    upload_mass(d.mass, s);
    ORBC_HIP_CHECK(hipStreamSynchronize(s));
}

long long step_blocks(long long n) { return nblk(n); }

void launch_forces(const ForceArgs &a, bool mixed, bool energy, bool pair, bool bond, hipStream_t s) {
    if (a.n <= 0) return;
    const dim3 g(nblk(a.n)), b(BLOCK);
    if (mixed) {
        if (energy) hipLaunchKernelGGL((k_forces<true, true>), g, b, 0, s, a, (int)pair, (int)bond);
        else hipLaunchKernelGGL((k_forces<true, false>), g, b, 0, s, a, (int)pair, (int)bond);
    } else {
        if (energy) hipLaunchKernelGGL((k_forces<false, true>), g, b, 0, s, a, (int)pair, (int)bond);
        else hipLaunchKernelGGL((k_forces<false, false>), g, b, 0, s, a, (int)pair, (int)bond);
    }
}

void launch_step_langevin(const StepArgs &a, bool mixed, hipStream_t s) {
    if (a.n <= 0) return;
    const dim3 g(nblk(a.n)), b(BLOCK);
    const bool wide = a.halo_cap == HALO_CAP_WIDE;
    if (mixed && a.halo_cnt && wide) hipLaunchKernelGGL(k_step_langevin_halo_wide, g, b, 0, s, a);
    else if (mixed && a.halo_cnt) hipLaunchKernelGGL(k_step_langevin_halo, g, b, 0, s, a);
    else if (mixed) hipLaunchKernelGGL(k_step_langevin<true>, g, b, 0, s, a);
    else if (a.halo_cnt && wide) hipLaunchKernelGGL(k_step_langevin64_halo_wide, g, b, 0, s, a);
    else if (a.halo_cnt) hipLaunchKernelGGL(k_step_langevin64_halo, g, b, 0, s, a);
    else hipLaunchKernelGGL(k_step_langevin<false>, g, b, 0, s, a);
}

void launch_step_nh_second(const StepArgs &a, bool mixed, float *f32, float *t32, double *f64, double *t64,
                           double *ke_partial, const double *zeta, hipStream_t s) {
    if (a.n <= 0) return;
    const dim3 g(nblk(a.n)), b(BLOCK);
    const bool wide = a.halo_cap == HALO_CAP_WIDE;
    if (mixed && a.halo_cnt && wide)
        hipLaunchKernelGGL(k_step_nh_second_halo_wide, g, b, 0, s, a, f32, t32, f64, t64, ke_partial, zeta);
    else if (mixed && a.halo_cnt)
        hipLaunchKernelGGL(k_step_nh_second_halo, g, b, 0, s, a, f32, t32, f64, t64, ke_partial, zeta);
    else if (mixed) hipLaunchKernelGGL(k_step_nh_second<true>, g, b, 0, s, a, f32, t32, f64, t64, ke_partial, zeta);
    else if (a.halo_cnt && wide)
        hipLaunchKernelGGL(k_step_nh_second64_halo_wide, g, b, 0, s, a, f32, t32, f64, t64, ke_partial, zeta);
    else if (a.halo_cnt)
        hipLaunchKernelGGL(k_step_nh_second64_halo, g, b, 0, s, a, f32, t32, f64, t64, ke_partial, zeta);
    else hipLaunchKernelGGL(k_step_nh_second<false>, g, b, 0, s, a, f32, t32, f64, t64, ke_partial, zeta);
}

void launch_pairs_export(const ForceArgs &a, bool mixed, unsigned *ids_i, unsigned *ids_j, unsigned long long *count,
                         long long cap, hipStream_t s) {
    if (a.n <= 0) return;
    const dim3 g(nblk(a.n)), b(BLOCK);
    if (mixed) hipLaunchKernelGGL(k_pairs_export<true>, g, b, 0, s, a, ids_i, ids_j, count, cap);
    else hipLaunchKernelGGL(k_pairs_export<false>, g, b, 0, s, a, ids_i, ids_j, count, cap);
}

}  // namespace orbc

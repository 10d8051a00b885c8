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

// One directed pair (owner i, neighbour j, d = x_i - x_j).  Returns true if
// the pair interacts (same test as the oracle, bit for bit).
template <typename R, bool ENERGY>
__device__ __forceinline__ bool pair_ij(int ti, int tj, R dx, R dy, R dz, R r2, R nix, R niy, R niz, R njx,
                                        R njy, R njz, Acc<R> &acc) {
    if (!(r2 > R(1e-5))) return false;
    const int k = ti * NT + tj;
    R cut2p, cutp, rep, att, alpha;
    int aniso;
    if (k == 0) {
        cut2p = R(kCut2LL); cutp = R(kCutLL); rep = R(kRepLL); att = R(kAttLL); alpha = R(kAlphaLL); aniso = 1;
    } else {
        const FFTab<R> &t = ff<R>();
        cut2p = t.cut2_poly[k]; cutp = t.cut_poly[k]; rep = t.rep[k]; att = t.att[k]; alpha = t.alpha[k];
        aniso = t.aniso[k];
    }
    R fx, fy, fz;
    if (r2 < cut2p) {
        const R rinv = rsq(r2);
        const R r = r2 * rinv;
        const R ex = dx * rinv, ey = dy * rinv, ez = dz * rinv;
        const R s = cutp - r;
        const R s3 = s * s * s, s4 = s * s3, s7 = s3 * s4;
        if (aniso) {
            const R ninj = nix * njx + niy * njy + niz * njz;
            const R nie = nix * ex + niy * ey + niz * ez;
            const R nje = njx * ex + njy * ey + njz * ez;
            const R a = ninj - nie * nje;
            const R A = R(1) + alpha * (a - R(1));
            const R pix = nix - nie * ex, piy = niy - nie * ey, piz = niz - nie * ez;
            const R pjx = njx - nje * ex, pjy = njy - nje * ey, pjz = njz - nje * ez;
            const R ua = att * s4;
            const R aua = alpha * ua;
            const R auar = aua * rinv;
            const R fra = R(8) * rep * s7 + A * R(4) * att * s3;
            fx = fra * ex + auar * (nje * pix + nie * pjx);
            fy = fra * ey + auar * (nje * piy + nie * pjy);
            fz = fra * ez + auar * (nje * piz + nie * pjz);
            acc.tx -= aua * pjx;
            acc.ty -= aua * pjy;
            acc.tz -= aua * pjz;
            if (ENERGY) acc.e += R(0.5) * (rep * s4 * s4 + A * ua);
        } else {
            const R fr = R(8) * rep * s7;
            fx = fr * ex; fy = fr * ey; fz = fr * ez;
            if (ENERGY) acc.e += R(0.5) * rep * s4 * s4;
        }
    } else {
        const FFTab<R> &t = ff<R>();
        if (!(r2 < t.cut2_lj[k])) return false;
        const R r2inv = R(1) / r2;
        const R r6inv = r2inv * r2inv * r2inv;
        const R fpair = r6inv * (t.lj1[k] * r6inv - t.lj2[k]) * r2inv;
        fx = dx * fpair; fy = dy * fpair; fz = dz * fpair;
        if (ENERGY) acc.e += R(0.5) * r6inv * (t.lj1[k] * (R(1) / R(12)) * r6inv - t.lj2[k] * (R(1) / R(6)));
    }
    acc.fx += fx; acc.fy += fy; acc.fz += fz;
    if (ENERGY) acc.w += R(0.5) * (dx * fx + dy * fy + dz * fz);
    return true;
}

// Bond on its owner i: d = x_j - x_i.
template <typename R, bool ENERGY>
__device__ __forceinline__ void bond_ij(int bt, R dx, R dy, R dz, Acc<R> &acc) {
    const FFTab<R> &t = ff<R>();
    const R r2 = dx * dx + dy * dy + dz * dz;
    const R rinv = rsq(r2);
    const R K = t.bond_K[bt], r0 = t.bond_r0[bt];
    const R c = K * (R(1) - r0 * rinv);
    acc.fx += c * dx; acc.fy += c * dy; acc.fz += c * dz;
    if (ENERGY) {
        const R dr = r2 * rinv - r0;
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

struct AnchorI {
    int x, y, z;
};
__device__ inline AnchorI unpack(uint32_t w, const Anchor &a) {
    return AnchorI{(int)(w & a.mx), (int)((w >> a.shy) & a.my), (int)(w >> a.shz)};
}

// ---- per-particle force accumulation (both precisions) ------------------
// Mixed: i's own mirror values are passed in; FP64 reads x, n directly.
template <bool ENERGY>
__device__ __forceinline__ void accumulate_mixed(long long i, long long stride, const float4 *xq, const float4 *nq,
                                                 const uint32_t *nbr, const uint32_t *nnbr, const uint32_t *bnd,
                                                 const uint8_t *nbnd, const Anchor &ag, float4 xi, float4 ni,
                                                 Acc<float> &acc, bool pair, bool bond) {
    const int ti = __float_as_int(ni.w);
    const AnchorI ai = unpack(__float_as_uint(xi.w), ag);
    if (pair) {
        const int cnt = (int)nnbr[i];
        const uint32_t *nb = nbr + i;
        for (int k = 0; k < cnt; ++k) {
            const uint32_t j = nb[(long long)k * stride];
            const float4 xj = xq[j];
            const float4 nj = nq[j];
            const AnchorI aj = unpack(__float_as_uint(xj.w), ag);
            const float dx = add_rn(sub_rn(xi.x, xj.x), ag.h * (float)(ai.x - aj.x));
            const float dy = add_rn(sub_rn(xi.y, xj.y), ag.h * (float)(ai.y - aj.y));
            const float dz = add_rn(sub_rn(xi.z, xj.z), ag.h * (float)(ai.z - aj.z));
            const float r2 = r2_rn(dx, dy, dz);
            pair_ij<float, ENERGY>(ti, __float_as_int(nj.w), dx, dy, dz, r2, ni.x, ni.y, ni.z, nj.x, nj.y, nj.z,
                                   acc);
        }
    }
    if (bond && nbnd) {
        const int nb = nbnd[i];
        for (int k = 0; k < nb; ++k) {
            const uint32_t e = bnd[(long long)k * stride + i];
            const uint32_t j = e & BOND_INDEX_MASK;
            const float4 xj = xq[j];
            const AnchorI aj = unpack(__float_as_uint(xj.w), ag);
            const float dx = (xj.x - xi.x) + ag.h * (float)(aj.x - ai.x);
            const float dy = (xj.y - xi.y) + ag.h * (float)(aj.y - ai.y);
            const float dz = (xj.z - xi.z) + ag.h * (float)(aj.z - ai.z);
            bond_ij<float, ENERGY>((int)(e >> 30), dx, dy, dz, acc);
        }
    }
}

template <bool ENERGY>
__device__ __forceinline__ void accumulate_fp64(long long i, long long stride, const double *x, const double *nrm,
                                                const uint8_t *type, const uint32_t *nbr, const uint32_t *nnbr,
                                                const uint32_t *bnd, const uint8_t *nbnd, Acc<double> &acc, bool pair,
                                                bool bond) {
    const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
    if (pair) {
        const double nix = nrm[3 * i], niy = nrm[3 * i + 1], niz = nrm[3 * i + 2];
        const int ti = type[i];
        const int cnt = (int)nnbr[i];
        const uint32_t *nb = nbr + i;
        for (int k = 0; k < cnt; ++k) {
            const long long j = nb[(long long)k * stride];
            const double dx = sub_rn(xi, x[3 * j]);
            const double dy = sub_rn(yi, x[3 * j + 1]);
            const double dz = sub_rn(zi, x[3 * j + 2]);
            const double r2 = r2_rn(dx, dy, dz);
            pair_ij<double, ENERGY>(ti, type[j], dx, dy, dz, r2, nix, niy, niz, nrm[3 * j], nrm[3 * j + 1],
                                    nrm[3 * j + 2], acc);
        }
    }
    if (bond && nbnd) {
        const int nb = nbnd[i];
        for (int k = 0; k < nb; ++k) {
            const uint32_t e = bnd[(long long)k * stride + i];
            const long long j = e & BOND_INDEX_MASK;
            bond_ij<double, ENERGY>((int)(e >> 30), x[3 * j] - xi, x[3 * j + 1] - yi, x[3 * j + 2] - zi, acc);
        }
    }
}

// ---- standalone force evaluation (compute_pairwise_fused / compute_bonded) --
template <bool MIXED, bool ENERGY>
__global__ __launch_bounds__(BLOCK) void k_forces(ForceArgs a, int pair, int bond) {
    const long long i = block_index(a.swizzle) * BLOCK + threadIdx.x;
    if (i >= a.n) return;
    if (MIXED) {
        Acc<float> acc;
        accumulate_mixed<ENERGY>(i, a.stride, a.xq, a.nq, a.nbr, a.nnbr, a.bnd, a.nbnd, a.anc_grid, a.xq[i], a.nq[i],
                                 acc, pair, bond);
        a.f32[3 * i] += acc.fx; a.f32[3 * i + 1] += acc.fy; a.f32[3 * i + 2] += acc.fz;
        a.t32[3 * i] += acc.tx; a.t32[3 * i + 1] += acc.ty; a.t32[3 * i + 2] += acc.tz;
        if (ENERGY) { a.e[i] += acc.e; a.w[i] += acc.w; }
    } else {
        Acc<double> acc;
        accumulate_fp64<ENERGY>(i, a.stride, a.x, a.nrm, a.type, a.nbr, a.nnbr, a.bnd, a.nbnd, acc, pair, bond);
        a.f64[3 * i] += acc.fx; a.f64[3 * i + 1] += acc.fy; a.f64[3 * i + 2] += acc.fz;
        a.t64[3 * i] += acc.tx; a.t64[3 * i + 1] += acc.ty; a.t64[3 * i + 2] += acc.tz;
        if (ENERGY) { a.e[i] += acc.e; a.w[i] += acc.w; }
    }
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
            const float4 xi = a.xq[i], xj = a.xq[j];
            const AnchorI ai = unpack(__float_as_uint(xi.w), a.anc_grid), aj = unpack(__float_as_uint(xj.w), a.anc_grid);
            const float dx = add_rn(sub_rn(xi.x, xj.x), a.anc_grid.h * (float)(ai.x - aj.x));
            const float dy = add_rn(sub_rn(xi.y, xj.y), a.anc_grid.h * (float)(ai.y - aj.y));
            const float dz = add_rn(sub_rn(xi.z, xj.z), a.anc_grid.h * (float)(ai.z - aj.z));
            const float r2 = r2_rn(dx, dy, dz);
            const int kk = __float_as_int(a.nq[i].w) * NT + __float_as_int(a.nq[j].w);
            hit = r2 > 1e-5f && (r2 < c_fff.cut2_poly[kk] || r2 < c_fff.cut2_lj[kk]);
        } else {
            const double dx = sub_rn(a.x[3 * i], a.x[3 * (long long)j]);
            const double dy = sub_rn(a.x[3 * i + 1], a.x[3 * (long long)j + 1]);
            const double dz = sub_rn(a.x[3 * i + 2], a.x[3 * (long long)j + 2]);
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

__device__ inline float wave_max(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

// ---- fused timestep: forces + bonds + verlet_langevin ---------------------
// src/openrbc.cpp:219-233 (compute_pairwise_fused, compute_bonded,
// integrate(verlet_langevin)) in one pass; the Langevin update restates
// src/integrate_langevin.h:275-297 in FP64 with counter-based noise.
template <bool MIXED>
__global__ __launch_bounds__(BLOCK) void k_step_langevin(StepArgs a) {
    if (a.skip && *a.skip) return;
    const long long i = block_index(a.swizzle) * BLOCK + threadIdx.x;
    const bool act = i < a.n;
    float d2 = 0.f;
    bool bad = false;
    if (act) {
        double fx, fy, fz, tx, ty, tz;
        float4 xi4, ni4;
        int ti;
        uint32_t wi;
        if (MIXED) {
            xi4 = a.xq_in[i];
            ni4 = a.nq_in[i];
            Acc<float> acc;
            accumulate_mixed<false>(i, a.stride, a.xq_in, a.nq_in, a.nbr, a.nnbr, a.bnd, a.nbnd, a.anc_grid, xi4,
                                    ni4, acc, true, a.has_bonds != 0);
            fx = acc.fx; fy = acc.fy; fz = acc.fz; tx = acc.tx; ty = acc.ty; tz = acc.tz;
            ti = __float_as_int(ni4.w);
            wi = __float_as_uint(xi4.w);
        } else {
            Acc<double> acc;
            accumulate_fp64<false>(i, a.stride, a.x_in, a.n_in, a.type, a.nbr, a.nnbr, a.bnd, a.nbnd, acc, true,
                                   a.has_bonds != 0);
            fx = acc.fx; fy = acc.fy; fz = acc.fz; tx = acc.tx; ty = acc.ty; tz = acc.tz;
            ti = a.type[i];
            wi = a.anc[i];
        }
        if (a.f_ext) { fx += a.f_ext[3 * i]; fy += a.f_ext[3 * i + 1]; fz += a.f_ext[3 * i + 2]; }
        // rotation: t = n x t; o += dt t; n = normalize(n + (o x n) dt)
        const double nx = a.n_in[3 * i], ny = a.n_in[3 * i + 1], nz = a.n_in[3 * i + 2];
        double ox = a.o[3 * i], oy = a.o[3 * i + 1], oz = a.o[3 * i + 2];
        const double dt = a.dt;
        const double qx = ny * tz - nz * ty, qy = nz * tx - nx * tz, qz = nx * ty - ny * tx;
        ox += dt * qx; oy += dt * qy; oz += dt * qz;
        double gx = nx + (oy * nz - oz * ny) * dt;
        double gy = ny + (oz * nx - ox * nz) * dt;
        double gz = nz + (ox * ny - oy * nx) * dt;
        const double gs = 1.0 / sqrt(gx * gx + gy * gy + gz * gz);
        gx *= gs; gy *= gs; gz *= gs;
        // translation with uniform [-1,1) noise (src/integrate_langevin.h:282-296)
        uint32_t c0 = a.id[i], c1 = (uint32_t)a.step, c2 = (uint32_t)(a.step >> 32), c3 = 0u;
        philox4x32_10(c0, c1, c2, c3, (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
        const double u = 4.656612873077392578125e-10;
        const double rx = (double)c0 * u - 1.0, ry = (double)c1 * u - 1.0, rz = (double)c2 * u - 1.0;
        double vx = a.v[3 * i], vy = a.v[3 * i + 1], vz = a.v[3 * i + 2];
        const double gam = a.gamma[ti], sig = a.sigma[ti];
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
        const AnchorI ai = unpack(wi, a.anc_grid);
        const float r0 = __double2float_rn(px - a.anc_grid.hd * (double)(ai.x + a.anc_grid.A0[0]));
        const float r1 = __double2float_rn(py - a.anc_grid.hd * (double)(ai.y + a.anc_grid.A0[1]));
        const float r2 = __double2float_rn(pz - a.anc_grid.hd * (double)(ai.z + a.anc_grid.A0[2]));
        if (MIXED) {
            a.xq_out[i] = make_float4(r0, r1, r2, __uint_as_float(wi));
            a.nq_out[i] = make_float4(__double2float_rn(gx), __double2float_rn(gy), __double2float_rn(gz),
                                      __int_as_float(ti));
        }
        const float e0 = r0 - a.rb[3 * i], e1 = r1 - a.rb[3 * i + 1], e2 = r2 - a.rb[3 * i + 2];
        d2 = e0 * e0 + e1 * e1 + e2 * e2;
        bad = !(isfinite(px) && isfinite(py) && isfinite(pz) && isfinite(gx));
        if (bad) d2 = 3.0e38f;
    }
    d2 = wave_max(d2);
    if (__lane_id() == 0 && d2 > 0.f) atomicMax(a.disp_max, __float_as_uint(d2));
    if (bad) atomicOr(a.bad, 1u);
    if (a.done && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(a.done, 1);
}

// ---- fused forces + post_toque_final_update (NH second half) --------------
// src/integrate_nh.h:237-273 after compute_pairwise_fused/compute_bonded:
// t = n x t; v += dt/2 (f/m - zeta v); o += dt/2 t; KE partial per block.
// f and the torque are stored for the next first half.
template <bool MIXED>
__global__ __launch_bounds__(BLOCK) void k_step_nh_second(StepArgs a, double *f64, double *t64,
                                                          double *ke_partial, const double *zeta_p) {
    const long long i = block_index(a.swizzle) * BLOCK + threadIdx.x;
    const bool act = i < a.n;
    double ke = 0.0;
    if (act) {
        double fx, fy, fz, tx, ty, tz;
        int ti;
        if (MIXED) {
            Acc<float> acc;
            accumulate_mixed<false>(i, a.stride, a.xq_in, a.nq_in, a.nbr, a.nnbr, a.bnd, a.nbnd, a.anc_grid,
                                    a.xq_in[i], a.nq_in[i], acc, true, a.has_bonds != 0);
            fx = acc.fx; fy = acc.fy; fz = acc.fz; tx = acc.tx; ty = acc.ty; tz = acc.tz;
            ti = __float_as_int(a.nq_in[i].w);
        } else {
            Acc<double> acc;
            accumulate_fp64<false>(i, a.stride, a.x_in, a.n_in, a.type, a.nbr, a.nnbr, a.bnd, a.nbnd, acc, true,
                                   a.has_bonds != 0);
            fx = acc.fx; fy = acc.fy; fz = acc.fz; tx = acc.tx; ty = acc.ty; tz = acc.tz;
            ti = a.type[i];
        }
        if (a.f_ext) { fx += a.f_ext[3 * i]; fy += a.f_ext[3 * i + 1]; fz += a.f_ext[3 * i + 2]; }
        const double zeta = *zeta_p, dt = a.dt;
        const double nx = a.n_in[3 * i], ny = a.n_in[3 * i + 1], nz = a.n_in[3 * i + 2];
        const double qx = ny * tz - nz * ty, qy = nz * tx - nx * tz, qz = nx * ty - ny * tx;
        const double im = c_ffd.inv_mass[ti];
        double vx = a.v[3 * i], vy = a.v[3 * i + 1], vz = a.v[3 * i + 2];
        vx += 0.5 * dt * (fx * im - zeta * vx);
        vy += 0.5 * dt * (fy * im - zeta * vy);
        vz += 0.5 * dt * (fz * im - zeta * vz);
        a.v[3 * i] = vx; a.v[3 * i + 1] = vy; a.v[3 * i + 2] = vz;
        a.o[3 * i] += 0.5 * dt * qx; a.o[3 * i + 1] += 0.5 * dt * qy; a.o[3 * i + 2] += 0.5 * dt * qz;
        f64[3 * i] = fx; f64[3 * i + 1] = fy; f64[3 * i + 2] = fz;
        t64[3 * i] = qx; t64[3 * i + 1] = qy; t64[3 * i + 2] = qz;
        ke = 0.5 * c_ffd.mass[ti] * (vx * vx + vy * vy + vz * vz);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) ke += __shfl_xor(ke, off, 64);
    __shared__ double ws[BLOCK / 64];
    if (__lane_id() == 0) ws[threadIdx.x >> 6] = ke;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int w = 0; w < BLOCK / 64; ++w) s += ws[w];
        ke_partial[blockIdx.x] = s;
    }
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
    if (mixed) hipLaunchKernelGGL(k_step_langevin<true>, g, b, 0, s, a);
    else hipLaunchKernelGGL(k_step_langevin<false>, g, b, 0, s, a);
}

void launch_step_nh_second(const StepArgs &a, bool mixed, double *f64, double *t64, double *ke_partial,
                           const double *zeta, hipStream_t s) {
    if (a.n <= 0) return;
    const dim3 g(nblk(a.n)), b(BLOCK);
    if (mixed) hipLaunchKernelGGL(k_step_nh_second<true>, g, b, 0, s, a, f64, t64, ke_partial, zeta);
    else hipLaunchKernelGGL(k_step_nh_second<false>, g, b, 0, s, a, f64, t64, ke_partial, zeta);
}

void launch_pairs_export(const ForceArgs &a, bool mixed, unsigned *ids_i, unsigned *ids_j, unsigned long long *count,
                         long long cap, hipStream_t s) {
    if (a.n <= 0) return;
    const dim3 g(nblk(a.n)), b(BLOCK);
    if (mixed) hipLaunchKernelGGL(k_pairs_export<true>, g, b, 0, s, a, ids_i, ids_j, count, cap);
    else hipLaunchKernelGGL(k_pairs_export<false>, g, b, 0, s, a, ids_i, ids_j, count, cap);
}

}  // namespace orbc

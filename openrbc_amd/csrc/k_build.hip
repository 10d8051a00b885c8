// k_build.hip -- neighbour rebuild on gfx950: bounding box, Morton keys of
// the binning grid, permutation of the particle store, cell ranges, the
// Verlet list and the per-particle bond lists.
//
// Replaces, for the hot loop, the reference's Voronoi/k-d machinery
// (src/voronoi.h:77-237, src/kdtree.h:116-285, src/reorder.h:30-149,
// src/reorder_morton.h:44-122): particles are kept in Morton order of a
// uniform grid (cell edge >= lipid cutoff + skin), every particle lists all
// partners within rc(ti,tj) + skin, so the force kernel sees every pair the
// cutoff admits (the reference's centroid stencil can miss some; the
// Voronoi filter below reproduces its set when asked to).
#include <algorithm>

#include "orbc_internal.h"

namespace orbc {

namespace {

__device__ inline float normsq_unfused(float a, float b, float c) {
    return __fadd_rn(__fadd_rn(__fmul_rn(a, a), __fmul_rn(b, b)), __fmul_rn(c, c));
}

__constant__ float c_list_cut2[36];   // (rc(a,b) + skin)^2, <= 0 if a,b never interact
__constant__ int c_rings[NT];         // grid rings a type searches
__constant__ float c_pair_cut[36];    // interaction cutoff of a type pair (0: none)
__constant__ int c_budget_rings[NT];  // rings searched for the skin budget of an isolated particle (0: none)

__global__ __launch_bounds__(BLOCK) void k_bbox_partial(const double *x, long long n, double *partial) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (long long i = (long long)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (long long)gridDim.x * BLOCK) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double v = x[3 * i + d];
            lo[d] = fmin(lo[d], v);
            hi[d] = fmax(hi[d], v);
        }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        for (int off = 32; off > 0; off >>= 1) {
            lo[d] = fmin(lo[d], __shfl_xor(lo[d], off, 64));
            hi[d] = fmax(hi[d], __shfl_xor(hi[d], off, 64));
        }
    }
    __shared__ double s[BLOCK / 64][6];
    const int wave = threadIdx.x >> 6;
    if (__lane_id() == 0)
        for (int d = 0; d < 3; ++d) { s[wave][d] = lo[d]; s[wave][3 + d] = hi[d]; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < BLOCK / 64; ++w)
            for (int d = 0; d < 3; ++d) {
                s[0][d] = fmin(s[0][d], s[w][d]);
                s[0][3 + d] = fmax(s[0][3 + d], s[w][3 + d]);
            }
        for (int d = 0; d < 6; ++d) partial[6 * blockIdx.x + d] = s[0][d];
    }
}

struct WrapArgs {
    double lo[3], L[3];
    int per[3];
};
__global__ __launch_bounds__(BLOCK) void k_wrap(double *x, long long n, WrapArgs w) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    for (int d = 0; d < 3; ++d)
        if (w.per[d]) {
            const double u = x[3 * i + d] - w.lo[d];
            x[3 * i + d] = w.lo[d] + (u - w.L[d] * floor(u / w.L[d]));
        }
}

// one wavefront: lanes stride over the block partials, then a shuffle reduction
__global__ void k_bbox_final(const double *partial, int nb, double *out) {
    double r[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
    for (int b = threadIdx.x; b < nb; b += 64)
        for (int d = 0; d < 3; ++d) {
            r[d] = fmin(r[d], partial[6 * b + d]);
            r[3 + d] = fmax(r[3 + d], partial[6 * b + 3 + d]);
        }
    for (int off = 32; off > 0; off >>= 1)
        for (int d = 0; d < 3; ++d) {
            r[d] = fmin(r[d], __shfl_xor(r[d], off, 64));
            r[3 + d] = fmax(r[3 + d], __shfl_xor(r[3 + d], off, 64));
        }
    if (threadIdx.x == 0)
        for (int d = 0; d < 6; ++d) out[d] = r[d];
}

__device__ inline int cell_coord(double v, double lo, double cs, int dim) {
    int c = (int)floor((v - lo) / cs);
    return c < 0 ? 0 : (c >= dim ? dim - 1 : c);
}

__device__ inline uint32_t morton_var(int cx, int cy, int cz, int bx, int by, int bz) {
    uint32_t key = 0;
    int pos = 0;
    const int mb = max(bx, max(by, bz));
    for (int b = 0; b < mb; ++b) {
        if (b < bx) key |= (uint32_t)((cx >> b) & 1) << pos++;
        if (b < by) key |= (uint32_t)((cy >> b) & 1) << pos++;
        if (b < bz) key |= (uint32_t)((cz >> b) & 1) << pos++;
    }
    return key;
}

__global__ __launch_bounds__(BLOCK) void k_keys(const double *x, long long n, Grid g, uint32_t *keys,
                                                uint32_t *vals) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const int cx = cell_coord(x[3 * i], g.lo[0], g.csd[0], g.dim[0]);
    const int cy = cell_coord(x[3 * i + 1], g.lo[1], g.csd[1], g.dim[1]);
    const int cz = cell_coord(x[3 * i + 2], g.lo[2], g.csd[2], g.dim[2]);
    keys[i] = morton_var(cx, cy, cz, g.bits[0], g.bits[1], g.bits[2]);
    vals[i] = (uint32_t)i;
}

// AoS triples: one thread per component, so three consecutive threads read
// one particle's triple and every store instruction writes a contiguous run
__global__ __launch_bounds__(BLOCK) void k_permute3(PermuteArgs a) {
    const long long e = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (e >= 3 * a.n) return;
    const long long i = e / 3;
    const long long src = 3 * (long long)a.perm[i] + (e - 3 * i);
    for (int k = 0; k < a.n3; ++k) a.dst3[k][e] = a.src3[k][src];
    for (int k = 0; k < a.nf3; ++k) a.dstf3[k][e] = a.srcf3[k][src];
}

// per-particle fields
__global__ __launch_bounds__(BLOCK) void k_permute(PermuteArgs a) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= a.n) return;
    const long long p = a.perm[i];
    a.type_dst[i] = a.type_src[p];
    a.id_dst[i] = a.id_src[p];
    a.slot_dst[i] = a.slot_src[p];
    if (a.aff_src) a.aff_dst[i] = a.aff_src[p];
    if (a.grp_src) a.grp_dst[i] = a.grp_src[p];
}

// ---- removal of particles (stray-lipid cleanup) -------------------------------
// flag[slot] = 1 for the deleted particles (flag zeroed on entry); keep[i] = !del[i]
__global__ __launch_bounds__(BLOCK) void k_del_flags(long long n, const uint32_t *slot, const uint32_t *del,
                                                     uint32_t *flag, uint32_t *keep) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    if (del[i]) flag[slot[i]] = 1u;
    keep[i] = del[i] ? 0u : 1u;
}

// kept particle i goes to pos[i]; its slot, if below pslot0 (a lipid), closes
// the gaps of the deleted slots below it (slot_pre: exclusive scan of flag)
__global__ __launch_bounds__(BLOCK) void k_compact(long long n, const uint32_t *del, const uint32_t *pos,
                                                   uint32_t *perm, uint32_t *slot, const uint32_t *slot_pre,
                                                   uint32_t pslot0) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n || del[i]) return;
    perm[pos[i]] = (uint32_t)i;
    const uint32_t sl = slot[i];
    if (sl < pslot0) slot[i] = sl - slot_pre[sl];
}

__global__ __launch_bounds__(BLOCK) void k_inverse(const uint32_t *slot, uint32_t *inv, long long n) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) inv[slot[i]] = (uint32_t)i;
}

__global__ __launch_bounds__(BLOCK) void k_refresh(long long n, const double *x, const double *nrm,
                                                   const uint8_t *type, Anchor a, PQ *pq, float *rb,
                                                   uint32_t *anc) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    long long A[3];
    float r[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double v = x[3 * i + d];
        long long q = (long long)floor((v - a.lo[d]) / a.hd[d]) - a.A0[d];
        if (a.wrap[d]) {
            // periodic axis: x is wrapped into the box; guard the last anchor against rounding
            const long long M = 1LL << (32 - a.wrap[d]);
            q = q < 0 ? 0 : (q > M - 1 ? M - 1 : q);
        }
        A[d] = q;
        r[d] = __double2float_rn(v - anchor_pos(a, d, q));
    }
    const uint32_t w = (uint32_t)A[0] | ((uint32_t)A[1] << a.shy) | ((uint32_t)A[2] << a.shz);
    anc[i] = w;
    rb[3 * i] = r[0];
    rb[3 * i + 1] = r[1];
    rb[3 * i + 2] = r[2];
    if (pq) {
        PQ q;
        q.x = make_float4(r[0], r[1], r[2], __uint_as_float(w));
        q.n = make_float4(__double2float_rn(nrm[3 * i]), __double2float_rn(nrm[3 * i + 1]),
                          __double2float_rn(nrm[3 * i + 2]), __int_as_float((int)type[i]));
        pq[i] = q;
    }
}

__device__ inline long long cell_lin(int cx, int cy, int cz, const Grid &g) {
    return ((long long)cz * g.dim[1] + cy) * g.dim[0] + cx;
}

// cell coordinate c + dc along axis d: wraps on a periodic axis; false if it
// leaves an open axis
__device__ inline bool cell_step(const Grid &g, int d, int c, int dc, int &out) {
    int v = c + dc;
    if (g.per[d]) {
        v %= g.dim[d];
        if (v < 0) v += g.dim[d];
    } else if (v < 0 || v >= g.dim[d]) {
        return false;
    }
    out = v;
    return true;
}

__global__ __launch_bounds__(BLOCK) void k_cell_ranges(const double *x, long long n, Grid g, uint32_t *cstart,
                                                       uint32_t *cend) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    auto lin = [&](long long k) {
        return cell_lin(cell_coord(x[3 * k], g.lo[0], g.csd[0], g.dim[0]), cell_coord(x[3 * k + 1], g.lo[1], g.csd[1], g.dim[1]),
                        cell_coord(x[3 * k + 2], g.lo[2], g.csd[2], g.dim[2]), g);
    };
    const long long c = lin(i);
    const long long cp = i > 0 ? lin(i - 1) : -1;
    const long long cn = i + 1 < n ? lin(i + 1) : -1;
    if (c != cp) cstart[c] = (uint32_t)i;
    if (c != cn) cend[c] = (uint32_t)(i + 1);
}

// ---- Verlet list -------------------------------------------------------
// Candidate test in the FP32 anchor geometry of the force kernel (a superset
// test: r^2 < (rc + skin)^2), plus the Voronoi stencil filter when asked.
struct NlCtx {
    const NlistArgs &a;
    long long i;
    int ti;
    float4 xi;
    int ax, ay, az;
    float c0, c1, c2;
    int cnt;
};

__device__ inline bool nl_accept(const NlCtx &c, uint32_t j, float4 xj, int tj, const float *cut2tab) {
    if (j == (uint32_t)c.i) return false;
    const float cut2 = cut2tab[c.ti * NT + tj];
    if (!(cut2 > 0.f)) return false;
    const Anchor &g = c.a.anc;
    const uint32_t w = __float_as_uint(xj.w);
    const float dx = (c.xi.x - xj.x) + anchor_dx(g, 0, c.ax - (int)(w & g.mx));
    const float dy = (c.xi.y - xj.y) + anchor_dx(g, 1, c.ay - (int)((w >> g.shy) & g.my));
    const float dz = (c.xi.z - xj.z) + anchor_dx(g, 2, c.az - (int)(w >> g.shz));
    if (!(dx * dx + dy * dy + dz * dz < cut2)) return false;
    if (c.a.aff) {
        // reference stencil radii: pp 9, pl 8, ll 6 (src/compute_pairwise_fused.h:141-284)
        const int cb = c.a.aff[j];
        const float d2 = normsq_unfused(__fsub_rn(c.a.centroids[3 * cb], c.c0), __fsub_rn(c.a.centroids[3 * cb + 1], c.c1),
                                        __fsub_rn(c.a.centroids[3 * cb + 2], c.c2));
        const int nl = (c.ti == 0) + (tj == 0);
        const float Rs = nl == 2 ? 6.0f : (nl == 1 ? 8.0f : 9.0f);
        if (!(d2 < Rs * Rs)) return false;
    }
    return true;
}

constexpr int NL_CAP = 256;          // candidates staged per wave per pass

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline int wave_max_i(int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

// Cells at Chebyshev distance 2..R around (cx,cy,cz), scanned by the whole
// wave for ONE particle (held in c on every lane): 64 cells at a time have
// their ranges prefix-summed across the lanes, their particles are staged in
// LDS and tested 64 per round, hits appended after c.cnt in candidate order.
// Only band-3 / actin (the sigma-3.4 LJ pair) reach past the first ring, so
// this runs for a few percent of the particles.
__device__ inline void nl_wave_shell(NlCtx &c, int cx, int cy, int cz, int R, float4 *sx, uint32_t *sj,
                                     uint32_t *spref, uint32_t *scst, const float *cut2tab) {
    const Grid &g = c.a.g;
    const int lane = __lane_id();
    const int D = 2 * R + 1, ncube = D * D * D;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int base = 0; base < ncube; base += 64) {
        const int k = base + lane;
        uint32_t st = 0, cntc = 0;
        if (k < ncube) {
            const int dx = k % D - R, dy = (k / D) % D - R, dz = k / (D * D) - R;
            int xx, yy, zz;
            if (max(abs(dx), max(abs(dy), abs(dz))) >= 2 && cell_step(g, 0, cx, dx, xx) && cell_step(g, 1, cy, dy, yy) &&
                cell_step(g, 2, cz, dz, zz)) {
                const long long cl = cell_lin(xx, yy, zz, g);
                st = c.a.cstart[cl];
                cntc = c.a.cend[cl] - st;
            }
        }
        uint32_t inc = cntc;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(inc, off, 64);
            if (lane >= off) inc += y;
        }
        const uint32_t total = __shfl(inc, 63, 64);
        wave_sync();
        spref[lane + 1] = inc;
        scst[lane] = st;
        if (lane == 0) spref[0] = 0;
        wave_sync();
        for (uint32_t off = 0; off < total; off += NL_CAP) {
            const uint32_t m = min((uint32_t)NL_CAP, total - off);
            wave_sync();
            for (uint32_t t = lane; t < m; t += 64) {
                const uint32_t gidx = off + t;
                int lo = 0, hi = 64;                // largest k with spref[k] <= gidx
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (spref[mid] <= gidx) lo = mid; else hi = mid;
                }
                const uint32_t j = scst[lo] + (gidx - spref[lo]);
                sx[t] = c.a.pq[j].x;
                sj[t] = j | ((uint32_t)c.a.type[j] << 29);
            }
            wave_sync();
            for (uint32_t t0 = 0; t0 < m; t0 += 64) {
                const uint32_t t = t0 + lane;
                bool hit = false;
                uint32_t j = 0;
                if (t < m) {
                    const uint32_t jt = sj[t];
                    j = jt & 0x1fffffffu;
                    hit = nl_accept(c, j, sx[t], (int)(jt >> 29), cut2tab);
                }
                const unsigned long long hm = __ballot(hit);
                if (hit) {
                    const int pos = c.cnt + __popcll(hm & lt);
                    if (pos < c.a.kmax) c.a.nbr[(long long)pos * c.a.stride + c.i] = j;
                }
                c.cnt += __popcll(hm);
            }
        }
    }
}

// first particle of every occupied cell (particles are sorted by cell)
__global__ __launch_bounds__(BLOCK) void k_cell_flags(const double *x, long long n, Grid g, uint32_t *flag) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    auto lin = [&](long long k) {
        return cell_lin(cell_coord(x[3 * k], g.lo[0], g.csd[0], g.dim[0]), cell_coord(x[3 * k + 1], g.lo[1], g.csd[1], g.dim[1]),
                        cell_coord(x[3 * k + 2], g.lo[2], g.csd[2], g.dim[2]), g);
    };
    flag[i] = (i == 0 || lin(i) != lin(i - 1)) ? 1u : 0u;
}

__global__ __launch_bounds__(BLOCK) void k_cell_compact(const uint32_t *flag, const uint32_t *pos, long long n,
                                                        uint32_t *occ) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n && flag[i]) occ[pos[i]] = (uint32_t)i;
}

// One wavefront per occupied grid cell.  The particles of the cell's 3x3x3
// neighbourhood are staged in LDS; the 64 lanes are split into (particle,
// stripe) pairs -- P = cell particles rounded up to a power of two, S = 64/P
// stripes -- and each round tests S consecutive candidates for every
// particle.  Hits are appended in candidate order with a ballot-based
// compaction, so the list order is deterministic.  The test is branch-free:
// a candidate's index/type word and its position are read from LDS together,
// and the list cutoff comes from the particle type's row held in registers.
// Particle types whose cutoff + skin exceeds one cell (band-3 / actin with the
// sigma-3.4 LJ) get the outer shells appended by k_nlist_far afterwards.
// candidate rounds unrolled in the list build (LDS reads of the next round
// issued before this round's ballot)
#ifndef ORBC_NL_UNROLL
#define ORBC_NL_UNROLL 1
#endif
#define ORBC_STR(x) #x
#define ORBC_PRAGMA(x) _Pragma(ORBC_STR(x))
__global__ __launch_bounds__(BLOCK) void k_nlist(NlistArgs a, const uint32_t *occ, long long nocc) {
    __shared__ float4 s_x[BLOCK / 64][NL_CAP];
    __shared__ uint32_t s_pref[BLOCK / 64][28];
    __shared__ uint32_t s_cst[BLOCK / 64][27];
    __shared__ float s_cut2[36];
    if (threadIdx.x < 36) s_cut2[threadIdx.x] = c_list_cut2[threadIdx.x];
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = __lane_id();
    // XCD-aware remap: consecutive cells (whose list rows share cache lines)
    // are built by blocks of the same XCD, so the lines complete in one L2
    const unsigned nbk = gridDim.x, q = nbk >> 3, rr = nbk & 7u, xcd = blockIdx.x & 7u, kk = blockIdx.x >> 3;
    const long long lb = (xcd < rr) ? (long long)xcd * (q + 1) + kk : (long long)rr * (q + 1) + (long long)(xcd - rr) * q + kk;
    const long long cell = lb * (BLOCK / 64) + w;
    if (cell >= nocc) return;                       // whole wave; no block barrier below
    const Grid &g = a.g;
    const long long p0 = occ[cell], p1 = (cell + 1 < nocc) ? occ[cell + 1] : a.n;
    const int cx = cell_coord(a.x[3 * p0], g.lo[0], g.csd[0], g.dim[0]);
    const int cy = cell_coord(a.x[3 * p0 + 1], g.lo[1], g.csd[1], g.dim[1]);
    const int cz = cell_coord(a.x[3 * p0 + 2], g.lo[2], g.csd[2], g.dim[2]);
    // ranges of the 27 neighbour cells, prefix counts
    uint32_t st = 0, cntc = 0;
    if (lane < 27) {
        int xx, yy, zz;
        if (cell_step(g, 0, cx, lane % 3 - 1, xx) && cell_step(g, 1, cy, (lane / 3) % 3 - 1, yy) &&
            cell_step(g, 2, cz, lane / 9 - 1, zz)) {
            const long long cl = cell_lin(xx, yy, zz, g);
            st = a.cstart[cl];
            cntc = a.cend[cl] - st;
        }
    }
    uint32_t inc = cntc;
    for (int off = 1; off < 32; off <<= 1) {
        const uint32_t y = __shfl_up(inc, off, 64);
        if (lane >= off) inc += y;
    }
    const uint32_t total = __shfl(inc, 26, 64);
    if (lane < 27) { s_pref[w][lane + 1] = inc; s_cst[w][lane] = st; }
    if (lane == 0) s_pref[w][0] = 0;
    wave_sync();
    const long long np = p1 - p0;
    const Anchor &an = a.anc;
    // positions relative to the anchor of the cell's first particle, formed
    // once per particle / staged candidate: the pair test is then 3 subtracts
    const uint32_t wref = __float_as_uint(a.pq[p0].x.w);
    const int rx = (int)(wref & an.mx), ry = (int)((wref >> an.shy) & an.my), rz = (int)(wref >> an.shz);
    auto rel = [&](float4 q) {
        const uint32_t wq = __float_as_uint(q.w);
        return make_float3(anchor_add(an, 0, q.x, (int)(wq & an.mx) - rx),
                           anchor_add(an, 1, q.y, (int)((wq >> an.shy) & an.my) - ry),
                           anchor_add(an, 2, q.z, (int)(wq >> an.shz) - rz));
    };
    for (long long pc = 0; pc < np; pc += 64) {
        const int npc = (int)min(64LL, np - pc);
        int P = 1;
        while (P < npc) P <<= 1;
        const int S = 64 / P;
        const int pslot = lane / S, stripe = lane % S;
        const bool act = pslot < npc;
        const long long i = p0 + pc + (act ? pslot : 0);
        NlCtx c{a, i, 0, make_float4(0, 0, 0, 0), 0, 0, 0, 0.f, 0.f, 0.f, 0};
        float3 pi = make_float3(0.f, 0.f, 0.f);
        int crow = 0;                       // this particle's row of the cutoff table
        if (act) {
            c.ti = a.type[i];
            c.xi = a.pq[i].x;
            pi = rel(c.xi);
            crow = c.ti * NT;
            if (a.aff) {
                const int ca = a.aff[i];
                c.c0 = a.centroids[3 * ca]; c.c1 = a.centroids[3 * ca + 1]; c.c2 = a.centroids[3 * ca + 2];
            }
        }
        const unsigned long long gmask = (S == 64) ? ~0ull : (((1ull << S) - 1ull) << (pslot * S));
        const unsigned long long below = (stripe == 0) ? 0ull : (((1ull << stripe) - 1ull) << (pslot * S));
        uint32_t *const row = a.nbr + i;
        int cnt = 0;
        for (uint32_t off = 0; off < total; off += NL_CAP) {
            const uint32_t m = min((uint32_t)NL_CAP, total - off);
            wave_sync();
            for (uint32_t t = lane; t < m; t += 64) {
                const uint32_t gidx = off + t;
                int lo = 0, hi = 27;                // largest k with pref[k] <= gidx
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (s_pref[w][mid] <= gidx) lo = mid; else hi = mid;
                }
                const uint32_t j = s_cst[w][lo] + (gidx - s_pref[w][lo]);
                const float3 pj = rel(a.pq[j].x);
                s_x[w][t] = make_float4(pj.x, pj.y, pj.z, __uint_as_float(j | ((uint32_t)a.type[j] << 29)));
            }
            wave_sync();
ORBC_PRAGMA(unroll ORBC_NL_UNROLL)
            for (uint32_t t0 = 0; t0 < m; t0 += S) {
                const uint32_t t = t0 + stripe;
                const bool valid = act && t < m;
                const uint32_t tt = valid ? t : 0u;
                const float4 xj = s_x[w][tt];
                const uint32_t jt = __float_as_uint(xj.w);
                const uint32_t j = jt & 0x1fffffffu;
                const int tj = (int)(jt >> 29);
                const float cut2 = s_cut2[crow + tj];
                const float dx = pi.x - xj.x, dy = pi.y - xj.y, dz = pi.z - xj.z;
                // a pair without interaction has cut2 < 0 and never passes
                bool hit = valid && j != (uint32_t)i && dx * dx + dy * dy + dz * dz < cut2;
                if (a.aff && hit) {
                    // reference stencil radii: pp 9, pl 8, ll 6 (src/compute_pairwise_fused.h:141-284)
                    const int cb = a.aff[j];
                    const float d2 = normsq_unfused(__fsub_rn(a.centroids[3 * cb], c.c0),
                                                    __fsub_rn(a.centroids[3 * cb + 1], c.c1),
                                                    __fsub_rn(a.centroids[3 * cb + 2], c.c2));
                    const int nl = (c.ti == 0) + (tj == 0);
                    const float Rs = nl == 2 ? 6.0f : (nl == 1 ? 8.0f : 9.0f);
                    hit = d2 < Rs * Rs;
                }
                const unsigned long long hm = __ballot(hit);
                if (hit) {
                    const int pos = cnt + __popcll(hm & below);
                    if (pos < a.kmax) row[(size_t)pos * (size_t)a.stride] = j;
                }
                cnt += __popcll(hm & gmask);
            }
        }
        if (act && stripe == 0) a.nnbr[i] = (uint32_t)(cnt < a.kmax ? cnt : a.kmax);
        const int mc = wave_max_i((act && stripe == 0) ? cnt : 0);
        if (lane == 0 && mc > a.kmax) atomicMax(a.max_count, (unsigned)mc);
    }
}

// Outer shells (Chebyshev distance 2..R) for the particles whose type reaches
// past the first ring of cells, appended after the k_nlist entries.  One
// wavefront per 64 particles; each such particle is handled by the whole wave
// (nl_wave_shell).  nnbr holds the first-ring count on entry.
// the particles whose type searches past the first ring, in store order
__global__ __launch_bounds__(BLOCK) void k_far_flags(long long n, const uint8_t *type, uint32_t *flag) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) flag[i] = c_rings[type[i]] >= 2 ? 1u : 0u;
}
__global__ __launch_bounds__(BLOCK) void k_far_compact(long long n, const uint32_t *flag, const uint32_t *pos,
                                                       uint32_t *list, uint32_t *count) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    if (flag[i]) list[pos[i]] = (uint32_t)i;
    if (i == n - 1) *count = pos[i] + flag[i];
}

// list[0, *count): the particles to extend; the waves of the grid take them in
// turn (these proteins cluster in the store, so one wave per 64 consecutive
// particles would serialise them on a few waves)
__global__ __launch_bounds__(BLOCK) void k_nlist_far(NlistArgs a, const uint32_t *list, const uint32_t *count) {
    __shared__ float4 s_x[BLOCK / 64][NL_CAP];
    __shared__ uint32_t s_j[BLOCK / 64][NL_CAP];
    __shared__ uint32_t s_pref2[BLOCK / 64][65];
    __shared__ uint32_t s_cst2[BLOCK / 64][64];
    __shared__ float s_cut2[36];
    if (threadIdx.x < 36) s_cut2[threadIdx.x] = c_list_cut2[threadIdx.x];
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = __lane_id();
    const long long m = *count;
    const long long nw = (long long)gridDim.x * (BLOCK / 64);
    const Grid &g = a.g;
    const Anchor &an = a.anc;
    for (long long q = (long long)blockIdx.x * (BLOCK / 64) + w; q < m; q += nw) {
        const long long i = list[q];                 // wave-uniform
        NlCtx c{a, i, (int)a.type[i], a.pq[i].x, 0, 0, 0, 0.f, 0.f, 0.f, (int)a.nnbr[i]};
        const uint32_t wi = __float_as_uint(c.xi.w);
        c.ax = (int)(wi & an.mx); c.ay = (int)((wi >> an.shy) & an.my); c.az = (int)(wi >> an.shz);
        if (a.aff) {
            const int ca = a.aff[i];
            c.c0 = a.centroids[3 * ca]; c.c1 = a.centroids[3 * ca + 1]; c.c2 = a.centroids[3 * ca + 2];
        }
        const int cx = cell_coord(a.x[3 * i], g.lo[0], g.csd[0], g.dim[0]);
        const int cy = cell_coord(a.x[3 * i + 1], g.lo[1], g.csd[1], g.dim[1]);
        const int cz = cell_coord(a.x[3 * i + 2], g.lo[2], g.csd[2], g.dim[2]);
        nl_wave_shell(c, cx, cy, cz, c_rings[c.ti], s_x[w], s_j[w], s_pref2[w], s_cst2[w], s_cut2);
        if (lane == 0) {
            a.nnbr[i] = (uint32_t)(c.cnt < a.kmax ? c.cnt : a.kmax);
            if (c.cnt > a.kmax) atomicMax(a.max_count, (unsigned)c.cnt);
        }
    }
}

__global__ __launch_bounds__(BLOCK) void k_bondlist(long long n, long long stride, const uint32_t *slot,
                                                    const uint32_t *inv, const uint32_t *bstart,
                                                    const uint32_t *bpartner, const uint8_t *btype, uint32_t *bnd,
                                                    uint8_t *nbnd) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = slot[i];
    const uint32_t b0 = bstart[s], b1 = bstart[s + 1];
    for (uint32_t k = b0; k < b1; ++k)
        bnd[(long long)(k - b0) * stride + i] = inv[bpartner[k]] | ((uint32_t)btype[k] << 30);
    nbnd[i] = (uint8_t)(b1 - b0);
}

// max squared displacement since the last build, from the FP64 state; a
// particle with a skin budget (bud2, k_budget) counts 0 within it and
// "infinite" past it, as in the skin check of the step kernels
__global__ __launch_bounds__(BLOCK) void k_disp(long long n, const double *x, const uint32_t *anc, const float *rb,
                                                Anchor g, const float *bud2, const uint32_t *nnbr,
                                                unsigned *disp_max) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    float d2 = 0.f;
    if (i < n) {
        const uint32_t w = anc[i];
        const long long A[3] = {(long long)(w & g.mx), (long long)((w >> g.shy) & g.my), (long long)(w >> g.shz)};
        for (int d = 0; d < 3; ++d) {
            const float e = __double2float_rn(x[3 * i + d] - anchor_pos(g, d, A[d])) - rb[3 * i + d];
            d2 += e * e;
        }
        if (!isfinite(d2)) d2 = 3.0e38f;
        if (bud2 && nnbr[i] == 0 && bud2[i] > 0.f) d2 = d2 > bud2[i] ? 3.0e38f : 0.f;
    }
    for (int off = 32; off > 0; off >>= 1) d2 = fmaxf(d2, __shfl_xor(d2, off, 64));
    // read before the atomic: only waves that raise the maximum write
    if (__lane_id() == 0 && d2 > 0.f && __float_as_uint(d2) > __hip_atomic_load(disp_max, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMax(disp_max, __float_as_uint(d2));
}

// ---- block halos ------------------------------------------------------------
// One workgroup per particle block b (particles i0 = b * BLOCK ...): the
// distinct list entries outside the block go into an LDS hash set, and the
// order in which they first arrive numbers them: ext[b * HALO_EXT_MAX + e] is
// entry e.  Then each particle's entries are rewritten as 16-bit entries: the
// local index (t < BLOCK: particle i0 + t; BLOCK + e: ext entry e) in bits
// 0-9 and the neighbour's type in bits 10-12, two per word, word m of the
// row at nbr16[m * stride + column] (the column of the lane order below), rows padded to whole groups of
// HALO_GROUP entries with the particle's own local index and type 0
// (distance 0: no contribution).  Which ext slot an entry gets depends on the order of
// the LDS atomics; the values staged at those slots, and so the forces, do
// not.  The set stops taking entries once ext_cap (the halo capacity less
// BLOCK) distinct ones are in, so at most HALO_EXT_MAX + BLOCK of its HALO_HASH slots fill and every probe
// sequence ends at an empty or equal slot.
constexpr int HALO_HASH = 4096;
constexpr uint32_t HALO_EMPTY = 0xffffffffu;
static_assert(HALO_EXT_MAX + BLOCK < HALO_HASH, "halo hash set too small");
static_assert(HALO_CAP == BLOCK + HALO_EXT && HALO_CAP <= HALO_CAP_WIDE && HALO_CAP_WIDE <= 1024, "10-bit halo indices");
static_assert(NT <= 8, "3-bit types in the halo entries");

__device__ inline uint32_t halo_hash(uint32_t j) { return (j * 2654435761u) >> (32 - 12); }   // 12 = log2(HALO_HASH)

__global__ __launch_bounds__(BLOCK) void k_halo(long long n, long long stride, const uint32_t *nbr,
                                                const uint32_t *nnbr, const uint8_t *type, uint32_t *ext,
                                                uint32_t *ext_cnt, uint32_t *nbr16, uint8_t *hord, int ext_cap) {
    __shared__ uint32_t s_key[HALO_HASH];
    __shared__ uint16_t s_val[HALO_HASH];
    __shared__ int s_nn[BLOCK];
    __shared__ int s_cnt, s_over;
    const int t = threadIdx.x;
    const long long b = blockIdx.x, i0 = b * BLOCK, i = i0 + t;
    for (int h = t; h < HALO_HASH; h += BLOCK) s_key[h] = HALO_EMPTY;
    if (t == 0) { s_cnt = 0; s_over = 0; }
    __syncthreads();
    const int nn = i < n ? (int)nnbr[i] : 0;
    const uint32_t self = (uint32_t)i;
    uint32_t *ext_b = ext + b * HALO_EXT_MAX;
    // entries in batches of HB (the loads of a batch in flight together);
    // own-block entries and the padding (the particle itself) are skipped
    constexpr int HB = 8;
    for (int k0 = 0; k0 < nn; k0 += HB) {
        uint32_t jb[HB];
#pragma unroll
        for (int u = 0; u < HB; ++u) jb[u] = (k0 + u < nn) ? nbr[(long long)(k0 + u) * stride + i] : self;
        if (__hip_atomic_load(&s_over, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
#pragma unroll
        for (int u = 0; u < HB; ++u) {
            const uint32_t j = jb[u];
            if ((unsigned long long)((long long)j - i0) < (unsigned long long)BLOCK) continue;
            uint32_t h = halo_hash(j);
            for (int probe = 0; probe < HALO_HASH; ++probe) {
                const uint32_t prev = atomicCAS(&s_key[h], HALO_EMPTY, j);
                if (prev == HALO_EMPTY) {
                    const int e = atomicAdd(&s_cnt, 1);
                    if (e < ext_cap) {
                        s_val[h] = (uint16_t)e;
                        ext_b[e] = j;
                    } else {
                        atomicExch(&s_over, 1);
                    }
                    break;
                }
                if (prev == j) break;
                h = (h + 1) & (HALO_HASH - 1);
            }
        }
    }
    __syncthreads();
    if (s_over) {
        if (t == 0) ext_cnt[b] = HALO_OVERFLOW;
        return;
    }
    if (t == 0) ext_cnt[b] = (uint32_t)s_cnt;
    // lane order of the step kernels: the block's particles by list length,
    // longest first (ties by position; positions past n last), so that each
    // wave walks rows of similar length.  Lane `r` of the block takes particle
    // i0 + hord[i0 + r], whose row is written to column i0 + r.
    s_nn[t] = i < n ? nn : -1;
    __syncthreads();
    const int mine = s_nn[t];
    int rank = 0;
    for (int u = 0; u < BLOCK; ++u) {
        const int v = s_nn[u];
        rank += (v > mine) || (v == mine && u < t);
    }
    hord[i0 + rank] = (uint8_t)t;
    if (i >= n) return;
    const long long col = i0 + rank;        // < n: active particles rank first
    // local indices, HB entries (HB / 2 words) per batch; the row's words
    // [0, halo_row_words(nn)) are written, entries past nn being the particle
    // itself, and nn <= kmax keeps them inside halo_words(kmax)
    const int nw = halo_row_words(nn);
    for (int m0 = 0; m0 < nw; m0 += HB / 2) {
        uint32_t jb[HB], tb[HB];
#pragma unroll
        for (int u = 0; u < HB; ++u) jb[u] = (m0 * 2 + u < nn) ? nbr[(long long)(m0 * 2 + u) * stride + i] : self;
#pragma unroll
        for (int u = 0; u < HB; ++u) tb[u] = (m0 * 2 + u < nn) ? (uint32_t)type[jb[u]] : 0u;
#pragma unroll
        for (int q = 0; q < HB / 2; ++q) {
            if (m0 + q >= nw) break;
            uint32_t word = 0;
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) {
                const uint32_t j = jb[2 * q + hh];
                const long long d = (long long)j - i0;
                uint32_t l = (uint32_t)t;
                if (d >= 0 && d < BLOCK) {
                    l = (uint32_t)d;
                } else {
                    // j is in the set (no overflow): the probe ends at its slot
                    uint32_t h = halo_hash(j);
                    for (int probe = 0; probe < HALO_HASH; ++probe) {
                        if (s_key[h] == j) { l = (uint32_t)BLOCK + s_val[h]; break; }
                        h = (h + 1) & (HALO_HASH - 1);
                    }
                }
                word |= (l | (tb[2 * q + hh] << 10)) << (16 * hh);
            }
            nbr16[(long long)(m0 + q) * stride + col] = word;
        }
    }
}

// Skin budget of isolated particles.  The list stays valid while every pair
// not in it keeps d_i + d_j < r_ij(build) - rc_ij.  A particle with list
// entries gets the usual budget skin/2 (its unlisted pairs start at rc + skin
// or more).  A particle with no entries at all (a lipid ejected from a
// vesicle, drifting away) gets b_i = min_j (r_ij - rc_ij) / 2 over the
// particles j it could interact with in the cells within R rings, capped by
// (R h - max_j rc_ij) / 2 for everything beyond them (h the shortest cell
// edge; a particle outside the cube of R rings is at least R h away): then any
// two particles keep d_i + d_j <= r_ij - rc_ij as long as each stays within
// its own budget.  R is the type's ring count; when that finds nothing closer
// than its cap and few particles are isolated, the cube is widened to
// BUDGET_FAR_RINGS, so a lipid flying off at several cutoffs per ten steps
// keeps the list for tens of steps instead of two.  Every isolated particle
// of one build uses the same widening (it depends only on their count), which
// keeps the pairwise argument above.  Stores b_i^2 (FP32, minus a rounding
// margin) where it exceeds (skin/2)^2, else 0 ("skin/2").  Open boxes only
// (the caller skips periodic ones).
constexpr int BUDGET_FAR_RINGS = 12;
constexpr long long BUDGET_FAR_MAX = 512;   // isolated particles for which the cube is widened
constexpr unsigned BUDGET_GRID = 512;       // blocks of k_budget

__global__ __launch_bounds__(BLOCK) void k_budget_flags(long long n, const uint8_t *type, const uint32_t *nnbr,
                                                        uint32_t *flag, float *bud2) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    flag[i] = (c_budget_rings[type[i]] > 0 && nnbr[i] == 0) ? 1u : 0u;
    bud2[i] = 0.f;
}

struct BudgetSite {
    double x, y, z, cmax;
    int ti, cx, cy, cz;
};

__device__ inline BudgetSite budget_site(long long i, const double *x, const uint8_t *type, const Grid &g) {
    BudgetSite b;
    b.x = x[3 * i]; b.y = x[3 * i + 1]; b.z = x[3 * i + 2];
    b.ti = type[i];
    b.cmax = 0.0;
    for (int tj = 0; tj < NT; ++tj) b.cmax = fmax(b.cmax, (double)c_pair_cut[b.ti * NT + tj]);
    b.cx = cell_coord(b.x, g.lo[0], g.csd[0], g.dim[0]);
    b.cy = cell_coord(b.y, g.lo[1], g.csd[1], g.dim[1]);
    b.cz = cell_coord(b.z, g.lo[2], g.csd[2], g.dim[2]);
    return b;
}

// min over the particles j != i of the cells with cube offset q (0 <= q <
// (2R+1)^3, skipped when within `inner` rings) of (r_ij - rc_ij) / 2
__device__ inline double budget_cell(long long i, const BudgetSite &b, long long q, int R, int inner,
                                     const double *x, const uint8_t *type, const Grid &g, const uint32_t *cstart,
                                     const uint32_t *cend, double best) {
    const int side = 2 * R + 1;
    const int dx = (int)(q % side) - R, dy = (int)((q / side) % side) - R, dz = (int)(q / ((long long)side * side)) - R;
    if (max(abs(dx), max(abs(dy), abs(dz))) <= inner) return best;
    int xx, yy, zz;
    if (!cell_step(g, 0, b.cx, dx, xx) || !cell_step(g, 1, b.cy, dy, yy) || !cell_step(g, 2, b.cz, dz, zz)) return best;
    const long long cl = cell_lin(xx, yy, zz, g);
    for (uint32_t j = cstart[cl]; j < cend[cl]; ++j) {
        if (j == (uint32_t)i) continue;
        const double c = c_pair_cut[b.ti * NT + type[j]];
        if (!(c > 0.0)) continue;
        const double ex = b.x - x[3 * j], ey = b.y - x[3 * j + 1], ez = b.z - x[3 * j + 2];
        best = fmin(best, 0.5 * (sqrt(ex * ex + ey * ey + ez * ez) - c));
    }
    return best;
}

__device__ inline double block_min(double v, double *s_red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    if (__lane_id() == 0) s_red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = s_red[0];
    for (int w = 1; w < BLOCK / 64; ++w) r = fmin(r, s_red[w]);
    __syncthreads();
    return r;
}

// list[0, *count): the isolated particles (k_budget_flags, compacted in index
// order).  Few: one block per particle, the cube of its type's rings and, if
// nothing there is closer than that cube's cap, the cube of BUDGET_FAR_RINGS.
// Many: one thread per particle over its type's rings.
__global__ __launch_bounds__(BLOCK) void k_budget(const double *x, const uint8_t *type, Grid g,
                                                  const uint32_t *cstart, const uint32_t *cend, const uint32_t *list,
                                                  const uint32_t *count, float half_skin, float *bud2) {
    __shared__ double s_red[BLOCK / 64];
    const long long m = *count;
    const double h = fmin(g.csd[0], fmin(g.csd[1], g.csd[2]));
    if (m > BUDGET_FAR_MAX) {
        for (long long q = (long long)blockIdx.x * BLOCK + threadIdx.x; q < m; q += (long long)gridDim.x * BLOCK) {
            const long long i = list[q];
            const BudgetSite b = budget_site(i, x, type, g);
            const int R = c_budget_rings[b.ti];
            const long long side = 2 * R + 1;
            double best = 0.5 * (R * h - b.cmax);
            for (long long k = 0; k < side * side * side; ++k)
                best = budget_cell(i, b, k, R, -1, x, type, g, cstart, cend, best);
            const float r = (float)best - 1e-4f;
            if (r > half_skin) bud2[i] = r * r;
        }
        return;
    }
    for (long long q = blockIdx.x; q < m; q += gridDim.x) {          // block-uniform
        const long long i = list[q];
        const BudgetSite b = budget_site(i, x, type, g);
        const int R = c_budget_rings[b.ti];
        long long side = 2 * R + 1;
        double best = 1e300;
        for (long long k = threadIdx.x; k < side * side * side; k += BLOCK)
            best = budget_cell(i, b, k, R, -1, x, type, g, cstart, cend, best);
        best = block_min(best, s_red);
        double cap = 0.5 * (R * h - b.cmax);
        if (best >= cap && BUDGET_FAR_RINGS > R) {
            // nothing within R rings limits the budget: the cells between R and
            // BUDGET_FAR_RINGS rings
            side = 2 * BUDGET_FAR_RINGS + 1;
            double far = 1e300;
            for (long long k = threadIdx.x; k < side * side * side; k += BLOCK)
                far = budget_cell(i, b, k, BUDGET_FAR_RINGS, R, x, type, g, cstart, cend, far);
            best = fmin(best, block_min(far, s_red));
            cap = 0.5 * (BUDGET_FAR_RINGS * h - b.cmax);
        }
        const float r = (float)fmin(best, cap) - 1e-4f;
        if (threadIdx.x == 0 && r > half_skin) bud2[i] = r * r;
    }
}

inline unsigned nblk(long long n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

}  // namespace

void launch_budget(long long n, const double *x, const uint8_t *type, const uint32_t *nnbr, const Grid &g,
                   const uint32_t *cstart, const uint32_t *cend, float half_skin, float *bud2, const FarScratch &fs,
                   hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_budget_flags, dim3(nblk(n)), dim3(BLOCK), 0, s, n, type, nnbr, fs.flag, bud2);
    exclusive_scan_u32(fs.flag, fs.pos, n, fs.scan_temp, s);
    hipLaunchKernelGGL(k_far_compact, dim3(nblk(n)), dim3(BLOCK), 0, s, n, fs.flag, fs.pos, fs.list, fs.count);
    hipLaunchKernelGGL(k_budget, dim3(BUDGET_GRID), dim3(BLOCK), 0, s, x, type, g, cstart, cend, fs.list, fs.count,
                       half_skin, bud2);
}

void launch_halo(long long n, long long stride, const uint32_t *nbr, const uint32_t *nnbr, const uint8_t *type,
                 uint32_t *ext, uint32_t *cnt, uint32_t *nbr16, uint8_t *hord, int cap, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_halo, dim3(nblk(n)), dim3(BLOCK), 0, s, n, stride, nbr, nnbr, type, ext, cnt, nbr16, hord,
                       cap - BLOCK);
}

void launch_disp(long long n, const double *x, const uint32_t *anc, const float *rb, Anchor g, const float *bud2,
                 const uint32_t *nnbr, unsigned *disp_max, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_disp, dim3(nblk(n)), dim3(BLOCK), 0, s, n, x, anc, rb, g, bud2, nnbr, disp_max);
}

void set_list_cutoffs(const float *cut2_list, const int *rings, const float *pair_cut, const int *budget_rings,
                      hipStream_t s) {
    ORBC_HIP_CHECK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_list_cut2), cut2_list, sizeof(float) * 36, 0,
                                          hipMemcpyHostToDevice, s));
    ORBC_HIP_CHECK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_rings), rings, sizeof(int) * NT, 0,
                                          hipMemcpyHostToDevice, s));
    ORBC_HIP_CHECK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_pair_cut), pair_cut, sizeof(float) * 36, 0,
                                          hipMemcpyHostToDevice, s));
    ORBC_HIP_CHECK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_budget_rings), budget_rings, sizeof(int) * NT, 0,
                                          hipMemcpyHostToDevice, s));
}

void launch_wrap(double *x, long long n, const double lo[3], const double L[3], const int per[3], hipStream_t s) {
    if (n <= 0) return;
    WrapArgs w{};
    for (int d = 0; d < 3; ++d) { w.lo[d] = lo[d]; w.L[d] = L[d]; w.per[d] = per[d]; }
    hipLaunchKernelGGL(k_wrap, dim3(nblk(n)), dim3(BLOCK), 0, s, x, n, w);
}

void launch_bbox(const double *x, long long n, double *out6, hipStream_t s) {
    // out6 must have room for 6 * 257 doubles: partials then result
    const int nb = 256;
    double *partial = out6 + 6;
    hipLaunchKernelGGL(k_bbox_partial, dim3(nb), dim3(BLOCK), 0, s, x, n, partial);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(64), 0, s, partial, nb, out6);
}

void launch_keys(const double *x, long long n, Grid g, uint32_t *keys, uint32_t *vals, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_keys, dim3(nblk(n)), dim3(BLOCK), 0, s, x, n, g, keys, vals);
}

void launch_permute(const PermuteArgs &a, hipStream_t s) {
    if (a.n <= 0) return;
    if (a.n3 > 0 || a.nf3 > 0) hipLaunchKernelGGL(k_permute3, dim3(nblk(3 * a.n)), dim3(BLOCK), 0, s, a);
    hipLaunchKernelGGL(k_permute, dim3(nblk(a.n)), dim3(BLOCK), 0, s, a);
}

void launch_compact(long long n, const uint32_t *del, uint32_t *slot, uint32_t pslot0, uint32_t *flag,
                    uint32_t *slot_pre, uint32_t *keep, uint32_t *pos, uint32_t *perm, void *scan_temp,
                    hipStream_t s) {
    if (n <= 0) return;
    launch_fill_u32(flag, 0u, n, s);
    hipLaunchKernelGGL(k_del_flags, dim3(nblk(n)), dim3(BLOCK), 0, s, n, slot, del, flag, keep);
    exclusive_scan_u32(flag, slot_pre, n, scan_temp, s);
    exclusive_scan_u32(keep, pos, n, scan_temp, s);
    hipLaunchKernelGGL(k_compact, dim3(nblk(n)), dim3(BLOCK), 0, s, n, del, pos, perm, slot, slot_pre, pslot0);
}

void launch_inverse(const uint32_t *slot, uint32_t *inv, long long n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_inverse, dim3(nblk(n)), dim3(BLOCK), 0, s, slot, inv, n);
}

void launch_refresh(long long n, const double *x, const double *nrm, const uint8_t *type, Anchor a, PQ *pq,
                    float *rb, uint32_t *anc, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_refresh, dim3(nblk(n)), dim3(BLOCK), 0, s, n, x, nrm, type, a, pq, rb, anc);
}

void launch_cell_ranges(const double *x, long long n, Grid g, uint32_t *cstart, uint32_t *cend, hipStream_t s) {
    launch_fill_u32(cstart, 0u, g.ncell, s);
    launch_fill_u32(cend, 0u, g.ncell, s);
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cell_ranges, dim3(nblk(n)), dim3(BLOCK), 0, s, x, n, g, cstart, cend);
}

long long launch_cell_list(const double *x, long long n, Grid g, uint32_t *flag, uint32_t *pos, uint32_t *occ,
                           void *scan_temp, hipStream_t s) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_cell_flags, dim3(nblk(n)), dim3(BLOCK), 0, s, x, n, g, flag);
    exclusive_scan_u32(flag, pos, n, scan_temp, s);
    hipLaunchKernelGGL(k_cell_compact, dim3(nblk(n)), dim3(BLOCK), 0, s, flag, pos, n, occ);
    uint32_t last[2];
    ORBC_HIP_CHECK(hipMemcpyAsync(&last[0], pos + n - 1, 4, hipMemcpyDeviceToHost, s));
    ORBC_HIP_CHECK(hipMemcpyAsync(&last[1], flag + n - 1, 4, hipMemcpyDeviceToHost, s));
    ORBC_HIP_CHECK(hipStreamSynchronize(s));
    return (long long)last[0] + last[1];
}

void launch_nlist(const NlistArgs &a, const uint32_t *occ, long long nocc, bool far, const FarScratch &fs,
                  hipStream_t s) {
    if (a.n <= 0 || nocc <= 0) return;
    const long long nb = (nocc + BLOCK / 64 - 1) / (BLOCK / 64);
    hipLaunchKernelGGL(k_nlist, dim3((unsigned)nb), dim3(BLOCK), 0, s, a, occ, nocc);
    if (far) {
        hipLaunchKernelGGL(k_far_flags, dim3(nblk(a.n)), dim3(BLOCK), 0, s, a.n, a.type, fs.flag);
        exclusive_scan_u32(fs.flag, fs.pos, a.n, fs.scan_temp, s);
        hipLaunchKernelGGL(k_far_compact, dim3(nblk(a.n)), dim3(BLOCK), 0, s, a.n, fs.flag, fs.pos, fs.list, fs.count);
        // enough waves to fill the GPU; each takes list entries in turn
        const unsigned grid = (unsigned)std::min<long long>(nblk(a.n), 4096);
        hipLaunchKernelGGL(k_nlist_far, dim3(grid), dim3(BLOCK), 0, s, a, fs.list, fs.count);
    }
}

void launch_bondlist(long long n, long long stride, const uint32_t *slot, const uint32_t *inv,
                     const uint32_t *bstart, const uint32_t *bpartner, const uint8_t *btype, uint32_t *bnd,
                     uint8_t *nbnd, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_bondlist, dim3(nblk(n)), dim3(BLOCK), 0, s, n, stride, slot, inv, bstart, bpartner,
                       btype, bnd, nbnd);
}

}  // namespace orbc

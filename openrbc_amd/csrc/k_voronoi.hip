// k_voronoi.hip -- Voronoi cells of lipids on gfx950.
//
// The reference tessellates lipids into ~14-lipid cells by Lloyd iteration
// (src/voronoi.h:53-140), answers nearest-centroid and radius queries with a
// k-d tree (src/kdtree.h:206-285) and uses the cells for its pair stencil and
// for constrain_volume (src/constrain_volume.h:26-83).  Here the centroids are
// binned in a uniform grid (Morton-sorted with the same radix sort) and
// queries scan grid rings outward, which is exact: after ring R every
// unvisited centroid is farther than R cell edges from the query.
//
// Per-cell sums (centroids, normals) use 64-bit fixed-point atomics: integer
// addition is associative, so the sums do not depend on the order in which
// waves arrive (deterministic), and the quantum (2^-32 for positions, 2^-40
// for directors) is far below FP32 resolution.
#include "orbc_internal.h"

namespace orbc {

namespace {

constexpr double kPosScale = 4294967296.0;        // 2^32
constexpr double kNrmScale = 1099511627776.0;     // 2^40

inline unsigned nblk(long long n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

__device__ inline float normsq_unfused(float a, float b, float c) {
    return __fadd_rn(__fadd_rn(__fmul_rn(a, a), __fmul_rn(b, b)), __fmul_rn(c, c));
}

__device__ inline int gcoord(float v, const Grid &g, int d) {
    int c = (int)floorf((v - (float)g.lo[d]) / (float)g.cs);
    return c < 0 ? 0 : (c >= g.dim[d] ? g.dim[d] - 1 : c);
}

__device__ inline long long glin(int x, int y, int z, const Grid &g) {
    return ((long long)z * g.dim[1] + y) * g.dim[0] + x;
}

// reference Morton key, src/reorder_morton.h:25-42
__device__ inline unsigned bit_space3(unsigned x) {
    x = (x | (x << 12)) & 0x00FC003Fu;
    x = (x | (x << 6)) & 0x381C0E07u;
    x = (x | (x << 4)) & 0x190C8643u;
    x = (x | (x << 2)) & 0x49249249u;
    return x;
}

__global__ __launch_bounds__(BLOCK) void k_morton_ref(const float *p, long long n, double bsize, uint32_t *keys) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const unsigned a = (unsigned)((double)(2.0f * p[3 * i]) + bsize);
    const unsigned b = (unsigned)((double)(2.0f * p[3 * i + 1]) + bsize);
    const unsigned c = (unsigned)((double)(2.0f * p[3 * i + 2]) + bsize);
    keys[i] = bit_space3(a) | (bit_space3(b) << 1) | (bit_space3(c) << 2);
}

__global__ __launch_bounds__(BLOCK) void k_centroid_keys(const float *c, long long m, Grid g, uint32_t *keys,
                                                         uint32_t *vals, Gate gt) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= m || gate_closed(gt)) return;
    keys[i] = (uint32_t)glin(gcoord(c[3 * i], g, 0), gcoord(c[3 * i + 1], g, 1), gcoord(c[3 * i + 2], g, 2), g);
    vals[i] = (uint32_t)i;
}

__global__ __launch_bounds__(BLOCK) void k_cgrid_ranges(const uint32_t *k, long long m, uint32_t *start,
                                                        uint32_t *end, Gate gt) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= m || gate_closed(gt)) return;
    if (i == 0 || k[i] != k[i - 1]) start[k[i]] = (uint32_t)i;
    if (i + 1 == m || k[i] != k[i + 1]) end[k[i]] = (uint32_t)(i + 1);
}

// drift of the centroids since they were binned (0 when the index is current)
__device__ inline float centroid_drift(const VoroArgs &a) {
    if (!a.cdisp) return 0.f;
    const float d2 = __uint_as_float(*a.cdisp);
    return d2 > 0.f ? sqrtf(d2) * 1.0001f + 1e-5f : 0.f;
}

// nearest centroid (lowest index on equal squared distance), grid rings outward
__device__ int nearest_one(const VoroArgs &a, float p0, float p1, float p2) {
    const float drift = centroid_drift(a);
    const Grid g = *a.cg_dev;
    const int cx = gcoord(p0, g, 0), cy = gcoord(p1, g, 1), cz = gcoord(p2, g, 2);
    float best = INFINITY;
    int bi = -1;
    const int rmax = max(g.dim[0], max(g.dim[1], g.dim[2]));
    for (int R = 0; R <= rmax; ++R) {
        for (int dz = -R; dz <= R; ++dz) {
            const int z = cz + dz;
            if (z < 0 || z >= g.dim[2]) continue;
            for (int dy = -R; dy <= R; ++dy) {
                const int y = cy + dy;
                if (y < 0 || y >= g.dim[1]) continue;
                for (int dx = -R; dx <= R; ++dx) {
                    if (max(abs(dx), max(abs(dy), abs(dz))) != R) continue;   // shell of ring R only
                    const int x = cx + dx;
                    if (x < 0 || x >= g.dim[0]) continue;
                    const long long cl = glin(x, y, z, g);
                    for (uint32_t q = a.cg_start[cl]; q < a.cg_end[cl]; ++q) {
                        const int k = (int)a.cg_idx[q];
                        const float d = normsq_unfused(p0 - a.centroids[3 * k], p1 - a.centroids[3 * k + 1],
                                                       p2 - a.centroids[3 * k + 2]);
                        if (d < best || (d == best && k < bi)) { best = d; bi = k; }
                    }
                }
            }
        }
        // every unvisited centroid was binned at least R cell edges away from
        // the query and has drifted by at most `drift` since
        const float reach = (float)R * (float)g.cs - drift;
        if (bi >= 0 && reach > 0.f && best <= reach * reach) break;
    }
    return bi;
}

// Same answer as nearest_one, seeded with a guess k0 (the previous
// affiliation): every centroid at least as near as k0 lies in the ball of
// radius |p - c_k0| around p, so only the grid cells overlapping that ball's
// box are scanned (typically 1-4 instead of 27).  The radius carries a
// relative margin so FP32 rounding of the box ends cannot drop a candidate.
__device__ int nearest_warm(const VoroArgs &a, float p0, float p1, float p2, int k0) {
    const Grid g = *a.cg_dev;
    float best = normsq_unfused(p0 - a.centroids[3 * k0], p1 - a.centroids[3 * k0 + 1], p2 - a.centroids[3 * k0 + 2]);
    int bi = k0;
    // a centroid as near as k0 was binned within r + drift of p
    const float r = sqrtf(best) * 1.0001f + 1e-4f + centroid_drift(a);
    const int x0 = gcoord(p0 - r, g, 0), x1 = gcoord(p0 + r, g, 0);
    const int y0 = gcoord(p1 - r, g, 1), y1 = gcoord(p1 + r, g, 1);
    const int z0 = gcoord(p2 - r, g, 2), z1 = gcoord(p2 + r, g, 2);
    for (int z = z0; z <= z1; ++z)
        for (int y = y0; y <= y1; ++y)
            for (int x = x0; x <= x1; ++x) {
                const long long cl = glin(x, y, z, g);
                for (uint32_t q = a.cg_start[cl]; q < a.cg_end[cl]; ++q) {
                    const int k = (int)a.cg_idx[q];
                    const float d = normsq_unfused(p0 - a.centroids[3 * k], p1 - a.centroids[3 * k + 1],
                                                   p2 - a.centroids[3 * k + 2]);
                    if (d < best || (d == best && k < bi)) { best = d; bi = k; }
                }
            }
    return bi;
}

// Warm partition of the particle store, one wavefront per 64 consecutive
// (Morton-ordered, hence nearby) particles.  Each lane seeds its search ball
// from its previous cell (radius |p - c_k0| plus the centroid drift, as in
// nearest_warm); the union of the 64 balls is a box, the centroids binned in
// the grid cells overlapping it are staged in LDS, and every lane scans them
// all by broadcast reads.  Same answer as nearest_warm (lowest index on equal
// distance); a wave whose box holds too many cells or centroids falls back to
// nearest_warm per lane.
#ifndef ORBC_NW_CAP
#define ORBC_NW_CAP 192
#endif
constexpr int NW_CAP = ORBC_NW_CAP;   // staged centroids per wave (larger caps measured slower: fewer waves resident)
__global__ __launch_bounds__(BLOCK) void k_nearest_wave(VoroArgs a, const double *qd, long long nq, int32_t *out) {
    if (gate_closed(a.gate)) return;
    __shared__ float4 s_c[BLOCK / 64][NW_CAP];
    __shared__ uint32_t s_pref[BLOCK / 64][65];
    __shared__ uint32_t s_st[BLOCK / 64][64];
    const int w = threadIdx.x >> 6, lane = __lane_id();
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    const bool act = i < nq;
    const Grid g = *a.cg_dev;
    float p0 = 0.f, p1 = 0.f, p2 = 0.f, best = INFINITY, r = 0.f;
    int bi = -1, k0 = -1;
    if (act) {
        p0 = (float)qd[3 * i]; p1 = (float)qd[3 * i + 1]; p2 = (float)qd[3 * i + 2];
        k0 = out[i];
        if (k0 >= 0 && k0 < a.ncell) {
            best = normsq_unfused(p0 - a.centroids[3 * k0], p1 - a.centroids[3 * k0 + 1], p2 - a.centroids[3 * k0 + 2]);
            bi = k0;
            r = sqrtf(best) * 1.0001f + 1e-4f + centroid_drift(a);
        }
    }
    // wave box of the search balls (lanes without a guess use the full search below)
    const bool guessed = act && bi >= 0;
    const bool all_guessed = __ballot(guessed) == __ballot(act);
    float lo0 = guessed ? p0 - r : INFINITY, lo1 = guessed ? p1 - r : INFINITY, lo2 = guessed ? p2 - r : INFINITY;
    float hi0 = guessed ? p0 + r : -INFINITY, hi1 = guessed ? p1 + r : -INFINITY, hi2 = guessed ? p2 + r : -INFINITY;
    for (int off = 32; off > 0; off >>= 1) {
        lo0 = fminf(lo0, __shfl_xor(lo0, off, 64)); lo1 = fminf(lo1, __shfl_xor(lo1, off, 64));
        lo2 = fminf(lo2, __shfl_xor(lo2, off, 64)); hi0 = fmaxf(hi0, __shfl_xor(hi0, off, 64));
        hi1 = fmaxf(hi1, __shfl_xor(hi1, off, 64)); hi2 = fmaxf(hi2, __shfl_xor(hi2, off, 64));
    }
    bool wave_ok = all_guessed && __ballot(act) != 0ull;
    int x0 = 0, x1 = -1, y0 = 0, y1 = -1, z0 = 0, z1 = -1, nx = 0, ny = 0, ncl = 0;
    if (wave_ok) {
        x0 = gcoord(lo0, g, 0); x1 = gcoord(hi0, g, 0);
        y0 = gcoord(lo1, g, 1); y1 = gcoord(hi1, g, 1);
        z0 = gcoord(lo2, g, 2); z1 = gcoord(hi2, g, 2);
        nx = x1 - x0 + 1; ny = y1 - y0 + 1;
        ncl = nx * ny * (z1 - z0 + 1);
        wave_ok = ncl <= 64;
    }
    uint32_t total = 0;
    if (wave_ok) {
        uint32_t st = 0, cnt = 0;
        if (lane < ncl) {
            const int cx = x0 + lane % nx, cy = y0 + (lane / nx) % ny, cz = z0 + lane / (nx * ny);
            const long long cl = glin(cx, cy, cz, g);
            st = a.cg_start[cl];
            cnt = a.cg_end[cl] - st;
        }
        uint32_t inc = cnt;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(inc, off, 64);
            if (lane >= off) inc += y;
        }
        total = __shfl(inc, 63, 64);
        s_pref[w][lane + 1] = inc;
        s_st[w][lane] = st;
        if (lane == 0) s_pref[w][0] = 0;
        wave_ok = total <= (uint32_t)NW_CAP;
    }
    if (wave_ok) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (uint32_t t = lane; t < total; t += 64) {
            int lo = 0, hi = 64;                    // largest k with pref[k] <= t
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_pref[w][mid] <= t) lo = mid; else hi = mid;
            }
            const int k = (int)a.cg_idx[s_st[w][lo] + (t - s_pref[w][lo])];
            s_c[w][t] = make_float4(a.centroids[3 * k], a.centroids[3 * k + 1], a.centroids[3 * k + 2],
                                    __int_as_float(k));
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (uint32_t t = 0; t < total; ++t) {
            const float4 c = s_c[w][t];
            const int k = __float_as_int(c.w);
            const float d = normsq_unfused(p0 - c.x, p1 - c.y, p2 - c.z);
            if (d < best || (d == best && k < bi)) { best = d; bi = k; }
        }
        if (act) out[i] = bi;
    } else if (act) {
        out[i] = (bi >= 0) ? nearest_warm(a, p0, p1, p2, k0) : nearest_one(a, p0, p1, p2);
    }
}

// warm: out[] holds the previous affiliation and is used as the guess
__global__ __launch_bounds__(BLOCK) void k_nearest(VoroArgs a, const float *q, long long nq, const double *qd,
                                                   int32_t *out, int warm) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nq || gate_closed(a.gate)) return;
    float p0, p1, p2;
    if (qd) {
        p0 = (float)qd[3 * i]; p1 = (float)qd[3 * i + 1]; p2 = (float)qd[3 * i + 2];
    } else {
        p0 = q[3 * i]; p1 = q[3 * i + 1]; p2 = q[3 * i + 2];
    }
    const int k0 = warm ? out[i] : -1;
    out[i] = (k0 >= 0 && k0 < a.ncell) ? nearest_warm(a, p0, p1, p2, k0) : nearest_one(a, p0, p1, p2);
}

__global__ __launch_bounds__(BLOCK) void k_within(VoroArgs a, const float *q, long long nq, float rmax, int cap,
                                                  int *counts, int *out) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nq) return;
    const Grid g = *a.cg_dev;
    const float p0 = q[3 * i], p1 = q[3 * i + 1], p2 = q[3 * i + 2];
    const float r2 = rmax * rmax;
    const int R = (int)ceilf(rmax / (float)g.cs) + 1;
    const int cx = gcoord(p0, g, 0), cy = gcoord(p1, g, 1), cz = gcoord(p2, g, 2);
    int cnt = 0;
    for (int z = max(0, cz - R); z <= min(g.dim[2] - 1, cz + R); ++z)
        for (int y = max(0, cy - R); y <= min(g.dim[1] - 1, cy + R); ++y)
            for (int x = max(0, cx - R); x <= min(g.dim[0] - 1, cx + R); ++x) {
                const long long cl = glin(x, y, z, g);
                for (uint32_t s = a.cg_start[cl]; s < a.cg_end[cl]; ++s) {
                    const int k = (int)a.cg_idx[s];
                    // src/kdtree.h:274-275: normsq(pts - q) < max_r * max_r
                    const float d = normsq_unfused(a.centroids[3 * k] - p0, a.centroids[3 * k + 1] - p1,
                                                   a.centroids[3 * k + 2] - p2);
                    if (d < r2) {
                        if (cnt < cap) out[(long long)i * cap + cnt] = k;
                        ++cnt;
                    }
                }
            }
    counts[i] = cnt;
}

// Per-cell fixed-point sums of a lipid field (positions or directors) plus the
// lipid count: sums[4k..4k+2] += round(field * scale), sums[4k+3] += 1.
//
// Device-scope 64-bit atomics are slow, and the cells (~14 lipids) do not
// follow the particle order (Morton order of the binning grid), so the sums are
// aggregated first: within a wave over runs of equal cell in consecutive lanes,
// then per block (SUM_CHUNK particles) in an LDS hash table keyed by cell, and
// only the table is flushed with global atomics.  A key that finds the table
// full goes straight to the global atomics.  Integer sums are exact, so the
// grouping does not change the result.
#ifndef ORBC_SUM_LOG_SLOTS
#define ORBC_SUM_LOG_SLOTS 9
#endif
constexpr int SUM_LOG_SLOTS = ORBC_SUM_LOG_SLOTS;
constexpr int SUM_SLOTS = 1 << SUM_LOG_SLOTS;   // LDS hash slots
constexpr int SUM_CHUNK = 4 * SUM_SLOTS;        // particles per block

// inclusive segmented scan over runs of equal keys; true on the last lane of a run
__device__ inline bool wave_run_reduce4(int key, unsigned long long v[4]) {
    const int lane = __lane_id();
    const int prev = __shfl_up(key, 1, 64), next = __shfl_down(key, 1, 64);
    const bool head = lane == 0 || prev != key;
    const bool tail = lane == 63 || next != key;
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const int start = 63 - __clzll((long long)(heads & upto));      // first lane of my run
    for (int off = 1; off < 64; off <<= 1) {
        unsigned long long t[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) t[c] = __shfl_up(v[c], off, 64);
        if (lane - off >= start)
#pragma unroll
            for (int c = 0; c < 4; ++c) v[c] += t[c];
    }
    return tail;
}

template <bool DIRECTORS>
__global__ __launch_bounds__(BLOCK) void k_cell_sums(VoroArgs a, unsigned long long *sums) {
    if (gate_closed(a.gate)) return;
    __shared__ int tkey[SUM_SLOTS];
    __shared__ unsigned long long tval[SUM_SLOTS][4];
    for (int t = threadIdx.x; t < SUM_SLOTS; t += BLOCK) {
        tkey[t] = -1;
        tval[t][0] = tval[t][1] = tval[t][2] = tval[t][3] = 0ull;
    }
    __syncthreads();
    const double *src = DIRECTORS ? a.nrm : a.x;
    const double scale = DIRECTORS ? kNrmScale : kPosScale;
    const long long base = (long long)blockIdx.x * SUM_CHUNK;
    for (int it = 0; it < SUM_CHUNK / BLOCK; ++it) {
        const long long i = base + (long long)it * BLOCK + threadIdx.x;
        int k = -1;
        unsigned long long v[4] = {0ull, 0ull, 0ull, 0ull};
        if (i < a.n && a.type[i] == 0) {
            k = a.aff[i];
            if (k >= 0) {
#pragma unroll
                for (int d = 0; d < 3; ++d) v[d] = (unsigned long long)(long long)llrint(src[3 * i + d] * scale);
                v[3] = 1ull;
            }
        }
        const bool tail = wave_run_reduce4(k, v);
        if (tail && k >= 0) {
            unsigned h = ((unsigned)k * 2654435761u) >> (32 - SUM_LOG_SLOTS);
            bool placed = false;
            for (int probe = 0; probe < SUM_SLOTS; ++probe) {
                const int prev = atomicCAS(&tkey[h], -1, k);
                if (prev == -1 || prev == k) {
#pragma unroll
                    for (int c = 0; c < 4; ++c) atomicAdd(&tval[h][c], v[c]);
                    placed = true;
                    break;
                }
                h = (h + 1) & (SUM_SLOTS - 1);
            }
            if (!placed)
#pragma unroll
                for (int c = 0; c < 4; ++c) atomicAdd(&sums[4 * k + c], v[c]);
        }
    }
    __syncthreads();
    // a cell whose every member is in this block's table has no other
    // contributor: its sums are stored (the sums are exact integers, so the
    // result is the same as the atomic path's)
    const bool sized = a.csize && a.csize[a.ncell] != 0u;
    for (int t = threadIdx.x; t < SUM_SLOTS; t += BLOCK) {
        const int k = tkey[t];
        if (k < 0) continue;
        if (sized && (unsigned long long)a.csize[k] == tval[t][3]) {
#pragma unroll
            for (int c = 0; c < 4; ++c) sums[4 * k + c] = tval[t][c];
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) atomicAdd(&sums[4 * k + c], tval[t][c]);
        }
    }
}

// update_centroid, src/voronoi.h:124-140 (empty cells keep their centroid).
// The sums are cleared after reading: they are zero whenever no kernel pair
// (sums, finish) is in flight, so no fill is needed before the next sums.
__global__ __launch_bounds__(BLOCK) void k_centroid_finish(VoroArgs a, unsigned long long *sums) {
    if (gate_closed(a.gate)) return;
    const long long k = (long long)blockIdx.x * BLOCK + threadIdx.x;
    float *c = a.centroids;
    float d2 = 0.f;
    if (k < a.ncell) {
        unsigned long long q[4];
        for (int d = 0; d < 4; ++d) { q[d] = sums[4 * k + d]; sums[4 * k + d] = 0ull; }
        const long long cnt = (long long)q[3];
        if (cnt > 0) {
            for (int d = 0; d < 3; ++d) {
                const double s = (double)(long long)q[d] / kPosScale;
                c[3 * k + d] = (float)(s / (double)cnt);
            }
        }
    }
    if (a.cdisp && k < a.ncell) {
        const float e0 = c[3 * k] - a.cref[3 * k], e1 = c[3 * k + 1] - a.cref[3 * k + 1], e2 = c[3 * k + 2] - a.cref[3 * k + 2];
        d2 = e0 * e0 + e1 * e1 + e2 * e2;
    }
    for (int off = 32; off > 0; off >>= 1) d2 = fmaxf(d2, __shfl_xor(d2, off, 64));
    if (a.cdisp && __lane_id() == 0 && d2 > 0.f &&
        __float_as_uint(d2) > __hip_atomic_load(a.cdisp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMax(a.cdisp, __float_as_uint(d2));
}

__device__ inline void block_sum_store(double v, double *partial) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __shared__ double ws[BLOCK / 64];
    if (__lane_id() == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int w = 0; w < BLOCK / 64; ++w) s += ws[w];
        partial[blockIdx.x] = s;
    }
}

// centre of the centroids from the block partials of k_centroid_center, once
// per centroid update (one wavefront; fixed summation order)
__global__ void k_center_finish(const double *center_partial, long long ncp, long long m, double *center, Gate gt) {
    if (gate_closed(gt)) return;
    for (int d = 0; d < 3; ++d) {
        double v = 0;
        for (long long b = threadIdx.x; b < ncp; b += 64) v += center_partial[d * ncp + b];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (threadIdx.x == 0) center[d] = v / (double)m;
    }
}

// per-cell normal, orientation and cone volume, src/constrain_volume.h:54-62,
// about the centre from k_center_finish; the cell sums are cleared after
// reading (see k_centroid_finish).  Volume partials go to vol_partial.
__global__ __launch_bounds__(BLOCK) void k_volume_finish(unsigned long long *sums, const float *c, long long m,
                                                         const double *center, double *normals,
                                                         double *vol_partial, uint32_t *csize, Gate gt) {
    if (gate_closed(gt)) return;
    const double s_center[3] = {center[0], center[1], center[2]};
    const long long k = (long long)blockIdx.x * BLOCK + threadIdx.x;
    double vol = 0.0;
    if (k < m) {
        unsigned long long q[4];
        for (int d = 0; d < 4; ++d) { q[d] = sums[4 * k + d]; sums[4 * k + d] = 0ull; }
        const long long cnt = (long long)q[3];
        if (csize) csize[k] = (uint32_t)cnt;
        double nx = (double)(long long)q[0] / kNrmScale;
        double ny = (double)(long long)q[1] / kNrmScale;
        double nz = (double)(long long)q[2] / kNrmScale;
        const double nn = sqrt(nx * nx + ny * ny + nz * nz);
        if (cnt > 0 && nn > 0) {
            nx /= nn; ny /= nn; nz /= nn;
            const double dx = c[3 * k] - s_center[0], dy = c[3 * k + 1] - s_center[1], dz = c[3 * k + 2] - s_center[2];
            if (nx * dx + ny * dy + nz * dz < 0) { nx = -nx; ny = -ny; nz = -nz; }
            const double h = dx * nx + dy * ny + dz * nz;
            vol = h * (double)cnt * 3.1415926 * 1.26 / 4.0 / 3.0 * 1e-6;
        } else {
            nx = ny = nz = 0.0;
        }
        normals[3 * k] = nx; normals[3 * k + 1] = ny; normals[3 * k + 2] = nz;
    }
    if (csize && blockIdx.x == 0 && threadIdx.x == 0) csize[m] = 1u;   // counts current (read by later kernels)
    block_sum_store(vol, vol_partial);
}

__global__ __launch_bounds__(BLOCK) void k_centroid_center(const float *c, long long m, double *partial, Gate gt) {
    if (gate_closed(gt)) return;
    const long long k = (long long)blockIdx.x * BLOCK + threadIdx.x;
    for (int d = 0; d < 3; ++d) {
        block_sum_store(k < m ? (double)c[3 * k + d] : 0.0, partial + d * gridDim.x);
        __syncthreads();
    }
}

// f += strength (V0 - V)/V0 * normal_cell * m_type (src/constrain_volume.h:67-75,
// with each particle's own type)
__global__ __launch_bounds__(BLOCK) void k_volume_push(VoroArgs a, const double *normals, const double *fpush,
                                                       double *f_ext, double *f64, float *f32) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= a.n || gate_closed(a.gate)) return;
    const int k = a.aff[i];
    double add[3] = {0, 0, 0};
    if (k >= 0) {
        const double s = (*fpush) * a.mass[a.type[i]];
        for (int d = 0; d < 3; ++d) add[d] = s * normals[3 * k + d];
    }
    for (int d = 0; d < 3; ++d) {
        if (f_ext) f_ext[3 * i + d] = add[d];
        if (f64) f64[3 * i + d] += add[d];
        if (f32) f32[3 * i + d] += (float)add[d];
    }
}

__device__ inline double wave_sum_d(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// one wavefront: V = sum of the block partials, push = strength (target - V) / target
__global__ void k_volume_force(const double *vol_partial, long long nb, double target, double strength, double *scal,
                               Gate gt) {
    if (gate_closed(gt)) return;
    double v = 0;
    for (long long b = threadIdx.x; b < nb; b += 64) v += vol_partial[b];
    v = wave_sum_d(v);
    if (threadIdx.x == 0) {
        scal[0] = v;
        scal[1] = strength * (target - v) / target;
    }
}

// ---- stray-lipid cleanup, src/cleanup.h:29-91 --------------------------------
// key: 1 + the lipid's cell (0 for proteins and lipids without one), so a
// stable sort by key groups the members of every cell; d2 = |x - c|^2 in FP64
// against the FP32 centroid.
__global__ __launch_bounds__(BLOCK) void k_stray_keys(VoroArgs a, uint32_t *keys, uint32_t *vals, double *d2) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= a.n) return;
    int k = a.type[i] == 0 ? a.aff[i] : -1;
    double d = 0.0;
    if (k >= 0 && k < a.ncell) {
        const double dx = a.x[3 * i] - (double)a.centroids[3 * k];
        const double dy = a.x[3 * i + 1] - (double)a.centroids[3 * k + 1];
        const double dz = a.x[3 * i + 2] - (double)a.centroids[3 * k + 2];
        d = __fma_rn(dz, dz, __fma_rn(dy, dy, dx * dx));
    } else {
        k = -1;
    }
    keys[i] = (uint32_t)(k + 1);
    vals[i] = (uint32_t)i;
    d2[i] = d;
}

// One thread per cell: the threshold is the member distance of rank size/2
// in ascending order (dr2[size / 2] after the reference's sort) times tol^2,
// and a member is deleted unless d2 < threshold.  del[] is written for every
// member; the count of deletions is added to *count.
__global__ __launch_bounds__(BLOCK) void k_stray_mark(long long ncell, const uint32_t *start, const uint32_t *end,
                                                      const uint32_t *svals, const double *d2, double tol2,
                                                      uint32_t *del, unsigned *count) {
    const long long k = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (k >= ncell) return;
    const uint32_t b = start[k + 1], e = end[k + 1];
    if (e <= b) return;
    const uint32_t half = (e - b) / 2;
    double med = 0.0;
    for (uint32_t p = b; p < e; ++p) {
        const double v = d2[svals[p]];
        uint32_t lt = 0, le = 0;
        for (uint32_t q = b; q < e; ++q) {
            const double w = d2[svals[q]];
            lt += w < v;
            le += w <= v;
        }
        if (lt <= half && half < le) { med = v; break; }
    }
    const double thr = med * tol2;
    unsigned nd = 0;
    for (uint32_t p = b; p < e; ++p) {
        const uint32_t i = svals[p];
        const bool kill = !(d2[i] < thr);
        del[i] = kill ? 1 : 0;
        nd += kill;
    }
    if (nd) atomicAdd(count, nd);
}

}  // namespace

void launch_morton_ref(const float *pts, long long n, double bsize, uint32_t *keys, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_morton_ref, dim3(nblk(n)), dim3(BLOCK), 0, s, pts, n, bsize, keys);
}

void launch_centroid_keys(const float *c, long long m, Grid g, uint32_t *keys, uint32_t *vals, hipStream_t s,
                          Gate gt) {
    if (m > 0) hipLaunchKernelGGL(k_centroid_keys, dim3(nblk(m)), dim3(BLOCK), 0, s, c, m, g, keys, vals, gt);
}

void launch_cgrid_ranges(const uint32_t *sorted_keys, long long m, uint32_t *start, uint32_t *end, hipStream_t s,
                         Gate gt) {
    if (m > 0) hipLaunchKernelGGL(k_cgrid_ranges, dim3(nblk(m)), dim3(BLOCK), 0, s, sorted_keys, m, start, end, gt);
}

void launch_nearest(const VoroArgs &a, const float *q, long long nq, const double *qd, int32_t *out, bool warm,
                    hipStream_t s) {
    if (nq <= 0) return;
    if (warm && qd)
        hipLaunchKernelGGL(k_nearest_wave, dim3(nblk(nq)), dim3(BLOCK), 0, s, a, qd, nq, out);
    else
        hipLaunchKernelGGL(k_nearest, dim3(nblk(nq)), dim3(BLOCK), 0, s, a, q, nq, qd, out, (int)warm);
}

void launch_within(const VoroArgs &a, const float *q, long long nq, float rmax, int cap, int *counts, int *out,
                   hipStream_t s) {
    if (nq > 0) hipLaunchKernelGGL(k_within, dim3(nblk(nq)), dim3(BLOCK), 0, s, a, q, nq, rmax, cap, counts, out);
}

// sums must be zero on entry (they are cleared by the finish kernels)
void launch_centroid_sums(const VoroArgs &a, unsigned long long *sums, hipStream_t s) {
    if (a.n > 0)
        hipLaunchKernelGGL(k_cell_sums<false>, dim3((unsigned)((a.n + SUM_CHUNK - 1) / SUM_CHUNK)), dim3(BLOCK), 0, s,
                           a, sums);
}

void launch_centroid_finish(const VoroArgs &a, unsigned long long *sums, hipStream_t s) {
    if (a.ncell > 0) hipLaunchKernelGGL(k_centroid_finish, dim3(nblk(a.ncell)), dim3(BLOCK), 0, s, a, sums);
}

// centre-of-centroids partials (3 x nblk(ncell)) at partial; rerun whenever the centroids change
void launch_centroid_center(const VoroArgs &a, double *partial, hipStream_t s) {
    if (a.ncell > 0) {
        const long long nbc = nblk(a.ncell);
        hipLaunchKernelGGL(k_centroid_center, dim3((unsigned)nbc), dim3(BLOCK), 0, s, a.centroids, a.ncell,
                           partial, a.gate);
        hipLaunchKernelGGL(k_center_finish, dim3(1), dim3(64), 0, s, partial, nbc, a.ncell, partial + 4 * nbc, a.gate);
    }
}

void launch_volume(const VoroArgs &a, unsigned long long *sums, double *normals, double *partial, double *scal,
                   double target, double strength, hipStream_t s) {
    // scal: [0] V, [1] push; partial: centre partials (launch_centroid_center), then volume
    // partials, then the centre (3 doubles at 4 * nbc)
    const long long nbc = nblk(a.ncell);
    if (a.n > 0)
        hipLaunchKernelGGL(k_cell_sums<true>, dim3((unsigned)((a.n + SUM_CHUNK - 1) / SUM_CHUNK)), dim3(BLOCK), 0, s,
                           a, sums);
    hipLaunchKernelGGL(k_volume_finish, dim3((unsigned)nbc), dim3(BLOCK), 0, s, sums, a.centroids, a.ncell,
                       partial + 4 * nbc, normals, partial + 3 * nbc, a.csize, a.gate);
    hipLaunchKernelGGL(k_volume_force, dim3(1), dim3(64), 0, s, partial + 3 * nbc, nbc, target, strength, scal,
                       a.gate);
}

void launch_volume_push(const VoroArgs &a, const double *normals, const double *fpush, double *f_ext, double *f64,
                        float *f32, hipStream_t s) {
    if (a.n > 0) hipLaunchKernelGGL(k_volume_push, dim3(nblk(a.n)), dim3(BLOCK), 0, s, a, normals, fpush, f_ext, f64, f32);
}


void launch_stray_keys(const VoroArgs &a, uint32_t *keys, uint32_t *vals, double *d2, hipStream_t s) {
    if (a.n > 0) hipLaunchKernelGGL(k_stray_keys, dim3(nblk(a.n)), dim3(BLOCK), 0, s, a, keys, vals, d2);
}

void launch_stray_mark(long long ncell, const uint32_t *start, const uint32_t *end, const uint32_t *svals,
                       const double *d2, double tol2, uint32_t *del, unsigned *count, hipStream_t s) {
    if (ncell > 0)
        hipLaunchKernelGGL(k_stray_mark, dim3(nblk(ncell)), dim3(BLOCK), 0, s, ncell, start, end, svals, d2, tol2,
                           del, count);
}

}  // namespace orbc

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
                                                         uint32_t *vals) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= m) return;
    keys[i] = (uint32_t)glin(gcoord(c[3 * i], g, 0), gcoord(c[3 * i + 1], g, 1), gcoord(c[3 * i + 2], g, 2), g);
    vals[i] = (uint32_t)i;
}

__global__ __launch_bounds__(BLOCK) void k_cgrid_ranges(const uint32_t *k, long long m, uint32_t *start,
                                                        uint32_t *end) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= m) return;
    if (i == 0 || k[i] != k[i - 1]) start[k[i]] = (uint32_t)i;
    if (i + 1 == m || k[i] != k[i + 1]) end[k[i]] = (uint32_t)(i + 1);
}

__device__ inline bool voro_gated(const VoroArgs &a) { return a.gstop && (*a.gstop != 0 || *a.gslot != 0u); }

// nearest centroid (lowest index on equal squared distance), grid rings outward
__device__ int nearest_one(const VoroArgs &a, float p0, float p1, float p2) {
    const Grid &g = a.cg;
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
        // every unvisited centroid is at least R cell edges away from the query
        const float reach = (float)R * (float)g.cs;
        if (bi >= 0 && best <= reach * reach) break;
    }
    return bi;
}

// Same answer as nearest_one, seeded with a guess k0 (the previous
// affiliation): every centroid at least as near as k0 lies in the ball of
// radius |p - c_k0| around p, so only the grid cells overlapping that ball's
// box are scanned (typically 1-4 instead of 27).  The radius carries a
// relative margin so FP32 rounding of the box ends cannot drop a candidate.
__device__ int nearest_warm(const VoroArgs &a, float p0, float p1, float p2, int k0) {
    const Grid &g = a.cg;
    float best = normsq_unfused(p0 - a.centroids[3 * k0], p1 - a.centroids[3 * k0 + 1], p2 - a.centroids[3 * k0 + 2]);
    int bi = k0;
    const float r = sqrtf(best) * 1.0001f + 1e-4f;
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

// warm: out[] holds the previous affiliation and is used as the guess
__global__ __launch_bounds__(BLOCK) void k_nearest(VoroArgs a, const float *q, long long nq, const double *qd,
                                                   int32_t *out, int warm) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nq || voro_gated(a)) return;
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
    const Grid &g = a.cg;
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

// Segmented sum over runs of equal keys in consecutive lanes (particles are
// Morton-ordered and cells compact, so a wave holds a few runs): the last lane
// of each run adds the run's totals with one atomic per component.  Lanes with
// key < 0 contribute nothing.  Integer sums are exact, so the grouping does not
// change the result.
__device__ inline void wave_run_add4(int key, unsigned long long v0, unsigned long long v1, unsigned long long v2,
                                     unsigned long long v3, unsigned long long *sums) {
    const int lane = __lane_id();
    const int prev = __shfl_up(key, 1, 64), next = __shfl_down(key, 1, 64);
    const bool head = lane == 0 || prev != key;
    const bool tail = lane == 63 || next != key;
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const int start = 63 - __clzll((long long)(heads & upto));      // first lane of my run
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned long long a0 = __shfl_up(v0, off, 64), a1 = __shfl_up(v1, off, 64);
        const unsigned long long a2 = __shfl_up(v2, off, 64), a3 = __shfl_up(v3, off, 64);
        if (lane - off >= start) { v0 += a0; v1 += a1; v2 += a2; v3 += a3; }
    }
    if (tail && key >= 0) {
        atomicAdd(&sums[4 * key], v0);
        atomicAdd(&sums[4 * key + 1], v1);
        atomicAdd(&sums[4 * key + 2], v2);
        atomicAdd(&sums[4 * key + 3], v3);
    }
}

// fixed-point per-cell sums of lipid positions: sums[4k..4k+2] = sum x, [4k+3] = count
__global__ __launch_bounds__(BLOCK) void k_centroid_sums(VoroArgs a, unsigned long long *sums) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    int k = -1;
    unsigned long long v[3] = {0, 0, 0};
    if (i < a.n && a.type[i] == 0) {
        k = a.aff[i];
        if (k >= 0)
            for (int d = 0; d < 3; ++d) v[d] = (unsigned long long)(long long)llrint(a.x[3 * i + d] * kPosScale);
    }
    wave_run_add4(k, v[0], v[1], v[2], k >= 0 ? 1ull : 0ull, sums);
}

// update_centroid, src/voronoi.h:124-140 (empty cells keep their centroid)
__global__ __launch_bounds__(BLOCK) void k_centroid_finish(VoroArgs a, const unsigned long long *sums) {
    const long long k = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (k >= a.ncell || voro_gated(a)) return;
    float *c = a.centroids;
    const long long cnt = (long long)sums[4 * k + 3];
    if (cnt == 0) return;
    for (int d = 0; d < 3; ++d) {
        const double s = (double)(long long)sums[4 * k + d] / kPosScale;
        c[3 * k + d] = (float)(s / (double)cnt);
    }
}

// sum of lipid directors and lipid count per cell, src/constrain_volume.h:55
__global__ __launch_bounds__(BLOCK) void k_normal_sums(VoroArgs a, unsigned long long *sums) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    int k = -1;
    unsigned long long v[3] = {0, 0, 0};
    if (i < a.n && a.type[i] == 0) {
        k = a.aff[i];
        if (k >= 0)
            for (int d = 0; d < 3; ++d) v[d] = (unsigned long long)(long long)llrint(a.nrm[3 * i + d] * kNrmScale);
    }
    wave_run_add4(k, v[0], v[1], v[2], k >= 0 ? 1ull : 0ull, sums);
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

// per-cell normal, orientation and cone volume, src/constrain_volume.h:54-62
__global__ __launch_bounds__(BLOCK) void k_volume_finish(const unsigned long long *sums, const float *c, long long m,
                                                         const double *center, double *normals, double *partial) {
    const long long k = (long long)blockIdx.x * BLOCK + threadIdx.x;
    double vol = 0.0;
    if (k < m) {
        const long long cnt = (long long)sums[4 * k + 3];
        double nx = (double)(long long)sums[4 * k] / kNrmScale;
        double ny = (double)(long long)sums[4 * k + 1] / kNrmScale;
        double nz = (double)(long long)sums[4 * k + 2] / kNrmScale;
        const double nn = sqrt(nx * nx + ny * ny + nz * nz);
        if (cnt > 0 && nn > 0) {
            nx /= nn; ny /= nn; nz /= nn;
            const double dx = c[3 * k] - center[0], dy = c[3 * k + 1] - center[1], dz = c[3 * k + 2] - center[2];
            if (nx * dx + ny * dy + nz * dz < 0) { nx = -nx; ny = -ny; nz = -nz; }
            const double h = dx * nx + dy * ny + dz * nz;
            vol = h * (double)cnt * 3.1415926 * 1.26 / 4.0 / 3.0 * 1e-6;
        } else {
            nx = ny = nz = 0.0;
        }
        normals[3 * k] = nx; normals[3 * k + 1] = ny; normals[3 * k + 2] = nz;
    }
    block_sum_store(vol, partial);
}

__global__ __launch_bounds__(BLOCK) void k_centroid_center(const float *c, long long m, double *partial) {
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
    if (i >= a.n || voro_gated(a)) return;
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
__global__ void k_volume_force(const double *vol_partial, long long nb, double target, double strength, double *scal) {
    double v = 0;
    for (long long b = threadIdx.x; b < nb; b += 64) v += vol_partial[b];
    v = wave_sum_d(v);
    if (threadIdx.x == 0) {
        scal[0] = v;
        scal[1] = strength * (target - v) / target;
    }
}

// one wavefront: centre of the centroids
__global__ void k_center_finish(const double *partial, long long nb, long long m, double *center) {
    for (int d = 0; d < 3; ++d) {
        double v = 0;
        for (long long b = threadIdx.x; b < nb; b += 64) v += partial[d * nb + b];
        v = wave_sum_d(v);
        if (threadIdx.x == 0) center[d] = v / (double)m;
    }
}

}  // namespace

void launch_morton_ref(const float *pts, long long n, double bsize, uint32_t *keys, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_morton_ref, dim3(nblk(n)), dim3(BLOCK), 0, s, pts, n, bsize, keys);
}

void launch_centroid_keys(const float *c, long long m, Grid g, uint32_t *keys, uint32_t *vals, hipStream_t s) {
    if (m > 0) hipLaunchKernelGGL(k_centroid_keys, dim3(nblk(m)), dim3(BLOCK), 0, s, c, m, g, keys, vals);
}

void launch_cgrid_ranges(const uint32_t *sorted_keys, long long m, uint32_t *start, uint32_t *end, hipStream_t s) {
    if (m > 0) hipLaunchKernelGGL(k_cgrid_ranges, dim3(nblk(m)), dim3(BLOCK), 0, s, sorted_keys, m, start, end);
}

void launch_nearest(const VoroArgs &a, const float *q, long long nq, const double *qd, int32_t *out, bool warm,
                    hipStream_t s) {
    if (nq > 0) hipLaunchKernelGGL(k_nearest, dim3(nblk(nq)), dim3(BLOCK), 0, s, a, q, nq, qd, out, (int)warm);
}

void launch_within(const VoroArgs &a, const float *q, long long nq, float rmax, int cap, int *counts, int *out,
                   hipStream_t s) {
    if (nq > 0) hipLaunchKernelGGL(k_within, dim3(nblk(nq)), dim3(BLOCK), 0, s, a, q, nq, rmax, cap, counts, out);
}

void launch_centroid_sums(const VoroArgs &a, unsigned long long *sums, hipStream_t s) {
    launch_fill_u32(reinterpret_cast<uint32_t *>(sums), 0u, 8 * a.ncell, s);
    if (a.n > 0) hipLaunchKernelGGL(k_centroid_sums, dim3(nblk(a.n)), dim3(BLOCK), 0, s, a, sums);
}

void launch_centroid_finish(const VoroArgs &a, const unsigned long long *sums, hipStream_t s) {
    if (a.ncell > 0) hipLaunchKernelGGL(k_centroid_finish, dim3(nblk(a.ncell)), dim3(BLOCK), 0, s, a, sums);
}

void launch_volume(const VoroArgs &a, unsigned long long *sums, double *normals, double *partial, double *scal,
                   double target, double strength, hipStream_t s) {
    // scal: [0] V, [1] push, [2..4] center; partial >= 3 * nblk(ncell) doubles
    const long long nbc = nblk(a.ncell);
    hipLaunchKernelGGL(k_centroid_center, dim3((unsigned)nbc), dim3(BLOCK), 0, s, a.centroids, a.ncell, partial);
    hipLaunchKernelGGL(k_center_finish, dim3(1), dim3(64), 0, s, partial, nbc, a.ncell, scal + 2);
    launch_fill_u32(reinterpret_cast<uint32_t *>(sums), 0u, 8 * a.ncell, s);
    if (a.n > 0) hipLaunchKernelGGL(k_normal_sums, dim3(nblk(a.n)), dim3(BLOCK), 0, s, a, sums);
    hipLaunchKernelGGL(k_volume_finish, dim3((unsigned)nbc), dim3(BLOCK), 0, s, sums, a.centroids, a.ncell, scal + 2,
                       normals, partial);
    hipLaunchKernelGGL(k_volume_force, dim3(1), dim3(64), 0, s, partial, nbc, target, strength, scal);
}

void launch_volume_push(const VoroArgs &a, const double *normals, const double *fpush, double *f_ext, double *f64,
                        float *f32, hipStream_t s) {
    if (a.n > 0) hipLaunchKernelGGL(k_volume_push, dim3(nblk(a.n)), dim3(BLOCK), 0, s, a, normals, fpush, f_ext, f64, f32);
}

}  // namespace orbc

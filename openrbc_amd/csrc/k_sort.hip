// k_sort.hip -- exclusive scan and stable LSD radix sort for gfx950.
//
// These replace the reference's per-thread std::sort + parallel merge of
// Morton keys (src/reorder_morton.h:44-122) and its partial_sum counting
// passes (src/voronoi.h:224-233, src/reorder.h:50-61).
//
// Radix sort: 8-bit digits, tiles of 2048 keys per 256-thread block.  Per
// pass: (1) LDS histogram per tile, stored digit-major so that (2) one
// exclusive scan gives every (digit, tile) its output base, and (3) a stable
// scatter ranks equal digits inside each 64-lane wavefront with ballots (the
// wave's lanes agreeing on all digit bits are its "peers"), then across the
// block's four waves through LDS.  Deterministic: no atomics decide order.
#include <mutex>
#include <stdexcept>

#include "orbc_internal.h"

namespace orbc {

namespace {

constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = BLOCK * SCAN_ITEMS;   // 4096
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = BLOCK * RS_ITEMS;       // 2048

__device__ inline uint32_t wave_inclusive_scan(uint32_t x) {
    const int lane = __lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    return x;
}

// Exclusive scan of one value per thread over a 256-thread block.
__device__ inline uint32_t block_exclusive_scan(uint32_t x, uint32_t *total) {
    __shared__ uint32_t wsum[BLOCK / 64];
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    const uint32_t inc = wave_inclusive_scan(x);
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t wbase = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) {
        const uint32_t s = wsum[w];
        if (w < wave) wbase += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return wbase + inc - x;
}

__global__ __launch_bounds__(BLOCK) void k_scan_tiles(const uint32_t *in, uint32_t *out, long long n,
                                                      uint32_t *tilesums, Gate gt) {
    if (gate_closed(gt)) return;
    const long long base = (long long)blockIdx.x * SCAN_TILE + (long long)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const long long p = base + k;
        v[k] = p < n ? in[p] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(s, &total);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const long long p = base + k;
        if (p < n) out[p] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) tilesums[blockIdx.x] = total;
}

// single block: exclusive scan of the tile sums in place (any length)
__global__ __launch_bounds__(BLOCK) void k_scan_top(uint32_t *sums, long long m, Gate gt) {
    if (gate_closed(gt)) return;
    __shared__ uint32_t carry_s;
    if (threadIdx.x == 0) carry_s = 0;
    __syncthreads();
    for (long long c0 = 0; c0 < m; c0 += SCAN_TILE) {
        const long long base = c0 + (long long)threadIdx.x * SCAN_ITEMS;
        uint32_t v[SCAN_ITEMS];
        uint32_t s = 0;
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k) {
            const long long p = base + k;
            v[k] = p < m ? sums[p] : 0u;
            s += v[k];
        }
        uint32_t total;
        uint32_t run = block_exclusive_scan(s, &total) + carry_s;
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k) {
            const long long p = base + k;
            if (p < m) sums[p] = run;
            run += v[k];
        }
        __syncthreads();
        if (threadIdx.x == 0) carry_s += total;
        __syncthreads();
    }
}

__global__ __launch_bounds__(BLOCK) void k_scan_add(uint32_t *out, long long n, const uint32_t *tilesums, Gate gt) {
    if (gate_closed(gt)) return;
    const uint32_t add = tilesums[blockIdx.x];
    const long long base = (long long)blockIdx.x * SCAN_TILE;
    for (int k = threadIdx.x; k < SCAN_TILE; k += BLOCK) {
        const long long p = base + k;
        if (p < n) out[p] += add;
    }
}

__global__ __launch_bounds__(BLOCK) void k_rs_hist(const uint32_t *kin, long long n, int shift, uint32_t mask,
                                                   uint32_t *hist, long long ntiles, Gate gt) {
    if (gate_closed(gt)) return;
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const long long t0 = (long long)blockIdx.x * RS_TILE;
#pragma unroll
    for (int it = 0; it < RS_ITEMS; ++it) {
        const long long p = t0 + it * BLOCK + threadIdx.x;
        if (p < n) atomicAdd(&h[(kin[p] >> shift) & mask], 1u);
    }
    __syncthreads();
    hist[(long long)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(BLOCK) void k_rs_scatter(const uint32_t *kin, const uint32_t *vin, uint32_t *kout,
                                                      uint32_t *vout, long long n, int shift, int nbits,
                                                      const uint32_t *offs, long long ntiles, Gate gt) {
    if (gate_closed(gt)) return;
    __shared__ uint32_t base[256];
    __shared__ uint32_t wcnt[BLOCK / 64][256];
    const int tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
    const uint32_t mask = (nbits >= 32) ? 0xffffffffu : ((1u << nbits) - 1u);
    base[tid] = offs[(long long)tid * ntiles + blockIdx.x];
    const long long t0 = (long long)blockIdx.x * RS_TILE;
    const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int it = 0; it < RS_ITEMS; ++it) {
#pragma unroll
        for (int w = 0; w < BLOCK / 64; ++w) wcnt[w][tid] = 0;
        __syncthreads();
        const long long p = t0 + it * BLOCK + tid;
        const bool act = p < n;
        const uint32_t key = act ? kin[p] : 0u;
        const uint32_t val = act ? vin[p] : 0u;
        const uint32_t d = (key >> shift) & mask;
        unsigned long long peers = __ballot(act);
        for (int b = 0; b < nbits; ++b) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = __popcll(peers & lt_mask);
        if (act && rank == 0) wcnt[wave][d] = __popcll(peers);
        __syncthreads();
        uint32_t run = base[tid];
#pragma unroll
        for (int w = 0; w < BLOCK / 64; ++w) {
            const uint32_t c = wcnt[w][tid];
            wcnt[w][tid] = run;
            run += c;
        }
        base[tid] = run;
        __syncthreads();
        if (act) {
            const uint32_t dst = wcnt[wave][d] + rank;
            kout[dst] = key;
            vout[dst] = val;
        }
        __syncthreads();
    }
}

long long cdiv(long long a, long long b) { return (a + b - 1) / b; }

}  // namespace

size_t scan_temp_bytes(long long n) { return sizeof(uint32_t) * (size_t)(cdiv(n, SCAN_TILE) + 1); }

void exclusive_scan_u32(const uint32_t *in, uint32_t *out, long long n, void *temp, hipStream_t s, Gate gt) {
    if (n <= 0) return;
    const long long nt = cdiv(n, SCAN_TILE);
    uint32_t *tilesums = static_cast<uint32_t *>(temp);
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)nt), dim3(BLOCK), 0, s, in, out, n, tilesums, gt);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(BLOCK), 0, s, tilesums, nt, gt);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nt), dim3(BLOCK), 0, s, out, n, tilesums, gt);
}

size_t radix_temp_bytes(long long n) {
    const long long nt = cdiv(n > 0 ? n : 1, RS_TILE);
    const long long hn = 256 * nt;
    return sizeof(uint32_t) * (size_t)(2 * hn) + scan_temp_bytes(hn) + 256;
}

void radix_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t *keys_alt, uint32_t *vals_alt, long long n,
                      int bits, void *temp, hipStream_t s, bool *result_in_alt, Gate gt) {
    *result_in_alt = false;
    if (n <= 1 || bits <= 0) return;
    const long long nt = cdiv(n, RS_TILE);
    const long long hn = 256 * nt;
    uint32_t *hist = static_cast<uint32_t *>(temp);
    uint32_t *offs = hist + hn;
    void *stemp = offs + hn;
    uint32_t *ki = keys, *vi = vals, *ko = keys_alt, *vo = vals_alt;
    bool alt = false;
    for (int shift = 0; shift < bits; shift += 8) {
        const int nb = (bits - shift) < 8 ? (bits - shift) : 8;
        const uint32_t mask = (1u << nb) - 1u;
        hipLaunchKernelGGL(k_rs_hist, dim3((unsigned)nt), dim3(BLOCK), 0, s, ki, n, shift, mask, hist, nt, gt);
        exclusive_scan_u32(hist, offs, hn, stemp, s, gt);
        hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned)nt), dim3(BLOCK), 0, s, ki, vi, ko, vo, n, shift, nb,
                           offs, nt, gt);
        uint32_t *tk = ki, *tv = vi;
        ki = ko; vi = vo; ko = tk; vo = tv;
        alt = !alt;
    }
    *result_in_alt = alt;
}

// ---- fixed scratch of the allocation-free sort (orbc_sort_pairs_noalloc) ----
// The sort's device memory beyond the caller's keys and values -- the
// ping-pong key and value arrays and the histogram / scan temporaries for up
// to NOALLOC_SORT_MAX pairs -- lives in this file-scope array instead of the
// context's allocations.  Every GPU the library runs on holds its own copy
// until the process ends.  In host code the array's name is not a device
// address: scratch_address() asks the runtime for this device's copy once,
// with the device current, and callers pass that pointer to the kernels.
namespace {
constexpr size_t NOALLOC_PAIR_BYTES = 2 * sizeof(uint32_t) * (size_t)NOALLOC_SORT_MAX;   // k2, v2
constexpr size_t NOALLOC_TEMP_BYTES = (size_t)1 << 20 | (size_t)4096;                      // >= radix_temp_bytes(max)
constexpr size_t NOALLOC_BYTES = NOALLOC_PAIR_BYTES + NOALLOC_TEMP_BYTES;
__device__ __align__(256) uint4 scratch[NOALLOC_BYTES / sizeof(uint4)];
std::mutex g_scratch_mu;
void *g_scratch_ptr[64] = {};
}  // namespace

void *scratch_address(int device) {
    if (radix_temp_bytes(NOALLOC_SORT_MAX) > NOALLOC_TEMP_BYTES)
        throw std::runtime_error("fixed sort scratch smaller than the radix temporaries");
    if (device < 0 || device >= 64) throw std::runtime_error("device ordinal out of range for the fixed scratch");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) throw HipError(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

void sort_pairs_in_scratch(uint32_t *keys, uint32_t *vals, long long n, int bits, void *scr, hipStream_t s) {
    if (n > NOALLOC_SORT_MAX) throw std::runtime_error("more pairs than the fixed scratch holds");
    uint32_t *k2 = static_cast<uint32_t *>(scr);
    uint32_t *v2 = k2 + NOALLOC_SORT_MAX;
    void *temp = v2 + NOALLOC_SORT_MAX;
    bool alt = false;
    radix_sort_pairs(keys, vals, k2, v2, n, bits, temp, s, &alt);
    if (alt) {
        ORBC_HIP_CHECK(hipMemcpyAsync(keys, k2, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice, s));
        ORBC_HIP_CHECK(hipMemcpyAsync(vals, v2, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice, s));
    }
}

}  // namespace orbc

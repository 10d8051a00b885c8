// k_integrate.hip -- elementwise integrator kernels and FP64 reductions.
//
// Each kernel restates one reference integrate() functor
// (src/integrate_nh.h:29-55 dispatches them over both containers; here both
// species live in one store).  State is FP64; f,t come from the FP32 (mixed)
// or FP64 accumulators.  Global sums (kinetic energy, temperature, energy,
// virial) are two-pass block reductions in a fixed order, so they are
// deterministic (the reference uses omp atomics, src/integrate_nh.h:85-88).
#include "orbc_internal.h"

namespace orbc {

namespace {

__constant__ double c_mass[NT];
__constant__ double c_inv_mass[NT];

inline unsigned nblk(long long n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

__device__ inline void philox(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
}

template <typename F>
__device__ inline void load_ft(const IntegArgs &a, long long i, double *f, double *t);
template <>
__device__ inline void load_ft<float>(const IntegArgs &a, long long i, double *f, double *t) {
    for (int d = 0; d < 3; ++d) { f[d] = a.f32[3 * i + d]; t[d] = a.t32[3 * i + d]; }
}
template <>
__device__ inline void load_ft<double>(const IntegArgs &a, long long i, double *f, double *t) {
    for (int d = 0; d < 3; ++d) { f[d] = a.f64[3 * i + d]; t[d] = a.t64[3 * i + d]; }
}
template <typename F>
__device__ inline void store_ft(const IntegArgs &a, long long i, const double *f, const double *t);
template <>
__device__ inline void store_ft<float>(const IntegArgs &a, long long i, const double *f, const double *t) {
    for (int d = 0; d < 3; ++d) { a.f32[3 * i + d] = (float)f[d]; a.t32[3 * i + d] = (float)t[d]; }
}
template <>
__device__ inline void store_ft<double>(const IntegArgs &a, long long i, const double *f, const double *t) {
    for (int d = 0; d < 3; ++d) { a.f64[3 * i + d] = f[d]; a.t64[3 * i + d] = t[d]; }
}

__device__ inline void cross(const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ inline void rotate_director(double *n, const double *o, double dt) {
    double on[3];
    cross(o, n, on);
    double g[3] = {n[0] + on[0] * dt, n[1] + on[1] * dt, n[2] + on[2] * dt};
    const double s = 1.0 / sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    n[0] = g[0] * s; n[1] = g[1] * s; n[2] = g[2] * s;
}

// refresh the FP32 mirrors of particle i from its FP64 state (build-time anchor)
__device__ inline void write_mirror(const IntegArgs &a, long long i, const double *x, const double *n, int ty) {
    if (!a.pq) return;
    const uint32_t w = a.anc[i];
    const Anchor &g = a.anc_grid;
    const long long A[3] = {(long long)(w & g.mx), (long long)((w >> g.shy) & g.my), (long long)(w >> g.shz)};
    PQ q;
    q.x = make_float4(__double2float_rn(x[0] - anchor_pos(g, 0, A[0])), __double2float_rn(x[1] - anchor_pos(g, 1, A[1])),
                      __double2float_rn(x[2] - anchor_pos(g, 2, A[2])), __uint_as_float(w));
    q.n = make_float4(__double2float_rn(n[0]), __double2float_rn(n[1]), __double2float_rn(n[2]), __int_as_float(ty));
    a.pq[i] = q;
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

template <typename F>
__global__ __launch_bounds__(BLOCK) void k_clear(IntegArgs a) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= a.n) return;
    const double z[3] = {0, 0, 0};
    store_ft<F>(a, i, z, z);
}

// verlet_langevin, src/integrate_langevin.h:275-297
template <typename F>
__global__ __launch_bounds__(BLOCK) void k_langevin(IntegArgs a) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= a.n) return;
    const int ty = a.type[i];
    double f[3], t[3], x[3], v[3], n[3], o[3], q[3];
    load_ft<F>(a, i, f, t);
    for (int d = 0; d < 3; ++d) { x[d] = a.x[3 * i + d]; v[d] = a.v[3 * i + d]; n[d] = a.nrm[3 * i + d]; o[d] = a.o[3 * i + d]; }
    cross(n, t, q);
    for (int d = 0; d < 3; ++d) o[d] += a.dt * q[d];
    rotate_director(n, o, a.dt);
    uint32_t c0 = a.id[i], c1 = (uint32_t)a.step, c2 = (uint32_t)(a.step >> 32), c3 = 0u;
    philox(c0, c1, c2, c3, (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
    const double u = 4.656612873077392578125e-10;
    const double r[3] = {(double)c0 * u - 1.0, (double)c1 * u - 1.0, (double)c2 * u - 1.0};
    const double dtm = a.dt * c_inv_mass[ty];
    for (int d = 0; d < 3; ++d) {
        f[d] -= a.gamma[ty] * v[d] + (a.sig_group ? a.sig_group[a.grp[i] * NT + ty] : a.sigma[ty]) * r[d];
        v[d] += f[d] * dtm;
        x[d] += v[d] * a.dt;
    }
    for (int d = 0; d < 3; ++d) { a.x[3 * i + d] = x[d]; a.v[3 * i + d] = v[d]; a.nrm[3 * i + d] = n[d]; a.o[3 * i + d] = o[d]; }
    const double z[3] = {0, 0, 0};
    store_ft<F>(a, i, z, z);
    write_mirror(a, i, x, n, ty);
}

// verlet_initial_bounce_clearforce_update, src/integrate_nh.h:187-230
template <typename F>
__global__ __launch_bounds__(BLOCK) void k_nh_first(IntegArgs a) {
    if (a.stop && *a.stop) return;          // a batched run stopped before this step (whole grid)
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (a.skin_clear && i == 0) *a.skin_clear = 0u;
    double ke = 0.0;
    float d2 = 0.f;
    bool past = false;
    if (i < a.n) {
        const int ty = a.type[i];
        const double m = c_mass[ty];
        const double g = 1.0 / (1.0 + 0.5 * a.dt * (*a.zeta));
        double f[3], t[3], x[3], v[3], n[3], o[3];
        load_ft<F>(a, i, f, t);
        for (int d = 0; d < 3; ++d) { x[d] = a.x[3 * i + d]; v[d] = a.v[3 * i + d]; n[d] = a.nrm[3 * i + d]; o[d] = a.o[3 * i + d]; }
        for (int d = 0; d < 3; ++d) {
            v[d] = (v[d] + 0.5 / m * a.dt * f[d]) * g;
            x[d] += v[d] * a.dt;
        }
        for (int d = 0; d < 3; ++d) {
            if (a.anc_grid.L[d] > 0) continue;  // periodic axis: no walls
            if (x[d] < a.box_lo[d]) { x[d] = a.box_lo[d] + (a.box_lo[d] - x[d]); v[d] = -v[d]; }
            else if (x[d] > a.box_hi[d]) { x[d] = a.box_hi[d] - (x[d] - a.box_hi[d]); v[d] = -v[d]; }
        }
        ke = 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int d = 0; d < 3; ++d) o[d] += 0.5 * a.dt * t[d];
        rotate_director(n, o, a.dt);
        for (int d = 0; d < 3; ++d) { a.x[3 * i + d] = x[d]; a.v[3 * i + d] = v[d]; a.nrm[3 * i + d] = n[d]; a.o[3 * i + d] = o[d]; }
        const double z[3] = {0, 0, 0};
        store_ft<F>(a, i, z, z);
        write_mirror(a, i, x, n, ty);
        if (a.disp_max || a.skin_slot) {
            const uint32_t w = a.anc[i];
            const Anchor &g = a.anc_grid;
            const long long A[3] = {(long long)(w & g.mx), (long long)((w >> g.shy) & g.my), (long long)(w >> g.shz)};
            for (int d = 0; d < 3; ++d) {
                const float e = __double2float_rn(x[d] - anchor_pos(g, d, A[d])) - a.rb[3 * i + d];
                d2 += e * e;
            }
            if (!isfinite(d2)) d2 = 3.0e38f;
            past = d2 > a.half_skin2;
            if (past && a.bud2 && d2 < 3.0e38f && a.nnbr[i] == 0) past = d2 > a.bud2[i];
        }
    }
    if (a.skin_slot) {
        const unsigned long long over = __ballot(past);
        if (over && __lane_id() == (int)__ffsll((long long)over) - 1) atomicOr(a.skin_slot, 1u);
    }
    if (a.disp_max) {
        // one atomic per wave, only where the wave raises the maximum (as k_disp)
        for (int off = 32; off > 0; off >>= 1) d2 = fmaxf(d2, __shfl_xor(d2, off, 64));
        if (__lane_id() == 0 && d2 > 0.f &&
            __float_as_uint(d2) > __hip_atomic_load(a.disp_max, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(a.disp_max, __float_as_uint(d2));
    }
    block_sum_store(ke, a.ke_partial);
}

// post_toque_final_update, src/integrate_nh.h:246-268
template <typename F>
__global__ __launch_bounds__(BLOCK) void k_nh_second(IntegArgs a) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    double ke = 0.0;
    if (i < a.n) {
        const int ty = a.type[i];
        const double m = c_mass[ty];
        const double zeta = *a.zeta;
        double f[3], t[3], v[3], n[3], q[3];
        load_ft<F>(a, i, f, t);
        for (int d = 0; d < 3; ++d) { v[d] = a.v[3 * i + d]; n[d] = a.nrm[3 * i + d]; }
        cross(n, t, q);
        for (int d = 0; d < 3; ++d) {
            v[d] += 0.5 * a.dt * (f[d] / m - zeta * v[d]);
            a.o[3 * i + d] += 0.5 * a.dt * q[d];
            a.v[3 * i + d] = v[d];
        }
        store_ft<F>(a, i, f, q);
        ke = 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    block_sum_store(ke, a.ke_partial);
}

// post_torque, src/integrate_nh.h:146-154
template <typename F>
__global__ __launch_bounds__(BLOCK) void k_post_torque(IntegArgs a) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= a.n) return;
    double f[3], t[3], n[3], q[3];
    load_ft<F>(a, i, f, t);
    for (int d = 0; d < 3; ++d) n[d] = a.nrm[3 * i + d];
    cross(n, t, q);
    store_ft<F>(a, i, f, q);
}

// bounce_back, src/integrate_nh.h:124-144
__global__ __launch_bounds__(BLOCK) void k_bounce(IntegArgs a) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= a.n) return;
    for (int d = 0; d < 3; ++d) {
        if (a.anc_grid.L[d] > 0) continue;      // periodic axis: no walls
        double x = a.x[3 * i + d];
        if (x < a.box_lo[d]) { a.x[3 * i + d] = a.box_lo[d] + (a.box_lo[d] - x); a.v[3 * i + d] = -a.v[3 * i + d]; }
        else if (x > a.box_hi[d]) { a.x[3 * i + d] = a.box_hi[d] - (x - a.box_hi[d]); a.v[3 * i + d] = -a.v[3 * i + d]; }
    }
    if (a.pq) {
        double x[3], n[3];
        for (int d = 0; d < 3; ++d) { x[d] = a.x[3 * i + d]; n[d] = a.nrm[3 * i + d]; }
        write_mirror(a, i, x, n, a.type[i]);
    }
}

// capped steepest descent, src/openrbc.cpp:114-131 (t already n x t)
template <typename F>
__global__ __launch_bounds__(BLOCK) void k_minimize(IntegArgs a) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= a.n) return;
    const int ty = a.type[i];
    const double m = c_mass[ty];
    double f[3], t[3], x[3], n[3], dn[3];
    load_ft<F>(a, i, f, t);
    for (int d = 0; d < 3; ++d) { x[d] = a.x[3 * i + d]; n[d] = a.nrm[3 * i + d]; }
    cross(t, n, dn);
    const double ndx = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]) / m;
    const double ndn = sqrt(dn[0] * dn[0] + dn[1] * dn[1] + dn[2] * dn[2]);
    double h = a.dt;
    if (ndx > a.dr_opt || ndn > a.dn_opt) h = fmin(a.dr_opt / ndx, a.dn_opt / ndn);
    for (int d = 0; d < 3; ++d) { x[d] += f[d] * (h / m); n[d] += dn[d] * h; }
    const double s = 1.0 / sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int d = 0; d < 3; ++d) { n[d] *= s; a.x[3 * i + d] = x[d]; a.nrm[3 * i + d] = n[d]; }
    write_mirror(a, i, x, n, ty);
}

__global__ __launch_bounds__(BLOCK) void k_reduce_ke(const double *v, const uint8_t *type, long long n,
                                                     double *partial) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    double ke = 0.0;
    if (i < n) ke = 0.5 * c_mass[type[i]] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
    block_sum_store(ke, partial);
}

__global__ __launch_bounds__(BLOCK) void k_sum_array(const double *a, long long n, double *partial) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    block_sum_store(i < n ? a[i] : 0.0, partial);
}

__global__ __launch_bounds__(BLOCK) void k_sum_partials(const double *partial, long long nb, double *out) {
    double s = 0.0;
    for (long long k = threadIdx.x; k < nb; k += BLOCK) s += partial[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    __shared__ double ws[BLOCK / 64];
    if (__lane_id() == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < BLOCK / 64; ++w) t += ws[w];
        *out = t;
    }
}

// zeta += dt/2 / Q (KE - 3/2 N kT), the functor destructors
// src/integrate_nh.h:182-185,241-244
// One workgroup of ZETA_THREADS sums the kinetic-energy partials in a fixed
// order: four independent running sums per thread (so the loads of a thread
// are in flight together), then the wave and workgroup trees.
constexpr int ZETA_THREADS = 1024;
__global__ __launch_bounds__(ZETA_THREADS) void k_zeta_update(const double *partial, long long nb, double *zeta,
                                                              double dt, double Q, double nkT15, const int *stop) {
    if (stop && *stop) return;
    double s4[4] = {0.0, 0.0, 0.0, 0.0};
    long long k = threadIdx.x;
    for (; k + 3 * ZETA_THREADS < nb; k += 4 * ZETA_THREADS) {
#pragma unroll
        for (int u = 0; u < 4; ++u) s4[u] += partial[k + u * ZETA_THREADS];
    }
#pragma unroll
    for (int u = 0; u < 3; ++u) {
        const long long q = k + (long long)u * ZETA_THREADS;
        if (q < nb) s4[u] += partial[q];
    }
    double s = (s4[0] + s4[1]) + (s4[2] + s4[3]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    __shared__ double ws[ZETA_THREADS / 64];
    if (__lane_id() == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double ke = 0;
        for (int w = 0; w < ZETA_THREADS / 64; ++w) ke += ws[w];
        *zeta += 0.5 * dt / Q * (ke - nkT15);
    }
}

__global__ __launch_bounds__(BLOCK) void k_check_finite(const double *x, long long n3, unsigned *bad) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n3 && !isfinite(x[i])) atomicOr(bad, 1u);
}

__global__ __launch_bounds__(BLOCK) void k_scatter3(long long n, const uint32_t *slot, const double *src, double *dst) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const long long s = slot[i];
    for (int d = 0; d < 3; ++d) dst[3 * s + d] = src[3 * i + d];
}
__global__ __launch_bounds__(BLOCK) void k_scatter3f(long long n, const uint32_t *slot, const float *src, double *dst) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const long long s = slot[i];
    for (int d = 0; d < 3; ++d) dst[3 * s + d] = src[3 * i + d];
}
__global__ __launch_bounds__(BLOCK) void k_scatter1(long long n, const uint32_t *slot, const double *src, double *dst) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) dst[slot[i]] = src[i];
}
__global__ __launch_bounds__(BLOCK) void k_gather3(long long n, const uint32_t *slot, const double *src, double *dst) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const long long s = slot[i];
    for (int d = 0; d < 3; ++d) dst[3 * i + d] = src[3 * s + d];
}
__global__ __launch_bounds__(BLOCK) void k_gather3f(long long n, const uint32_t *slot, const double *src, float *dst) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const long long s = slot[i];
    for (int d = 0; d < 3; ++d) dst[3 * i + d] = (float)src[3 * s + d];
}
// dst[i] = src[slot[i]]: a per-particle u16 array from upload order to device order
__global__ __launch_bounds__(BLOCK) void k_gather_u16(long long n, const uint32_t *slot, const uint16_t *src,
                                                      uint16_t *dst) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) dst[i] = src[slot[i]];
}

__global__ __launch_bounds__(BLOCK) void k_fill_u32(uint32_t *p, uint32_t v, long long n, Gate gt) {
    if (gate_closed(gt)) return;
    for (long long i = (long long)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (long long)gridDim.x * BLOCK) p[i] = v;
}

template <typename F>
void integrate_t(int kernel, const IntegArgs &a, hipStream_t s) {
    const dim3 g(nblk(a.n)), b(BLOCK);
    switch (kernel) {
    case 0: hipLaunchKernelGGL(k_clear<F>, g, b, 0, s, a); break;
    case 1: hipLaunchKernelGGL(k_langevin<F>, g, b, 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_nh_first<F>, g, b, 0, s, a); break;
    case 3: hipLaunchKernelGGL(k_nh_second<F>, g, b, 0, s, a); break;
    case 4: hipLaunchKernelGGL(k_post_torque<F>, g, b, 0, s, a); break;
    case 5: hipLaunchKernelGGL(k_bounce, g, b, 0, s, a); break;
    default: break;
    }
}

}  // namespace

void upload_mass(const double *mass, hipStream_t s) {
    double inv[NT];
    for (int k = 0; k < NT; ++k) inv[k] = 1.0 / mass[k];
    ORBC_HIP_CHECK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_mass), mass, sizeof(double) * NT, 0, hipMemcpyHostToDevice, s));
    ORBC_HIP_CHECK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_inv_mass), inv, sizeof(double) * NT, 0, hipMemcpyHostToDevice, s));
    ORBC_HIP_CHECK(hipStreamSynchronize(s));
}

long long reduce_blocks(long long n) { return nblk(n > 0 ? n : 1); }

void launch_integrate(int kernel, const IntegArgs &a, bool mixed, hipStream_t s) {
    if (a.n <= 0) return;
    if (mixed) integrate_t<float>(kernel, a, s);
    else integrate_t<double>(kernel, a, s);
}

void launch_minimize(const IntegArgs &a, bool mixed, hipStream_t s) {
    if (a.n <= 0) return;
    const dim3 g(nblk(a.n)), b(BLOCK);
    if (mixed) hipLaunchKernelGGL(k_minimize<float>, g, b, 0, s, a);
    else hipLaunchKernelGGL(k_minimize<double>, g, b, 0, s, a);
}

void launch_reduce_ke(const double *v, const uint8_t *type, long long n, double *partial, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_ke, dim3(nblk(n > 0 ? n : 1)), dim3(BLOCK), 0, s, v, type, n, partial);
}

void launch_sum_partials(const double *partial, long long nb, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(BLOCK), 0, s, partial, nb, out);
}

void launch_zeta_update(const double *partial, long long nb, double *zeta, double dt, double Q, double nkT15,
                        hipStream_t s, const int *stop) {
    hipLaunchKernelGGL(k_zeta_update, dim3(1), dim3(ZETA_THREADS), 0, s, partial, nb, zeta, dt, Q, nkT15, stop);
}

void launch_sum_array(const double *a, long long n, double *partial, hipStream_t s) {
    hipLaunchKernelGGL(k_sum_array, dim3(nblk(n > 0 ? n : 1)), dim3(BLOCK), 0, s, a, n, partial);
}

void launch_check_finite(const double *x, long long n3, unsigned *bad, hipStream_t s) {
    if (n3 <= 0) return;
    hipLaunchKernelGGL(k_check_finite, dim3(nblk(n3)), dim3(BLOCK), 0, s, x, n3, bad);
}

void launch_scatter_slots(long long n, const uint32_t *slot, const double *src3, double *dst3, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_scatter3, dim3(nblk(n)), dim3(BLOCK), 0, s, n, slot, src3, dst3);
}
void launch_scatter_slots_f(long long n, const uint32_t *slot, const float *src3, double *dst3, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_scatter3f, dim3(nblk(n)), dim3(BLOCK), 0, s, n, slot, src3, dst3);
}
void launch_scatter_slots1(long long n, const uint32_t *slot, const double *src, double *dst, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_scatter1, dim3(nblk(n)), dim3(BLOCK), 0, s, n, slot, src, dst);
}
void launch_gather_slots(long long n, const uint32_t *slot, const double *src3, double *dst3, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_gather3, dim3(nblk(n)), dim3(BLOCK), 0, s, n, slot, src3, dst3);
}
void launch_gather_slots_f(long long n, const uint32_t *slot, const double *src3, float *dst3, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_gather3f, dim3(nblk(n)), dim3(BLOCK), 0, s, n, slot, src3, dst3);
}
void launch_gather_u16(long long n, const uint32_t *slot, const uint16_t *src, uint16_t *dst, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_gather_u16, dim3(nblk(n)), dim3(BLOCK), 0, s, n, slot, src, dst);
}

namespace {
__global__ void k_set_u64(unsigned long long *p, unsigned long long v) { *p = v; }
__global__ void k_add_u64(unsigned long long *p, unsigned long long v) { *p += v; }
}  // namespace

void launch_set_u64(unsigned long long *p, unsigned long long v, hipStream_t s) {
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, s, p, v);
}
void launch_add_u64(unsigned long long *p, unsigned long long v, hipStream_t s) {
    hipLaunchKernelGGL(k_add_u64, dim3(1), dim3(1), 0, s, p, v);
}

void launch_fill_u32(uint32_t *p, uint32_t v, long long n, hipStream_t s, Gate gt) {
    if (n <= 0) return;
    long long nb = nblk(n);
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)nb), dim3(BLOCK), 0, s, p, v, n, gt);
}

}  // namespace orbc

// voronoi.cpp -- host side of the Voronoi cells, constrain_volume and the
// device-kernel entry points that replace reorder_morton / KDTree queries.
//
// VoronoiDiagram (src/voronoi.h:33-141) keeps n_lipid/14 centroids updated by
// Lloyd steps; VCellList::partition (src/voronoi.h:179-237) assigns every
// particle to its nearest centroid.  Here the affiliation is a per-particle
// array that travels with the particle store (it is permuted at every
// neighbour rebuild) and the centroids are indexed by a uniform grid.
#include <cmath>
#include <cstring>
#include <vector>

#include "ctx.h"

using namespace orbc;

#define guarded orbc_guarded

namespace {

int ceil_log2(long long v) {
    int b = 0;
    while ((1LL << b) < v) ++b;
    return b;
}

Grid make_grid(const double lo[3], const double hi[3], double cs) {
    Grid g{};
    g.cs = cs;
    for (int d = 0; d < 3; ++d) {
        g.csd[d] = cs;
        g.lo[d] = lo[d];
        g.dim[d] = std::max(1, (int)std::ceil((hi[d] - lo[d]) / cs) + 1);
        g.bits[d] = ceil_log2(g.dim[d]);
    }
    g.ncell = (long long)g.dim[0] * g.dim[1] * g.dim[2];
    return g;
}

// index m points (device, float xyz) into grid g: cg_idx sorted by cell
struct PointIndex {
    uint32_t *idx = nullptr, *start = nullptr, *end = nullptr;
};

void build_index(orbc_ctx *c, const float *pts, long long m, const Grid &g, uint32_t *k, uint32_t *v, uint32_t *k2,
                 uint32_t *v2, void *temp, uint32_t *start, uint32_t *end, uint32_t **idx_out, Gate gt = {}) {
    launch_fill_u32(start, 0u, g.ncell, c->stream, gt);
    launch_fill_u32(end, 0u, g.ncell, c->stream, gt);
    launch_centroid_keys(pts, m, g, k, v, c->stream, gt);
    bool alt = false;
    radix_sort_pairs(k, v, k2, v2, m, ceil_log2(g.ncell), temp, c->stream, &alt, gt);
    launch_cgrid_ranges(alt ? k2 : k, m, start, end, c->stream, gt);
    *idx_out = alt ? v2 : v;
}

VoroArgs voro_args(orbc_ctx *c) {
    VoroArgs a{};
    a.n = c->n;
    a.ncell = c->voro.ncell;
    a.x = c->x[c->cur];
    a.nrm = c->nrm[c->cur];
    a.type = c->type;
    a.aff = c->voro.aff;
    a.centroids = c->voro.c;
    for (int t = 0; t < NT; ++t) a.mass[t] = c->ffd.mass[t];
    a.cg = c->voro.cg;
    a.cg_start = c->voro.cg_start;
    a.cg_end = c->voro.cg_end;
    a.cg_idx = c->voro.cg_idx;
    a.gate = c->voro.gate;
    a.cref = c->voro.cref;
    a.cdisp = c->voro.cdisp;
    a.cg_dev = c->cg_dev;
    a.csize = c->voro.csize;
    return a;
}

void index_centroids(orbc_ctx *c) {
    auto &V = c->voro;
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = c->grid.lo[d];
        hi[d] = c->grid.lo[d] + c->grid.csd[d] * c->grid.dim[d];
    }
    V.cg = make_grid(lo, hi, c->cg_edge);
    ORBC_HIP_CHECK(hipMemcpyAsync(c->cg_dev, &V.cg, sizeof(Grid), hipMemcpyHostToDevice, c->stream));
    if (V.cg.ncell > V.cg_cap) {
        c->mem.release(V.cg_start); c->mem.release(V.cg_end);
        V.cg_cap = (long long)(1.25 * (double)V.cg.ncell) + 16;
        V.cg_start = c->mem.alloc<uint32_t>(V.cg_cap);
        V.cg_end = c->mem.alloc<uint32_t>(V.cg_cap);
    }
    uint32_t *idx = nullptr;
    build_index(c, V.c, V.ncell, V.cg, V.ck, V.cv, V.ck2, V.cv2, V.temp, V.cg_start, V.cg_end, &idx, V.gate);
    V.cg_idx = idx;
    // binned positions and zero drift
    ORBC_HIP_CHECK(hipMemcpyAsync(V.cref, V.c, 12 * V.ncell, hipMemcpyDeviceToDevice, c->stream));
    launch_fill_u32(V.cdisp, 0u, 1, c->stream);
}

void partition_all(orbc_ctx *c) {
    VoroArgs a = voro_args(c);
    launch_nearest(a, nullptr, c->n, c->x[c->cur], c->voro.aff, c->voro.aff_valid, c->stream);
    c->voro.aff_valid = true;
    // the cell counts of the old affiliation are no longer current
    if (c->voro.csize) launch_fill_u32(c->voro.csize + c->voro.ncell, 0u, 1, c->stream, a.gate);
}

void update_centroids(orbc_ctx *c) {
    VoroArgs a = voro_args(c);
    launch_centroid_sums(a, c->voro.sums, c->stream);
    launch_centroid_finish(a, c->voro.sums, c->stream);
    launch_centroid_center(a, c->voro.partial, c->stream);
}

void voro_alloc(orbc_ctx *c, long long ncell) {
    for (int d = 0; d < 3; ++d)
        if (c->cfg.periodic[d]) throw std::runtime_error("Voronoi cells are not defined on a periodic box");
    voronoi_free(c);
    c->vol_target_set = false;
    auto &V = c->voro;
    V.ncell = ncell;
    V.aff = c->mem.alloc<int32_t>(c->n);
    V.aff_alt = c->mem.alloc<int32_t>(c->n);
    V.c = c->mem.alloc<float>(3 * ncell);
    V.sums = c->mem.alloc<unsigned long long>(4 * ncell);
    // the cell sums are kept at zero between uses (the finish kernels clear them)
    ORBC_HIP_CHECK(hipMemsetAsync(V.sums, 0, sizeof(unsigned long long) * 4 * (size_t)ncell, c->stream));
    V.normals = c->mem.alloc<double>(3 * ncell);
    V.cref = c->mem.alloc<float>(3 * ncell);
    V.cdisp = c->mem.alloc<unsigned>(1);
    V.csize = c->mem.alloc<uint32_t>(ncell + 1);
    launch_fill_u32(V.csize + ncell, 0u, 1, c->stream);
    // centre partials (3 per block), volume partials (1 per block), the centre (3)
    V.partial = c->mem.alloc<double>(4 * (ncell / BLOCK + 2) + 16);
    V.ck = c->mem.alloc<uint32_t>(ncell); V.cv = c->mem.alloc<uint32_t>(ncell);
    V.ck2 = c->mem.alloc<uint32_t>(ncell); V.cv2 = c->mem.alloc<uint32_t>(ncell);
    V.temp = c->mem.alloc<char>(radix_temp_bytes(ncell));
    V.cg_cap = 0;
    V.updates = 0;
}

}  // namespace

namespace orbc {

void voronoi_free(orbc_ctx *c) {
    auto &V = c->voro;
    void *ps[] = {V.aff, V.aff_alt, V.c, V.sums, V.normals, V.partial, V.ck, V.cv, V.ck2, V.cv2, V.cg_start, V.cg_end,
                  V.temp, V.cref, V.cdisp, V.st_start, V.st_end, V.del, V.csize};
    for (void *p : ps) c->mem.release(p);
    V = orbc_ctx::Voro{};
}

void voronoi_rebuild_index(orbc_ctx *c) {
    if (c->voro.ncell > 0) index_centroids(c);
}

// delete_lipid, src/cleanup.h:29-91: per cell, the lipids whose squared
// distance to the centroid is not below tol^2 times the cell's median one
// (dr2[size / 2] of the sorted distances) are removed.  Uses the current
// centroids and affiliation, as the reference uses its current cell lists.
long long delete_stray_lipids(orbc_ctx *c, double tolerance) {
    auto &V = c->voro;
    if (V.ncell <= 0 || !V.aff_valid || c->nl <= 0) return 0;
    const long long n = c->n;
    hipStream_t s = c->stream;
    if (!V.st_start) {
        V.st_start = c->mem.alloc<uint32_t>(V.ncell + 1);
        V.st_end = c->mem.alloc<uint32_t>(V.ncell + 1);
        V.del = c->mem.alloc<uint32_t>(c->stride);
    }
    VoroArgs a = voro_args(c);
    double *d2 = c->stage;                        // 3n doubles of staging: n distances
    launch_stray_keys(a, c->keys, c->vals, d2, s);
    bool alt = false;
    radix_sort_pairs(c->keys, c->vals, c->keys_alt, c->vals_alt, n, ceil_log2(V.ncell + 1), c->sort_temp, s, &alt);
    const uint32_t *sk = alt ? c->keys_alt : c->keys, *sv = alt ? c->vals_alt : c->vals;
    launch_fill_u32(V.st_start, 0u, V.ncell + 1, s);
    launch_fill_u32(V.st_end, 0u, V.ncell + 1, s);
    launch_cgrid_ranges(sk, n, V.st_start, V.st_end, s);
    launch_fill_u32(V.del, 0u, n, s);
    zero_flag(c, 3);
    launch_stray_mark(V.ncell, V.st_start, V.st_end, sv, d2, tolerance * tolerance, V.del, c->flags + 3, s);
    const long long ndel = read_flag(c, 3);
    if (ndel > 0) compact_store(c, V.del, ndel);
    return ndel;
}

// VoronoiDiagram::update + VCellList::update of both containers
void voronoi_update_dev(orbc_ctx *c) {
    update_centroids(c);
    partition_all(c);
    c->voro.updates++;
}

// cone-sum volume, cell normals and push magnitude (scal + 8: V, push)
void volume_eval(orbc_ctx *c, double target, double strength) {
    VoroArgs a = voro_args(c);
    launch_volume(a, c->voro.sums, c->voro.normals, c->voro.partial, c->scal + 8, target, strength, c->stream);
}

// ... and the push force into f_ext (and/or the f accumulators)
double volume_push(orbc_ctx *c, double target, double strength, double *f_ext, bool to_f, bool read_back) {
    VoroArgs a = voro_args(c);
    double *scal = c->scal + 8;
    volume_eval(c, target, strength);
    launch_volume_push(a, c->voro.normals, scal + 1, f_ext, to_f && !c->mixed() ? c->f64 : nullptr,
                       to_f && c->mixed() ? c->f32 : nullptr, c->stream);
    if (!read_back) return 0.0;
    double V = 0;
    ORBC_HIP_CHECK(hipMemcpyAsync(&V, scal, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    sync(c);
    return V;
}

}  // namespace orbc

extern "C" {

int orbc_voronoi_init(orbc_ctx *c, long long n_cells, int n_iter) {
    return guarded(c, [&] {
        ensure_live(c);
        if (c->nl <= 0) return fail(c, ORBC_ERR_INVALID, "Voronoi cells need lipids");
        const long long ncell = n_cells > 0 ? n_cells : std::max<long long>(1, c->nl / c->cfg.voronoi_cell_size);
        if (ncell > c->nl) return fail(c, ORBC_ERR_INVALID, "more cells than lipids");
        if (n_iter < 0) return fail(c, ORBC_ERR_INVALID, "n_iter < 0");
        voro_alloc(c, ncell);
        // initial guess: every delta-th lipid of the container, src/voronoi.h:57-62
        std::vector<double> lx(3 * c->nl);
        launch_scatter_slots(c->n, c->slot, c->x[c->cur], c->stage, c->stream);
        ORBC_HIP_CHECK(hipMemcpyAsync(lx.data(), c->stage, 24 * c->nl, hipMemcpyDeviceToHost, c->stream));
        sync(c);
        const long long delta = c->nl / ncell;
        std::vector<float> cc(3 * ncell);
        for (long long i = 0, k = delta / 2; i < ncell; ++i, k = (k + delta) % c->nl)
            for (int d = 0; d < 3; ++d) cc[3 * i + d] = (float)lx[3 * k + d];
        ORBC_HIP_CHECK(hipMemcpyAsync(c->voro.c, cc.data(), 12 * ncell, hipMemcpyHostToDevice, c->stream));
        // Lloyd iterations, src/voronoi.h:64-71
        for (int it = 0; it < n_iter; ++it) {
            index_centroids(c);
            partition_all(c);
            update_centroids(c);
        }
        index_centroids(c);
        partition_all(c);
        launch_centroid_center(voro_args(c), c->voro.partial, c->stream);
        sync(c);
        if (c->cfg.neighbor_mode == ORBC_NEIGHBOR_VORONOI) rebuild(c);
        return (int)ORBC_OK;
    });
}

int orbc_voronoi_update(orbc_ctx *c) {
    return guarded(c, [&] {
        ensure_live(c);
        if (c->voro.ncell <= 0) return fail(c, ORBC_ERR_STATE, "orbc_voronoi_init first");
        voronoi_update_dev(c);
        if (c->cfg.neighbor_mode == ORBC_NEIGHBOR_VORONOI) rebuild(c);
        sync(c);
        return (int)ORBC_OK;
    });
}

int orbc_voronoi_set_centroids(orbc_ctx *c, long long n_cells, const float *cen) {
    return guarded(c, [&] {
        ensure_live(c);
        if (n_cells <= 0 || !cen) return fail(c, ORBC_ERR_INVALID, "bad centroids");
        voro_alloc(c, n_cells);
        ORBC_HIP_CHECK(hipMemcpyAsync(c->voro.c, cen, 12 * n_cells, hipMemcpyHostToDevice, c->stream));
        index_centroids(c);
        partition_all(c);
        launch_centroid_center(voro_args(c), c->voro.partial, c->stream);
        sync(c);
        if (c->cfg.neighbor_mode == ORBC_NEIGHBOR_VORONOI) rebuild(c);
        return (int)ORBC_OK;
    });
}

int orbc_voronoi_get(orbc_ctx *c, float *cen, int *aff_lipid, int *aff_protein) {
    return guarded(c, [&] {
        if (c->voro.ncell <= 0) return fail(c, ORBC_ERR_STATE, "no Voronoi cells");
        if (cen) ORBC_HIP_CHECK(hipMemcpyAsync(cen, c->voro.c, 12 * c->voro.ncell, hipMemcpyDeviceToHost, c->stream));
        if (aff_lipid || aff_protein) {
            // affiliation to upload order (as doubles through the staging buffer)
            std::vector<double> tmp(c->n);
            double *d = c->stage;
            std::vector<int32_t> hv(c->n);
            ORBC_HIP_CHECK(hipMemcpyAsync(hv.data(), c->voro.aff, 4 * c->n, hipMemcpyDeviceToHost, c->stream));
            std::vector<uint32_t> hs(c->n);
            ORBC_HIP_CHECK(hipMemcpyAsync(hs.data(), c->slot, 4 * c->n, hipMemcpyDeviceToHost, c->stream));
            sync(c);
            (void)d;
            for (long long i = 0; i < c->n; ++i) {
                const long long s = hs[i];
                if (s < c->pslot0) { if (aff_lipid) aff_lipid[s] = hv[i]; }
                else if (aff_protein) aff_protein[s - c->pslot0] = hv[i];
            }
        }
        sync(c);
        return (int)ORBC_OK;
    });
}

int orbc_delete_stray_lipids(orbc_ctx *c, double tolerance, long long *n_deleted) {
    return guarded(c, [&] {
        ensure_live(c);
        if (c->voro.ncell <= 0) return fail(c, ORBC_ERR_STATE, "orbc_voronoi_init first");
        if (!(tolerance > 0)) return fail(c, ORBC_ERR_INVALID, "tolerance must be positive");
        const long long nd = delete_stray_lipids(c, tolerance);
        if (n_deleted) *n_deleted = nd;
        return (int)ORBC_OK;
    });
}

int orbc_voronoi_count(const orbc_ctx *c, long long *n_cells) {
    if (!c || !n_cells) return ORBC_ERR_INVALID;
    *n_cells = c->voro.ncell;
    return ORBC_OK;
}

int orbc_constrain_volume(orbc_ctx *c, double target, double strength, double *volume) {
    return guarded(c, [&] {
        ensure_live(c);
        if (c->voro.ncell <= 0) return fail(c, ORBC_ERR_STATE, "orbc_voronoi_init first");
        if (!(target > 0)) return fail(c, ORBC_ERR_INVALID, "target volume must be positive");
        const double V = volume_push(c, target, strength, nullptr, true, true);
        if (volume) *volume = V;
        return (int)ORBC_OK;
    });
}

int orbc_morton_keys(orbc_ctx *c, long long n, const float *pts, double bsize, unsigned *keys) {
    return guarded(c, [&] {
        if (n < 0 || (n > 0 && (!pts || !keys))) return fail(c, ORBC_ERR_INVALID, "bad arguments");
        if (n == 0) return (int)ORBC_OK;
        float *dp = c->mem.alloc<float>(3 * n);
        uint32_t *dk = c->mem.alloc<uint32_t>(n);
        ORBC_HIP_CHECK(hipMemcpyAsync(dp, pts, 12 * n, hipMemcpyHostToDevice, c->stream));
        launch_morton_ref(dp, n, bsize, dk, c->stream);
        ORBC_HIP_CHECK(hipMemcpyAsync(keys, dk, 4 * n, hipMemcpyDeviceToHost, c->stream));
        sync(c);
        c->mem.release(dp); c->mem.release(dk);
        return (int)ORBC_OK;
    });
}

int orbc_sort_pairs(orbc_ctx *c, long long n, unsigned *keys, unsigned *values, int bits) {
    return guarded(c, [&] {
        if (n < 0 || bits < 1 || bits > 32 || (n > 0 && (!keys || !values))) return fail(c, ORBC_ERR_INVALID, "bad arguments");
        if (n == 0) return (int)ORBC_OK;
        uint32_t *k = c->mem.alloc<uint32_t>(n), *v = c->mem.alloc<uint32_t>(n);
        uint32_t *k2 = c->mem.alloc<uint32_t>(n), *v2 = c->mem.alloc<uint32_t>(n);
        void *t = c->mem.alloc<char>(radix_temp_bytes(n));
        ORBC_HIP_CHECK(hipMemcpyAsync(k, keys, 4 * n, hipMemcpyHostToDevice, c->stream));
        ORBC_HIP_CHECK(hipMemcpyAsync(v, values, 4 * n, hipMemcpyHostToDevice, c->stream));
        bool alt = false;
        radix_sort_pairs(k, v, k2, v2, n, bits, t, c->stream, &alt);
        ORBC_HIP_CHECK(hipMemcpyAsync(keys, alt ? k2 : k, 4 * n, hipMemcpyDeviceToHost, c->stream));
        ORBC_HIP_CHECK(hipMemcpyAsync(values, alt ? v2 : v, 4 * n, hipMemcpyDeviceToHost, c->stream));
        sync(c);
        void *ps[] = {k, v, k2, v2, t};
        for (void *p : ps) c->mem.release(p);
        return (int)ORBC_OK;
    });
}

int orbc_sort_pairs_noalloc(orbc_ctx *c, long long n, unsigned *keys, unsigned *values, int bits) {
    return guarded(c, [&] {
        if (n < 0 || bits < 1 || bits > 32 || (n > 0 && (!keys || !values))) return fail(c, ORBC_ERR_INVALID, "bad arguments");
        if (n > NOALLOC_SORT_MAX)
            return fail(c, ORBC_ERR_INVALID, "n exceeds ORBC_NOALLOC_SORT_MAX (1048576 pairs), the fixed scratch");
        if (!c->scratch) return fail(c, ORBC_ERR_HIP, "the fixed sort scratch has no device address on this device");
        if (n <= 1) return (int)ORBC_OK;
        sort_pairs_in_scratch(keys, values, n, bits, c->scratch, c->stream);
        return (int)ORBC_OK;
    });
}

namespace {

int point_queries(orbc_ctx *c, long long npts, const float *pts, long long nq, const float *q, bool nearest,
                  float rmax, int cap, int *idx, int *counts, int *out) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (long long i = 0; i < npts; ++i)
        for (int d = 0; d < 3; ++d) { lo[d] = std::min(lo[d], (double)pts[3 * i + d]); hi[d] = std::max(hi[d], (double)pts[3 * i + d]); }
    double ext = 0;
    for (int d = 0; d < 3; ++d) ext = std::max(ext, hi[d] - lo[d]);
    double cs = std::max(1e-3, ext / std::max(1.0, std::cbrt((double)npts)));
    if (!nearest) cs = std::max(cs, (double)rmax / 2);
    Grid g = make_grid(lo, hi, cs);
    while (g.ncell > 8 * npts + 4096) { cs *= 1.5; g = make_grid(lo, hi, cs); }
    float *dp = c->mem.alloc<float>(3 * npts), *dq = c->mem.alloc<float>(3 * nq);
    uint32_t *k = c->mem.alloc<uint32_t>(npts), *v = c->mem.alloc<uint32_t>(npts);
    uint32_t *k2 = c->mem.alloc<uint32_t>(npts), *v2 = c->mem.alloc<uint32_t>(npts);
    uint32_t *st = c->mem.alloc<uint32_t>(g.ncell), *en = c->mem.alloc<uint32_t>(g.ncell);
    void *t = c->mem.alloc<char>(radix_temp_bytes(npts));
    ORBC_HIP_CHECK(hipMemcpyAsync(dp, pts, 12 * npts, hipMemcpyHostToDevice, c->stream));
    ORBC_HIP_CHECK(hipMemcpyAsync(dq, q, 12 * nq, hipMemcpyHostToDevice, c->stream));
    uint32_t *idx_d = nullptr;
    build_index(c, dp, npts, g, k, v, k2, v2, t, st, en, &idx_d);
    VoroArgs a{};
    a.ncell = npts; a.centroids = dp; a.cg = g; a.cg_start = st; a.cg_end = en; a.cg_idx = idx_d;
    std::vector<void *> extra;
    Grid *gd = c->mem.alloc<Grid>(1);
    extra.push_back(gd);
    ORBC_HIP_CHECK(hipMemcpyAsync(gd, &a.cg, sizeof(Grid), hipMemcpyHostToDevice, c->stream));
    a.cg_dev = gd;
    if (nearest) {
        int32_t *o = c->mem.alloc<int32_t>(nq);
        extra.push_back(o);
        launch_nearest(a, dq, nq, nullptr, o, false, c->stream);
        ORBC_HIP_CHECK(hipMemcpyAsync(idx, o, 4 * nq, hipMemcpyDeviceToHost, c->stream));
    } else {
        int *dc = c->mem.alloc<int>(nq);
        int *dout = c->mem.alloc<int>((size_t)nq * (cap > 0 ? cap : 1));
        extra.push_back(dc); extra.push_back(dout);
        launch_within(a, dq, nq, rmax, cap, dc, dout, c->stream);
        ORBC_HIP_CHECK(hipMemcpyAsync(counts, dc, 4 * nq, hipMemcpyDeviceToHost, c->stream));
        if (cap > 0) ORBC_HIP_CHECK(hipMemcpyAsync(out, dout, 4 * (size_t)nq * cap, hipMemcpyDeviceToHost, c->stream));
    }
    sync(c);
    void *ps[] = {dp, dq, k, v, k2, v2, st, en, t};
    for (void *p : ps) c->mem.release(p);
    for (void *p : extra) c->mem.release(p);
    return ORBC_OK;
}

}  // namespace

int orbc_nearest(orbc_ctx *c, long long npts, const float *pts, long long nq, const float *q, int *idx) {
    return guarded(c, [&] {
        if (npts <= 0 || nq < 0 || !pts || (nq > 0 && (!q || !idx))) return fail(c, ORBC_ERR_INVALID, "bad arguments");
        if (nq == 0) return (int)ORBC_OK;
        return point_queries(c, npts, pts, nq, q, true, 0.f, 0, idx, nullptr, nullptr);
    });
}

int orbc_within(orbc_ctx *c, long long npts, const float *pts, long long nq, const float *q, float rmax, int cap,
                int *counts, int *out) {
    return guarded(c, [&] {
        if (npts <= 0 || nq < 0 || !pts || cap < 0 || !(rmax > 0) || (nq > 0 && (!q || !counts || (cap > 0 && !out))))
            return fail(c, ORBC_ERR_INVALID, "bad arguments");
        if (nq == 0) return (int)ORBC_OK;
        return point_queries(c, npts, pts, nq, q, false, rmax, cap, nullptr, counts, out);
    });
}

}  // extern "C"

// api.cpp -- orbc_ctx and the C ABI of liborbc_hip.so (include/orbc.h).
//
// The context owns one particle store on the device: both reference
// containers (LipidContainer, ProteContainer; src/container.h:117-157) merged
// into one array kept in Morton order of the binning grid, FP64 state x,v,n,o
// (x,n double-buffered for the fused step), FP32 mirrors for the mixed force
// path, the Verlet list, and per-particle bond lists.  The host keeps each
// species' upload order so results come back in the caller's order.
#include <algorithm>
#include <numeric>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "ctx.h"

using namespace orbc;

static const double kRadius[NT] = {0.56125, 1.12375, 1.12375, 0.56125, 1.5, 0.5};  // src/forcefield_canonical.h:41

namespace orbc {

int fail(orbc_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

}  // namespace orbc

#define guarded orbc_guarded

namespace {

int ceil_log2(long long v) {
    int b = 0;
    while ((1LL << b) < v) ++b;
    return b;
}

void langevin_coeffs(const orbc_config &cfg, double *gamma, double *sigma) {
    // src/integrate_langevin.h:262-266
    for (int k = 0; k < NT; ++k) {
        gamma[k] = 6.0 * M_PI * cfg.eta * kRadius[k];
        sigma[k] = std::sqrt(2.0 * cfg.kBT * gamma[k]) * std::sqrt(3.0 / cfg.dt);
    }
}

}  // namespace

namespace orbc {

void sync(orbc_ctx *c) { ORBC_HIP_CHECK(hipStreamSynchronize(c->stream)); }

unsigned read_flag(orbc_ctx *c, int k) {
    ORBC_HIP_CHECK(hipMemcpyAsync(c->pinned + k, c->flags + k, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    sync(c);
    return c->pinned[k];
}

void zero_flag(orbc_ctx *c, int k) { launch_fill_u32(c->flags + k, 0u, 1, c->stream); }

}  // namespace orbc

namespace {

ForceArgs force_args(orbc_ctx *c) {
    ForceArgs a{};
    a.n = c->n; a.stride = c->stride;
    a.x = c->x[c->cur]; a.nrm = c->nrm[c->cur];
    a.pq = c->pq[c->cur];
    a.type = c->type; a.id = c->id;
    a.nbr = c->nbr; a.nnbr = c->nnbr;
    a.bnd = c->nbonds ? c->bnd : nullptr;
    a.nbnd = c->nbonds ? c->nbnd : nullptr;
    a.anc_grid = c->ag;
    a.f32 = c->f32; a.t32 = c->t32; a.f64 = c->f64; a.t64 = c->t64;
    a.e = c->e; a.w = c->w;
    a.swizzle = 1;
    return a;
}

IntegArgs integ_args(orbc_ctx *c) {
    IntegArgs a{};
    a.n = c->n;
    a.x = c->x[c->cur]; a.v = c->v; a.nrm = c->nrm[c->cur]; a.o = c->o;
    a.f32 = c->f32; a.t32 = c->t32; a.f64 = c->f64; a.t64 = c->t64;
    a.type = c->type; a.id = c->id; a.anc = c->anc;
    a.pq = c->mixed() ? c->pq[c->cur] : nullptr;
    a.anc_grid = c->ag;
    a.dt = c->cfg.dt;
    langevin_coeffs(c->cfg, a.gamma, a.sigma);
    a.grp = c->grp; a.sig_group = c->sig_group;
    a.seed = c->cfg.seed;
    a.step = (unsigned long long)c->step;
    for (int d = 0; d < 3; ++d) { a.box_lo[d] = c->cfg.box_lo[d]; a.box_hi[d] = c->cfg.box_hi[d]; }
    a.zeta = c->scal;
    a.ke_partial = c->red;
    a.dr_opt = 5e-2; a.dn_opt = 5e-2;
    return a;
}

// Nose-Hoover mass: cfg.Q, or 0.01 N fixed at the first Nose-Hoover step of
// the current system, as the reference sets it once (`if (!parameter.Q)
// set_Q(n)`, src/runtime_parameter.h:76); a later stray-lipid cleanup that
// shrinks N keeps it
double nh_Q(orbc_ctx *c) {
    if (c->cfg.Q > 0) return c->cfg.Q;
    if (!(c->nh_q > 0)) c->nh_q = 0.01 * (double)c->n;
    return c->nh_q;
}

// ------------------------------------------------------------------ commit
void alloc_store(orbc_ctx *c) {
    const long long n = c->n;
    DevArena &m = c->mem;
    for (int b = 0; b < 2; ++b) {
        c->x[b] = m.alloc<double>(3 * n);
        c->nrm[b] = m.alloc<double>(3 * n);
        c->pq[b] = m.alloc<PQ>(n);
    }
    c->v = m.alloc<double>(3 * n); c->o = m.alloc<double>(3 * n);
    c->v_alt = m.alloc<double>(3 * n); c->o_alt = m.alloc<double>(3 * n);
    if (c->mixed()) {
        c->f32 = m.alloc<float>(3 * n); c->t32 = m.alloc<float>(3 * n);
        c->f32_alt = m.alloc<float>(3 * n); c->t32_alt = m.alloc<float>(3 * n);
        c->fe32 = m.alloc<float>(3 * n); c->te32 = m.alloc<float>(3 * n);
    } else {
        c->f64 = m.alloc<double>(3 * n); c->t64 = m.alloc<double>(3 * n);
        c->f64_alt = m.alloc<double>(3 * n); c->t64_alt = m.alloc<double>(3 * n);
        c->fe = m.alloc<double>(3 * n); c->te = m.alloc<double>(3 * n);
    }
    c->e = m.alloc<double>(n); c->w = m.alloc<double>(n);
    c->rb = m.alloc<float>(3 * n);
    c->anc = m.alloc<uint32_t>(n);
    c->type = m.alloc<uint8_t>(n); c->type_alt = m.alloc<uint8_t>(n);
    c->id = m.alloc<uint32_t>(n); c->id_alt = m.alloc<uint32_t>(n);
    c->slot = m.alloc<uint32_t>(n); c->slot_alt = m.alloc<uint32_t>(n);
    c->inv = m.alloc<uint32_t>(n);
    c->keys = m.alloc<uint32_t>(n); c->vals = m.alloc<uint32_t>(n);
    c->keys_alt = m.alloc<uint32_t>(n); c->vals_alt = m.alloc<uint32_t>(n);
    c->sort_temp = m.alloc<char>(radix_temp_bytes(n));
    c->nnbr = m.alloc<uint32_t>(n);
    c->nbnd = m.alloc<uint8_t>(n);
    c->flags = m.alloc<unsigned>(8);
    c->iflags = m.alloc<int>(8);
    c->red = m.alloc<double>(reduce_blocks(n) + 8);
    c->scal = m.alloc<double>(16);
    c->bbox = m.alloc<double>(6 * 257);
    c->stage = m.alloc<double>(3 * n);
    c->stagef = m.alloc<float>(3 * n);
    c->f_ext = m.alloc<double>(3 * n);
    c->step_dev = m.alloc<unsigned long long>(1);
    c->bud2 = m.alloc<float>(n);
    ORBC_HIP_CHECK(hipMemsetAsync(c->bud2, 0, sizeof(float) * n, c->stream));
    c->anc_dev = m.alloc<Anchor>(1);
    c->cg_dev = m.alloc<Grid>(1);
    ORBC_HIP_CHECK(hipMemsetAsync(c->flags, 0, 8 * sizeof(unsigned), c->stream));
    ORBC_HIP_CHECK(hipMemsetAsync(c->iflags, 0, 8 * sizeof(int), c->stream));
    ORBC_HIP_CHECK(hipMemsetAsync(c->scal, 0, 16 * sizeof(double), c->stream));
    ORBC_HIP_CHECK(hipMemsetAsync(c->e, 0, n * sizeof(double), c->stream));
    ORBC_HIP_CHECK(hipMemsetAsync(c->w, 0, n * sizeof(double), c->stream));
    ORBC_HIP_CHECK(hipMemsetAsync(c->nbnd, 0, n, c->stream));
    ORBC_HIP_CHECK(hipMemsetAsync(c->f_ext, 0, 3 * n * sizeof(double), c->stream));
    if (c->mixed()) {
        ORBC_HIP_CHECK(hipMemsetAsync(c->f32, 0, 3 * n * sizeof(float), c->stream));
        ORBC_HIP_CHECK(hipMemsetAsync(c->t32, 0, 3 * n * sizeof(float), c->stream));
    } else {
        ORBC_HIP_CHECK(hipMemsetAsync(c->f64, 0, 3 * n * sizeof(double), c->stream));
        ORBC_HIP_CHECK(hipMemsetAsync(c->t64, 0, 3 * n * sizeof(double), c->stream));
    }
}

void upload_bonds(orbc_ctx *c) {
    // static per-particle bond lists by upload slot, both ends, ordered by
    // (partner slot, type) so that the summation order is fixed
    std::unordered_map<uint32_t, uint32_t> id2slot;
    id2slot.reserve((size_t)c->np * 2 + 1);
    for (long long u = 0; u < c->np; ++u) id2slot[c->sp[1].id[u]] = (uint32_t)(c->pslot0 + u);
    for (long long u = 0; u < c->nl; ++u) id2slot[c->sp[0].id[u]] = (uint32_t)u;
    const long long nb = (long long)c->b_type.size();
    // slots: lipids [0, nl), proteins [pslot0, pslot0 + np) (pslot0 > nl after a stray cleanup)
    const long long nslot = c->pslot0 + c->np;
    std::vector<std::vector<std::pair<uint32_t, uint8_t>>> lists((size_t)nslot);
    for (long long k = 0; k < nb; ++k) {
        auto a = id2slot.find(c->b_i[k]), b = id2slot.find(c->b_j[k]);
        if (a == id2slot.end() || b == id2slot.end()) throw std::runtime_error("bond refers to an unknown particle id");
        if (c->b_type[k] < 0 || c->b_type[k] >= NBT) throw std::runtime_error("bond type out of range");
        lists[a->second].push_back({b->second, (uint8_t)c->b_type[k]});
        lists[b->second].push_back({a->second, (uint8_t)c->b_type[k]});
    }
    std::vector<uint32_t> start((size_t)nslot + 1, 0), partner;
    std::vector<uint8_t> bt;
    partner.reserve((size_t)(2 * nb));
    bt.reserve((size_t)(2 * nb));
    int kb = 0;
    for (long long s = 0; s < nslot; ++s) {
        auto &L = lists[s];
        std::sort(L.begin(), L.end());
        start[s] = (uint32_t)partner.size();
        for (auto &p : L) { partner.push_back(p.first); bt.push_back(p.second); }
        kb = std::max<int>(kb, (int)L.size());
    }
    start[nslot] = (uint32_t)partner.size();
    if (kb > 255) throw std::runtime_error("more than 255 bonds on one particle");
    c->mem.release(c->bstart); c->mem.release(c->bpartner); c->mem.release(c->btype_d); c->mem.release(c->bnd);
    c->bstart = c->mem.alloc<uint32_t>(start.size());
    c->bpartner = c->mem.alloc<uint32_t>(partner.size());
    c->btype_d = c->mem.alloc<uint8_t>(bt.size());
    ORBC_HIP_CHECK(hipMemcpyAsync(c->bstart, start.data(), start.size() * 4, hipMemcpyHostToDevice, c->stream));
    if (!partner.empty()) {
        ORBC_HIP_CHECK(hipMemcpyAsync(c->bpartner, partner.data(), partner.size() * 4, hipMemcpyHostToDevice, c->stream));
        ORBC_HIP_CHECK(hipMemcpyAsync(c->btype_d, bt.data(), bt.size(), hipMemcpyHostToDevice, c->stream));
    }
    c->kbmax = kb;
    c->nbonds = nb;
    c->bnd = kb > 0 ? c->mem.alloc<uint32_t>((size_t)kb * c->stride) : nullptr;
    sync(c);
}

// ensemble members: per-particle member in device order and the noise
// amplitude of every (member, type) from the member's kBT
void upload_groups(orbc_ctx *c) {
    if (c->group_kbt.empty()) return;
    const long long n = c->n;
    if (!c->grp) {
        c->grp = c->mem.alloc<uint16_t>(n);
        c->grp_alt = c->mem.alloc<uint16_t>(n);
        c->sig_group = c->mem.alloc<double>(c->group_kbt.size() * NT + NT);
    }
    // upload-order image: lipid slots from 0, protein slots from pslot0
    std::vector<uint16_t> img(c->pslot0 + c->np, 0);
    for (int s = 0; s < 2; ++s) {
        const Species &S = c->sp[s];
        const long long base = s == 0 ? 0 : c->pslot0;
        for (long long u = 0; u < S.n; ++u) img[base + u] = S.group.empty() ? 0 : S.group[u];
    }
    const size_t ng = c->group_kbt.size();
    std::vector<double> sig(ng * NT);
    for (size_t g = 0; g < ng; ++g) {
        orbc_config cg = c->cfg;
        cg.kBT = c->group_kbt[g];
        double gam[NT], sg[NT];
        langevin_coeffs(cg, gam, sg);
        for (int t = 0; t < NT; ++t) sig[g * NT + t] = sg[t];
    }
    // the staging buffer holds 3n doubles: ample for n u16 in upload order
    uint16_t *tmp = reinterpret_cast<uint16_t *>(c->stage);
    ORBC_HIP_CHECK(hipMemcpyAsync(tmp, img.data(), 2 * img.size(), hipMemcpyHostToDevice, c->stream));
    launch_gather_u16(n, c->slot, tmp, c->grp, c->stream);
    ORBC_HIP_CHECK(hipMemcpyAsync(c->sig_group, sig.data(), 8 * sig.size(), hipMemcpyHostToDevice, c->stream));
    sync(c);
}

// Frees every device array of the store and forgets their pointers and the
// capacities they were sized for (cell ranges, lists, block halos, bonds,
// members), so that the next commit allocates all of them afresh instead of
// reusing freed memory.
void drop_store(orbc_ctx *c) {
    free_graphs(c);
    voronoi_free(c);
    c->mem.free_all();
    c->live = false;
    for (int b = 0; b < 2; ++b) { c->x[b] = nullptr; c->nrm[b] = nullptr; c->pq[b] = nullptr; }
    c->v = c->o = c->v_alt = c->o_alt = nullptr;
    c->f64 = c->t64 = c->f64_alt = c->t64_alt = nullptr;
    c->f32 = c->t32 = c->f32_alt = c->t32_alt = nullptr;
    c->e = c->w = c->fe = c->te = nullptr;
    c->fe32 = c->te32 = nullptr;
    c->f_ext = nullptr;
    c->rb = nullptr;
    c->anc = nullptr;
    c->type = c->type_alt = nullptr;
    c->id = c->id_alt = c->slot = c->slot_alt = c->inv = nullptr;
    c->keys = c->vals = c->keys_alt = c->vals_alt = nullptr;
    c->sort_temp = nullptr;
    c->cstart = c->cend = nullptr;
    c->ncell_cap = 0;
    c->nbr = c->nnbr = nullptr;
    c->kmax = 0;
    c->halo_ext = c->halo_cnt = c->nbr16 = nullptr;
    c->hord = nullptr;
    c->halo_nblk_cap = c->nbr16_cap = 0;
    c->halo_ok = false;
    c->bnd = nullptr; c->nbnd = nullptr; c->kbmax = 0;
    c->bstart = c->bpartner = nullptr; c->btype_d = nullptr; c->nbonds = 0;
    c->flags = nullptr; c->iflags = nullptr;
    c->red = c->scal = c->bbox = c->stage = nullptr;
    c->stagef = nullptr;
    c->grp = c->grp_alt = nullptr;
    c->sig_group = nullptr;
    c->step_dev = nullptr;
    c->bud2 = nullptr;
    c->anc_dev = nullptr;
    c->cg_dev = nullptr;
}

void commit(orbc_ctx *c) {
    if (c->live) drop_store(c);
    c->nl = c->sp[0].n;
    c->np = c->sp[1].n;
    c->n = c->nl + c->np;
    c->pslot0 = c->nl;
    c->types_present = 0;
    if (c->n <= 0) throw std::runtime_error("no particles uploaded");
    if (c->n >= (1LL << 30)) throw std::runtime_error("more than 2^30 particles");
    c->stride = (c->n + 63) / 64 * 64;
    alloc_store(c);
    const long long n = c->n;
    std::vector<double> hx(3 * n), hv(3 * n), hn(3 * n), ho(3 * n);
    std::vector<uint8_t> ht(n);
    std::vector<uint32_t> hid(n), hslot(n);
    long long k = 0;
    for (int s = 0; s < 2; ++s) {
        const Species &S = c->sp[s];
        for (long long u = 0; u < S.n; ++u, ++k) {
            for (int d = 0; d < 3; ++d) {
                hx[3 * k + d] = S.x[3 * u + d];
                hv[3 * k + d] = S.v.empty() ? 0.0 : S.v[3 * u + d];
                hn[3 * k + d] = S.nrm[3 * u + d];
                ho[3 * k + d] = S.o.empty() ? 0.0 : S.o[3 * u + d];
            }
            const int t = s == 0 ? 0 : S.type[u];
            if (t < 0 || t >= NT || (s == 1 && t == 0)) throw std::runtime_error("protein type must be 1..5");
            ht[k] = (uint8_t)t;
            c->types_present |= 1u << t;
            hid[k] = S.id[u];
            hslot[k] = (uint32_t)k;
        }
    }
    ORBC_HIP_CHECK(hipMemcpyAsync(c->x[0], hx.data(), 24 * n, hipMemcpyHostToDevice, c->stream));
    ORBC_HIP_CHECK(hipMemcpyAsync(c->v, hv.data(), 24 * n, hipMemcpyHostToDevice, c->stream));
    ORBC_HIP_CHECK(hipMemcpyAsync(c->nrm[0], hn.data(), 24 * n, hipMemcpyHostToDevice, c->stream));
    ORBC_HIP_CHECK(hipMemcpyAsync(c->o, ho.data(), 24 * n, hipMemcpyHostToDevice, c->stream));
    ORBC_HIP_CHECK(hipMemcpyAsync(c->type, ht.data(), n, hipMemcpyHostToDevice, c->stream));
    ORBC_HIP_CHECK(hipMemcpyAsync(c->id, hid.data(), 4 * n, hipMemcpyHostToDevice, c->stream));
    ORBC_HIP_CHECK(hipMemcpyAsync(c->slot, hslot.data(), 4 * n, hipMemcpyHostToDevice, c->stream));
    const double z = c->cfg.zeta;
    ORBC_HIP_CHECK(hipMemcpyAsync(c->scal, &z, sizeof(double), hipMemcpyHostToDevice, c->stream));
    c->cur = 0;
    c->live = true;
    c->nh_started = false;
    c->nh_q = 0.0;
    c->nbonds = 0; c->kbmax = 0;
    c->grp = c->grp_alt = nullptr;
    c->sig_group = nullptr;
    upload_groups(c);
    if (!c->b_type.empty()) upload_bonds(c);
    c->bonds_dirty = false;
    c->need_rebuild = true;
    c->kmax = 0;
    sync(c);
    rebuild(c);
}

}  // namespace

namespace orbc {

// list cutoffs and search rings for the current skin
void set_cutoffs(orbc_ctx *c, double cs) {
    float cut2[36];
    int rings[NT];
    const double skin = c->cfg.skin;
    for (int a = 0; a < NT; ++a) {
        double rmax = 0;
        for (int b = 0; b < NT; ++b) {
            const double rc = ff_max_cut(a, b);
            // + 1e-4: the build tests FP32 positions relative to a cell anchor
            // (rounding < 1e-6), so the list stays a superset of the pairs within rc + skin
            cut2[a * NT + b] = rc > 0 ? (float)((rc + skin + 1e-4) * (rc + skin + 1e-4)) : -1.0f;
            // search reach only towards types present in the system
            if (rc > 0 && (c->types_present >> b & 1u)) rmax = std::max(rmax, rc + skin);
        }
        rings[a] = std::max(1, (int)std::ceil(rmax / cs - 1e-12));
    }
    c->far_rings = false;
    for (int a = 0; a < NT; ++a) {
        c->rings[a] = rings[a];
        if ((c->types_present >> a & 1u) && rings[a] >= 2) c->far_rings = true;
    }
    // skin budgets of isolated particles (k_budget): types that search one
    // ring of cells for their list and whose longest cutoff leaves less than a
    // skin to the cell edge (lipids, membrane proteins) search two rings, so
    // a drifting one gets a budget of up to (2 h - rc) / 2; other types keep skin/2
    float pcut[36];
    int brings[NT];
    for (int a = 0; a < NT; ++a) {
        double rmax = 0;
        for (int b = 0; b < NT; ++b) {
            const double rc = (c->types_present >> b & 1u) ? ff_max_cut(a, b) : 0.0;
            pcut[a * NT + b] = (float)rc;
            rmax = std::max(rmax, rc);
        }
        brings[a] = (rings[a] == 1 && rmax > 0 && cs - rmax < 2 * skin) ? 2 : 0;
    }
    set_list_cutoffs(cut2, rings, pcut, brings, c->stream);
}

// gathers every per-particle array of the store through perm (n_out
// entries: perm[i] is the old index of the new particle i)
static void permute_store(orbc_ctx *c, const uint32_t *perm, long long n_out, bool with_ft = true) {
    hipStream_t s = c->stream;
    PermuteArgs p{};
    p.n = n_out; p.perm = perm;
    const int nx = 1 - c->cur;
    p.src3[0] = c->x[c->cur]; p.dst3[0] = c->x[nx];
    p.src3[1] = c->nrm[c->cur]; p.dst3[1] = c->nrm[nx];
    p.src3[2] = c->v; p.dst3[2] = c->v_alt;
    p.src3[3] = c->o; p.dst3[3] = c->o_alt;
    if (!with_ft) {
        // the force accumulators hold nothing the caller can see (inside orbc_run)
        p.n3 = 4; p.nf3 = 0;
    } else if (c->mixed()) {
        p.n3 = 4;
        p.srcf3[0] = c->f32; p.dstf3[0] = c->f32_alt;
        p.srcf3[1] = c->t32; p.dstf3[1] = c->t32_alt;
        p.nf3 = 2;
    } else {
        p.src3[4] = c->f64; p.dst3[4] = c->f64_alt;
        p.src3[5] = c->t64; p.dst3[5] = c->t64_alt;
        p.n3 = 6; p.nf3 = 0;
    }
    p.type_src = c->type; p.type_dst = c->type_alt;
    p.id_src = c->id; p.id_dst = c->id_alt;
    p.slot_src = c->slot; p.slot_dst = c->slot_alt;
    p.aff_src = c->voro.aff; p.aff_dst = c->voro.aff_alt;
    p.grp_src = c->grp; p.grp_dst = c->grp_alt;
    launch_permute(p, s);
    c->cur = nx;
    std::swap(c->v, c->v_alt); std::swap(c->o, c->o_alt);
    if (with_ft) {
        std::swap(c->f32, c->f32_alt); std::swap(c->t32, c->t32_alt);
        std::swap(c->f64, c->f64_alt); std::swap(c->t64, c->t64_alt);
    }
    std::swap(c->type, c->type_alt); std::swap(c->id, c->id_alt); std::swap(c->slot, c->slot_alt);
    if (c->voro.aff) std::swap(c->voro.aff, c->voro.aff_alt);
    if (c->grp) std::swap(c->grp, c->grp_alt);
}

// The pointer pairs a rebuild's permutation swaps (and the double-buffer
// parity it flips), without moving any data: applied twice it is the
// identity.  Used to capture the batch graphs of the store as it will be
// after the next rebuild.
static void toggle_store_parity(orbc_ctx *c) {
    c->cur = 1 - c->cur;
    std::swap(c->v, c->v_alt); std::swap(c->o, c->o_alt);
    std::swap(c->type, c->type_alt); std::swap(c->id, c->id_alt); std::swap(c->slot, c->slot_alt);
    if (c->voro.aff) std::swap(c->voro.aff, c->voro.aff_alt);
    if (c->grp) std::swap(c->grp, c->grp_alt);
}

void rebuild(orbc_ctx *c) {
    const long long n = c->n;
    hipStream_t s = c->stream;
    int per[3];
    double Lbox[3];
    bool any_per = false;
    for (int d = 0; d < 3; ++d) {
        per[d] = c->cfg.periodic[d] != 0;
        Lbox[d] = c->cfg.box_hi[d] - c->cfg.box_lo[d];
        any_per = any_per || per[d];
    }
    // 0. periodic axes: positions into [lo, lo + L)
    if (any_per) launch_wrap(c->x[c->cur], n, c->cfg.box_lo, Lbox, per, s);
    // 1. bounding box
    launch_bbox(c->x[c->cur], n, c->bbox, s);
    double bb[6];
    ORBC_HIP_CHECK(hipMemcpyAsync(bb, c->bbox, sizeof(bb), hipMemcpyDeviceToHost, s));
    sync(c);
    for (int d = 0; d < 6; ++d)
        if (!std::isfinite(bb[d])) throw std::runtime_error("non-finite particle position at neighbour rebuild");
    // 2. binning grid: cell edge = lipid cutoff + skin
    const double cs = kCutLL + c->cfg.skin;
    Grid g{};
    g.cs = cs;
    for (int d = 0; d < 3; ++d) {
        if (per[d]) {
            // cells tile the box exactly, so neighbour cells wrap
            g.per[d] = 1;
            g.lo[d] = c->cfg.box_lo[d];
            g.dim[d] = std::max(1, (int)std::floor(Lbox[d] / cs));
            g.csd[d] = Lbox[d] / (double)g.dim[d];
        } else {
            g.lo[d] = bb[d] - 1e-6 * (1.0 + std::fabs(bb[d]));
            g.dim[d] = std::max(1, (int)std::floor((bb[3 + d] - g.lo[d]) / cs) + 1);
            g.csd[d] = cs;
        }
        g.bits[d] = ceil_log2(g.dim[d]);
    }
    if (g.bits[0] + g.bits[1] + g.bits[2] > 32) throw std::runtime_error("system extent exceeds the 32-bit Morton grid");
    g.ncell = (long long)g.dim[0] * g.dim[1] * g.dim[2];
    if (g.ncell > c->ncell_cap) {
        c->mem.release(c->cstart); c->mem.release(c->cend);
        c->ncell_cap = std::max(g.ncell, (long long)(1.25 * (double)g.ncell));
        c->cstart = c->mem.alloc<uint32_t>(c->ncell_cap);
        c->cend = c->mem.alloc<uint32_t>(c->ncell_cap);
    }
    c->grid = g;
    set_cutoffs(c, cs);
    if (any_per) {
        int rmax = 1;
        for (int a = 0; a < NT; ++a)
            if (c->types_present >> a & 1u) rmax = std::max(rmax, c->rings[a]);
        for (int d = 0; d < 3; ++d)
            if (per[d] && g.dim[d] < 2 * rmax + 1)
                throw std::runtime_error("periodic box too short for the cutoff: need at least 2 rings of cells");
    }
    // 3. anchor grid: the smallest power-of-two spacing h >= 1/4 whose anchors
    // pack in 32 bits.  Open axes use h itself; a periodic axis uses
    // 2^bits >= L/h anchors of spacing L / 2^bits, so differences wrap by bits.
    Anchor ag{};
    bool packed = false;
    for (int k = -2; k < 40 && !packed; ++k) {
        const double h = std::ldexp(1.0, k);
        int bits[3];
        long long A0[3];
        int tot = 0;
        for (int d = 0; d < 3; ++d) {
            if (per[d]) {
                A0[d] = 0;
                bits[d] = std::max(1, ceil_log2((long long)std::ceil(Lbox[d] / h)));
            } else {
                A0[d] = (long long)std::floor(bb[d] / h) - 1;
                const long long span = (long long)std::floor(bb[3 + d] / h) - A0[d] + 2;
                bits[d] = std::max(1, ceil_log2(span));
            }
            tot += bits[d];
        }
        if (tot <= 32) {
            for (int d = 0; d < 3; ++d) {
                ag.hd[d] = per[d] ? Lbox[d] / std::ldexp(1.0, bits[d]) : h;
                ag.hf[d] = (float)ag.hd[d];
                ag.lo[d] = per[d] ? c->cfg.box_lo[d] : 0.0;
                ag.A0[d] = A0[d];
                ag.wrap[d] = per[d] ? (unsigned)(32 - bits[d]) : 0u;
                ag.L[d] = per[d] ? Lbox[d] : 0.0;
                ag.Linv[d] = per[d] ? 1.0 / Lbox[d] : 0.0;
                c->anchor_bits[d] = bits[d];
            }
            ag.shy = bits[0]; ag.shz = bits[0] + bits[1];
            ag.mx = (1u << bits[0]) - 1u; ag.my = (1u << bits[1]) - 1u;
            ag.mz = bits[2] >= 32 ? 0xffffffffu : ((1u << bits[2]) - 1u);
            packed = true;
        }
    }
    if (!packed) throw std::runtime_error("cannot pack anchors");
    c->ag = ag;
    ORBC_HIP_CHECK(hipMemcpyAsync(c->anc_dev, &c->ag, sizeof(Anchor), hipMemcpyHostToDevice, s));
    // 4. Morton sort of the particles by grid cell
    launch_keys(c->x[c->cur], n, g, c->keys, c->vals, s);
    bool alt = false;
    radix_sort_pairs(c->keys, c->vals, c->keys_alt, c->vals_alt, n, g.bits[0] + g.bits[1] + g.bits[2], c->sort_temp, s,
                     &alt);
    const uint32_t *perm = alt ? c->vals_alt : c->vals;
    // 5. permute the store
    permute_store(c, perm, n, !c->ft_scratch);
    launch_inverse(c->slot, c->inv, n, s);
    // 6. mirrors, anchors, build-time offsets
    launch_refresh(n, c->x[c->cur], c->nrm[c->cur], c->type, ag, c->pq[c->cur], c->rb, c->anc, s);
    // 7. cell ranges and the Verlet list
    launch_cell_ranges(c->x[c->cur], n, g, c->cstart, c->cend, s);
    // occupied cells (the sort buffers are free again after the permutation)
    const long long nocc = launch_cell_list(c->x[c->cur], n, g, c->keys_alt, c->vals_alt, c->keys, c->sort_temp, s);
    for (int attempt = 0; attempt < 3; ++attempt) {
        if (c->kmax == 0) {
            c->kmax = 48;
            c->nbr = c->mem.alloc<uint32_t>((size_t)c->kmax * c->stride);
        }
        NlistArgs a{};
        a.n = n; a.stride = c->stride; a.x = c->x[c->cur]; a.type = c->type; a.g = g;
        a.pq = c->pq[c->cur]; a.anc = ag;
        for (int d = 0; d < 3; ++d) { a.box.lo[d] = c->cfg.box_lo[d]; a.box.hi[d] = c->cfg.box_hi[d];
                                      a.box.L[d] = c->cfg.box_hi[d] - c->cfg.box_lo[d]; a.box.per[d] = 0; }
        a.cstart = c->cstart; a.cend = c->cend; a.nbr = c->nbr; a.nnbr = c->nnbr; a.kmax = c->kmax;
        a.max_count = c->flags + 2; a.skin = c->cfg.skin;
        if (c->cfg.neighbor_mode == ORBC_NEIGHBOR_VORONOI && c->voro.aff) {
            a.aff = c->voro.aff; a.centroids = c->voro.c;
        }
        zero_flag(c, 2);
        // scratch of the outer-shell pass: the cell-list flags and positions are
        // spent, and the sort values after the permutation; flags[7] counts
        const FarScratch fs{c->keys_alt, c->vals_alt, c->vals, c->flags + 7, c->sort_temp};
        launch_nlist(a, c->keys, nocc, c->far_rings, fs, s);
        const unsigned mc = read_flag(c, 2);
        if ((int)mc <= c->kmax) break;
        c->mem.release(c->nbr);
        c->kmax = (int)mc + 16;
        c->nbr = c->mem.alloc<uint32_t>((size_t)c->kmax * c->stride);
    }
    // 7b. block halos of the fused steps.  Automatic mode builds them only while
    // a list serves several steps: the build costs about what three steps
    // save (full RBC: 0.39 ms per build against 0.12 ms per step).
    c->halo_ok = false;
    const double life = c->cfg.rebuild_period > 0 ? (double)c->cfg.rebuild_period : c->list_life;
    const bool halo = c->cfg.block_halo == 2 || (c->cfg.block_halo == 1 && (life <= 0.0 || life >= 4.0));
    if (halo) {
        const long long nb = (n + BLOCK - 1) / BLOCK;
        const long long words = halo_words(c->kmax) * c->stride;
        if (nb > c->halo_nblk_cap) {
            c->mem.release(c->halo_ext); c->mem.release(c->halo_cnt); c->mem.release(c->hord);
            c->halo_nblk_cap = std::max(nb, (c->stride + BLOCK - 1) / BLOCK);
            c->halo_ext = c->mem.alloc<uint32_t>((size_t)c->halo_nblk_cap * HALO_EXT_MAX);
            c->halo_cnt = c->mem.alloc<uint32_t>((size_t)c->halo_nblk_cap);
            c->hord = c->mem.alloc<uint8_t>((size_t)c->halo_nblk_cap * BLOCK);
        }
        if (words > c->nbr16_cap) {
            c->mem.release(c->nbr16);
            c->nbr16_cap = words;
            c->nbr16 = c->mem.alloc<uint32_t>((size_t)words);
        }
        launch_halo(n, c->stride, c->nbr, c->nnbr, c->type, c->halo_ext, c->halo_cnt, c->nbr16, c->hord, c->halo_cap, s);
        c->halo_cap_built = c->halo_cap;
        c->halo_ok = true;
    }
    // 7c. skin budgets of the particles left without list entries (open boxes)
    if (!any_per) {
        // scratch as for the outer-shell pass (free again after the list build)
        const FarScratch fs{c->keys_alt, c->vals_alt, c->vals, c->flags + 7, c->sort_temp};
        launch_budget(n, c->x[c->cur], c->type, c->nnbr, g, c->cstart, c->cend, (float)(0.5 * c->cfg.skin), c->bud2,
                      fs, s);
    }
    // 8. bond lists
    if (c->nbonds > 0)
        launch_bondlist(n, c->stride, c->slot, c->inv, c->bstart, c->bpartner, c->btype_d, c->bnd, c->nbnd, s);
    zero_flag(c, 0);
    ORBC_HIP_CHECK(hipMemsetAsync(c->flags + 4, 0, 3 * sizeof(unsigned), s));
    ORBC_HIP_CHECK(hipMemsetAsync(c->iflags, 0, sizeof(int), s));
    // 9. centroid grid of the Voronoi cells (kept between rebuilds, see VoroArgs::cdisp)
    voronoi_rebuild_index(c);
    c->need_rebuild = false;
    c->moved = false;
    c->since_build = 0;
    c->rebuilds++;
}

// Removes the particles with del[i] != 0 (lipids only) from the store: the
// kept ones keep their order, lipid upload slots close up (proteins keep
// theirs, so bond data indexed by protein slot stay valid), the host copies
// of the lipid species drop the same entries, and the list is rebuilt.
void compact_store(orbc_ctx *c, const uint32_t *del, long long ndel) {
    if (ndel <= 0) return;
    const long long n = c->n, n_new = n - ndel;
    hipStream_t s = c->stream;
    // scratch: keys (slot flags), vals (their scan), keys_alt (keep), vals_alt (positions), inv (perm)
    launch_compact(n, del, c->slot, (uint32_t)c->pslot0, c->keys, c->vals, c->keys_alt, c->vals_alt, c->inv,
                   c->sort_temp, s);
    std::vector<uint32_t> gone(c->nl);
    if (c->nl > 0)
        ORBC_HIP_CHECK(hipMemcpyAsync(gone.data(), c->keys, 4 * c->nl, hipMemcpyDeviceToHost, s));
    permute_store(c, c->inv, n_new);
    sync(c);
    Species &L = c->sp[ORBC_LIPID];
    long long w = 0;
    for (long long u = 0; u < L.n; ++u) {
        if (gone[u]) continue;
        for (int d = 0; d < 3; ++d) {
            L.x[3 * w + d] = L.x[3 * u + d]; L.v[3 * w + d] = L.v[3 * u + d];
            L.nrm[3 * w + d] = L.nrm[3 * u + d]; L.o[3 * w + d] = L.o[3 * u + d];
        }
        L.type[w] = L.type[u]; L.id[w] = L.id[u];
        if (!L.group.empty()) L.group[w] = L.group[u];
        ++w;
    }
    if (w != L.n - ndel) throw std::runtime_error("stray cleanup: inconsistent lipid count");
    L.n = w;
    L.x.resize(3 * w); L.v.resize(3 * w); L.nrm.resize(3 * w); L.o.resize(3 * w);
    L.type.resize(w); L.id.resize(w);
    if (!L.group.empty()) L.group.resize(w);
    c->n = n_new;
    c->nl -= ndel;
    c->deleted += ndel;
    rebuild(c);
}

void ensure_live(orbc_ctx *c) {
    if (!c->live) commit(c);
    if (c->bonds_dirty) { upload_bonds(c); c->bonds_dirty = false; c->need_rebuild = true; }
    if (c->need_rebuild) rebuild(c);
}

// make the list valid for the current positions (skin criterion)
void ensure_list(orbc_ctx *c) {
    ensure_live(c);
    if (c->moved) {
        zero_flag(c, 0);
        const bool per = c->cfg.periodic[0] || c->cfg.periodic[1] || c->cfg.periodic[2];
        launch_disp(c->n, c->x[c->cur], c->anc, c->rb, c->ag, per ? nullptr : c->bud2, c->nnbr, c->flags, c->stream);
        c->moved = false;
    }
    const unsigned bits = read_flag(c, 0);
    float d2;
    std::memcpy(&d2, &bits, 4);
    const double half = 0.5 * c->cfg.skin;
    if (!(d2 <= half * half)) rebuild(c);
}

}  // namespace orbc

// =================================================================== C ABI
extern "C" {

void orbc_config_default(orbc_config *cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->precision = ORBC_MIXED;
    cfg->kBT = 0.22;            // src/runtime_parameter.h:49
    cfg->eta = 0.01;            // :50
    cfg->dt = 1e-2;             // :55
    cfg->rho = 1.05;            // :48
    cfg->voronoi_cell_size = 14;// :47
    cfg->stray_tolerance = 2.5; // :51
    cfg->freq_voronoi = 2;      // :59
    cfg->freq_sort_ctrd = 24;   // :60
    cfg->freq_sort_bond = 120;  // :61
    cfg->freq_cleanup = 60;     // :62
    cfg->seed = 0xBAD5EED;      // :64
    cfg->Q = 0.0;
    cfg->zeta = 0.0;
    for (int d = 0; d < 3; ++d) { cfg->box_lo[d] = -1000; cfg->box_hi[d] = 1000; }   // :111-115
    cfg->skin = 0.4;
    cfg->rebuild_period = 0;
    cfg->neighbor_mode = ORBC_NEIGHBOR_EXACT;
    cfg->thermostat = ORBC_RUN_LANGEVIN;
    cfg->volume_target = 0.0;
    cfg->volume_strength = 0.05;   // src/openrbc.cpp:229
    cfg->device = 0;
    cfg->block_size = 0;
    cfg->block_halo = 1;
    cfg->graphs = 1;
}

static bool config_valid(const orbc_config &cfg_) {
    if (cfg_.dt <= 0 || cfg_.skin <= 0 || cfg_.kBT < 0 || cfg_.eta < 0) return false;
    if (cfg_.block_halo < 0 || cfg_.block_halo > 2) return false;
    if (cfg_.graphs < 0 || cfg_.graphs > 1) return false;
    if (cfg_.block_size != 0 && cfg_.block_size != BLOCK) return false;   // the kernels are built for 256
    if (cfg_.voronoi_cell_size <= 0) return false;                         // lipids per Voronoi cell
    if (cfg_.precision != ORBC_FP64 && cfg_.precision != ORBC_MIXED) return false;
    for (int d = 0; d < 3; ++d)
        if (cfg_.periodic[d] && !(cfg_.box_hi[d] > cfg_.box_lo[d])) return false;
    // src/runtime_parameter.h:118-120 (reported, not exit(0))
    if (cfg_.freq_voronoi <= 0 || cfg_.freq_sort_ctrd % cfg_.freq_voronoi || cfg_.freq_sort_bond % cfg_.freq_voronoi ||
        cfg_.freq_cleanup % cfg_.freq_voronoi)
        return false;
    return true;
}

int orbc_create(const orbc_config *cfg, orbc_ctx **out) {
    if (!out) return ORBC_ERR_INVALID;
    *out = nullptr;
    orbc_config c0;
    orbc_config_default(&c0);
    const orbc_config &cfg_ = cfg ? *cfg : c0;
    if (!config_valid(cfg_)) return ORBC_ERR_INVALID;
    orbc_ctx *c = new orbc_ctx();
    c->cfg = cfg_;
    hipError_t e = hipSetDevice(cfg_.device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipHostMalloc((void **)&c->pinned, 64, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void **)&c->pinned_d, 64 * sizeof(double), hipHostMallocDefault);
    if (e != hipSuccess) {
        delete c;
        return ORBC_ERR_HIP;
    }
    // test hook: the Nose-Hoover run loop one step at a time (the batched loop
    // must give the same result)
    const char *nh_env = std::getenv("ORBC_NH_PER_STEP");
    c->nh_per_step = nh_env && nh_env[0] == '1';
    ff_fill(c->ffd, c->fff);
    try {
        upload_ff(c->ffd, c->fff, c->stream);
    } catch (const HipError &) {
        delete c;
        return ORBC_ERR_HIP;
    }
    // the fixed scratch of the allocation-free sort: its address on this
    // device, looked up now with the device current (orbc_sort_pairs_noalloc
    // reports a missing one)
    try {
        c->scratch = scratch_address(cfg_.device);
    } catch (const std::exception &) {
        c->scratch = nullptr;
    }
    *out = c;
    return ORBC_OK;
}

}  // extern "C"

namespace orbc {

// a new configuration for an existing context (restart files): the device
// state is dropped and the run counters reset; the caller uploads the state
int orbc_set_config_internal(orbc_ctx *c, const orbc_config &cfg) {
    if (!config_valid(cfg) || cfg.device != c->cfg.device) return fail(c, ORBC_ERR_INVALID, "invalid configuration");
    if (c->live) drop_store(c);
    c->cfg = cfg;
    c->group_kbt.clear();
    for (auto &S : c->sp) S.group.clear();
    c->step = 0; c->nh_started = false; c->nh_q = 0.0; c->deleted = 0; c->last_cleanup = -1;
    c->vol_target_set = false; c->list_life = 0.0; c->since_build = 0;
    return ORBC_OK;
}

// cells of every particle, given per species in upload order
void voronoi_set_affiliation(orbc_ctx *c, const int *aff_lipid, const int *aff_protein) {
    std::vector<uint32_t> hs(c->n);
    ORBC_HIP_CHECK(hipMemcpyAsync(hs.data(), c->slot, 4 * c->n, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    std::vector<int32_t> hv(c->n);
    for (long long i = 0; i < c->n; ++i) {
        const long long s = hs[i];
        hv[i] = s < c->pslot0 ? aff_lipid[s] : aff_protein[s - c->pslot0];
    }
    ORBC_HIP_CHECK(hipMemcpyAsync(c->voro.aff, hv.data(), 4 * c->n, hipMemcpyHostToDevice, c->stream));
    c->voro.aff_valid = true;
    if (c->voro.csize) launch_fill_u32(c->voro.csize + c->voro.ncell, 0u, 1, c->stream);   // counts no longer current
    sync(c);
}

}  // namespace orbc

extern "C" {

int orbc_destroy(orbc_ctx *c) {
    if (!c) return ORBC_ERR_INVALID;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_store(c);
    for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->pinned_d) (void)hipHostFree(c->pinned_d);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return ORBC_OK;
}

const char *orbc_last_error(const orbc_ctx *c) { return c ? c->err.c_str() : "null context"; }

int orbc_get_config(const orbc_ctx *c, orbc_config *cfg) {
    if (!c || !cfg) return ORBC_ERR_INVALID;
    *cfg = c->cfg;
    return ORBC_OK;
}

int orbc_set_state(orbc_ctx *c, int species, long long n, const double *x, const double *v, const double *nrm,
                   const double *o, const int *type, const unsigned *id) {
    return guarded(c, [&] {
        if (species != ORBC_LIPID && species != ORBC_PROTEIN) return fail(c, ORBC_ERR_INVALID, "bad species");
        if (n < 0 || (n > 0 && (!x || !nrm || !id))) return fail(c, ORBC_ERR_INVALID, "x, nrm and id are required");
        if (species == ORBC_PROTEIN && n > 0 && !type) return fail(c, ORBC_ERR_INVALID, "protein type is required");
        Species &S = c->sp[species];
        bool same = c->live && S.n == n;
        if (same && n > 0) {
            // the device keeps tags and types per slot: a re-upload that permutes
            // them (the reference reorders its containers, src/reorder.h:73-149)
            // re-commits the store, so bonds, noise keys and force-field rows follow
            if (!std::equal(id, id + n, S.id.begin())) same = false;
            if (species == ORBC_PROTEIN && !std::equal(type, type + n, S.type.begin())) same = false;
        }
        c->vol_target_set = false;          // a volume target "< 0" is re-read from the new state
        S.n = n;
        S.x.assign(x, x + 3 * n);
        S.nrm.assign(nrm, nrm + 3 * n);
        if (v) S.v.assign(v, v + 3 * n); else S.v.assign(3 * n, 0.0);
        if (o) S.o.assign(o, o + 3 * n); else S.o.assign(3 * n, 0.0);
        S.type.assign(n, 0);
        if (species == ORBC_PROTEIN)
            for (long long k = 0; k < n; ++k) {
                if (type[k] < 1 || type[k] >= NT) return fail(c, ORBC_ERR_INVALID, "protein type must be 1..5");
                S.type[k] = type[k];
            }
        S.id.assign(id, id + n);
        if (same && n > 0) {
            // update in place: upload in slot order, gather into device order
            const long long off = species == ORBC_LIPID ? 0 : c->pslot0;
            double *fields[4] = {c->x[c->cur], c->v, c->nrm[c->cur], c->o};
            const std::vector<double> *src[4] = {&S.x, &S.v, &S.nrm, &S.o};
            for (int f = 0; f < 4; ++f) {
                // full upload-order image: fetch current, overwrite this species, gather back
                launch_scatter_slots(c->n, c->slot, fields[f], c->stage, c->stream);
                ORBC_HIP_CHECK(hipMemcpyAsync(c->stage + 3 * off, src[f]->data(), 24 * n, hipMemcpyHostToDevice, c->stream));
                launch_gather_slots(c->n, c->slot, c->stage, fields[f], c->stream);
            }
            launch_refresh(c->n, c->x[c->cur], c->nrm[c->cur], c->type, c->ag, c->mixed() ? c->pq[c->cur] : nullptr,
                           c->stagef, c->anc, c->stream);
            c->need_rebuild = true;
            sync(c);
        } else if (c->live) {
            drop_store(c);     // counts, tags or types changed: re-commit on next use
        }
        return (int)ORBC_OK;
    });
}

int orbc_set_groups(orbc_ctx *c, int species, long long n, const int *group, int n_groups, const double *kBT) {
    return guarded(c, [&] {
        if (species != ORBC_LIPID && species != ORBC_PROTEIN) return fail(c, ORBC_ERR_INVALID, "bad species");
        Species &S = c->sp[species];
        if (n != S.n) return fail(c, ORBC_ERR_INVALID, "group array length differs from the species' particle count");
        if (n_groups < 1 || n_groups > 65535 || !kBT) return fail(c, ORBC_ERR_INVALID, "n_groups must be 1..65535");
        for (int g = 0; g < n_groups; ++g)
            if (!(kBT[g] >= 0)) return fail(c, ORBC_ERR_INVALID, "negative member temperature");
        if (!c->group_kbt.empty() && (int)c->group_kbt.size() != n_groups)
            return fail(c, ORBC_ERR_INVALID, "n_groups differs from the earlier call");
        std::vector<uint16_t> gv((size_t)n);
        for (long long k = 0; k < n; ++k) {
            if (!group || group[k] < 0 || group[k] >= n_groups) return fail(c, ORBC_ERR_INVALID, "group out of range");
            gv[k] = (uint16_t)group[k];
        }
        S.group.swap(gv);
        c->group_kbt.assign(kBT, kBT + n_groups);
        if (c->live) upload_groups(c);
        return (int)ORBC_OK;
    });
}

int orbc_set_bonds(orbc_ctx *c, long long nb, const int *type, const unsigned *id_i, const unsigned *id_j) {
    return guarded(c, [&] {
        if (nb < 0 || (nb > 0 && (!type || !id_i || !id_j))) return fail(c, ORBC_ERR_INVALID, "bad bond arrays");
        c->b_type.assign(type, type + nb);
        c->b_i.assign(id_i, id_i + nb);
        c->b_j.assign(id_j, id_j + nb);
        c->bonds_dirty = c->live;
        return (int)ORBC_OK;
    });
}

int orbc_get_state(orbc_ctx *c, int species, double *x, double *v, double *nrm, double *o) {
    return guarded(c, [&] {
        if (species != ORBC_LIPID && species != ORBC_PROTEIN) return fail(c, ORBC_ERR_INVALID, "bad species");
        const Species &S = c->sp[species];
        if (!c->live) {
            if (x) std::memcpy(x, S.x.data(), 24 * S.n);
            if (v) std::memcpy(v, S.v.data(), 24 * S.n);
            if (nrm) std::memcpy(nrm, S.nrm.data(), 24 * S.n);
            if (o) std::memcpy(o, S.o.data(), 24 * S.n);
            return (int)ORBC_OK;
        }
        const long long off = species == ORBC_LIPID ? 0 : c->pslot0;
        const long long n = species == ORBC_LIPID ? c->nl : c->np;
        const double *fields[4] = {c->x[c->cur], c->v, c->nrm[c->cur], c->o};
        double *outs[4] = {x, v, nrm, o};
        for (int f = 0; f < 4; ++f) {
            if (!outs[f] || n == 0) continue;
            launch_scatter_slots(c->n, c->slot, fields[f], c->stage, c->stream);
            ORBC_HIP_CHECK(hipMemcpyAsync(outs[f], c->stage + 3 * off, 24 * n, hipMemcpyDeviceToHost, c->stream));
            sync(c);
        }
        return (int)ORBC_OK;
    });
}

int orbc_get_forces(orbc_ctx *c, int species, double *f, double *t, double *e, double *w) {
    return guarded(c, [&] {
        if (species != ORBC_LIPID && species != ORBC_PROTEIN) return fail(c, ORBC_ERR_INVALID, "bad species");
        if (!c->live) return fail(c, ORBC_ERR_STATE, "no device state");
        const long long off = species == ORBC_LIPID ? 0 : c->pslot0;
        const long long n = species == ORBC_LIPID ? c->nl : c->np;
        if (n == 0) return (int)ORBC_OK;
        for (int k = 0; k < 2; ++k) {
            double *out = k == 0 ? f : t;
            if (!out) continue;
            if (c->mixed()) launch_scatter_slots_f(c->n, c->slot, k == 0 ? c->f32 : c->t32, c->stage, c->stream);
            else launch_scatter_slots(c->n, c->slot, k == 0 ? c->f64 : c->t64, c->stage, c->stream);
            ORBC_HIP_CHECK(hipMemcpyAsync(out, c->stage + 3 * off, 24 * n, hipMemcpyDeviceToHost, c->stream));
            sync(c);
        }
        for (int k = 0; k < 2; ++k) {
            double *out = k == 0 ? e : w;
            if (!out) continue;
            launch_scatter_slots1(c->n, c->slot, k == 0 ? c->e : c->w, c->stage, c->stream);
            ORBC_HIP_CHECK(hipMemcpyAsync(out, c->stage + off, 8 * n, hipMemcpyDeviceToHost, c->stream));
            sync(c);
        }
        return (int)ORBC_OK;
    });
}

int orbc_set_forces(orbc_ctx *c, int species, const double *f, const double *t) {
    return guarded(c, [&] {
        if (species != ORBC_LIPID && species != ORBC_PROTEIN) return fail(c, ORBC_ERR_INVALID, "bad species");
        ensure_live(c);
        const long long off = species == ORBC_LIPID ? 0 : c->pslot0;
        const long long n = species == ORBC_LIPID ? c->nl : c->np;
        if (n == 0) return (int)ORBC_OK;
        for (int k = 0; k < 2; ++k) {
            const double *src = k == 0 ? f : t;
            if (!src) continue;
            if (c->mixed()) {
                launch_scatter_slots_f(c->n, c->slot, k == 0 ? c->f32 : c->t32, c->stage, c->stream);
                ORBC_HIP_CHECK(hipMemcpyAsync(c->stage + 3 * off, src, 24 * n, hipMemcpyHostToDevice, c->stream));
                launch_gather_slots_f(c->n, c->slot, c->stage, k == 0 ? c->f32 : c->t32, c->stream);
            } else {
                double *dst = k == 0 ? c->f64 : c->t64;
                launch_scatter_slots(c->n, c->slot, dst, c->stage, c->stream);
                ORBC_HIP_CHECK(hipMemcpyAsync(c->stage + 3 * off, src, 24 * n, hipMemcpyHostToDevice, c->stream));
                launch_gather_slots(c->n, c->slot, c->stage, dst, c->stream);
            }
        }
        sync(c);
        return (int)ORBC_OK;
    });
}

int orbc_build_neighbors(orbc_ctx *c) {
    return guarded(c, [&] {
        ensure_live(c);
        rebuild(c);
        sync(c);
        return (int)ORBC_OK;
    });
}

static void clear_ft(orbc_ctx *c, bool energy) {
    const long long n = c->n;
    if (c->mixed()) {
        ORBC_HIP_CHECK(hipMemsetAsync(c->f32, 0, 12 * n, c->stream));
        ORBC_HIP_CHECK(hipMemsetAsync(c->t32, 0, 12 * n, c->stream));
    } else {
        ORBC_HIP_CHECK(hipMemsetAsync(c->f64, 0, 24 * n, c->stream));
        ORBC_HIP_CHECK(hipMemsetAsync(c->t64, 0, 24 * n, c->stream));
    }
    if (energy) {
        ORBC_HIP_CHECK(hipMemsetAsync(c->e, 0, 8 * n, c->stream));
        ORBC_HIP_CHECK(hipMemsetAsync(c->w, 0, 8 * n, c->stream));
    }
}

int orbc_compute_pairwise(orbc_ctx *c, unsigned flags) {
    return guarded(c, [&] {
        ensure_list(c);
        const bool energy = flags & ORBC_F_ENERGY;
        if (flags & ORBC_F_ZERO_FIRST) clear_ft(c, energy);
        launch_forces(force_args(c), c->mixed(), energy, true, false, c->stream);
        return (int)ORBC_OK;
    });
}

int orbc_compute_bonded(orbc_ctx *c, unsigned flags) {
    return guarded(c, [&] {
        ensure_live(c);
        const bool energy = flags & ORBC_F_ENERGY;
        if (flags & ORBC_F_ZERO_FIRST) clear_ft(c, energy);
        if (c->nbonds > 0) launch_forces(force_args(c), c->mixed(), energy, false, true, c->stream);
        return (int)ORBC_OK;
    });
}

int orbc_integrate(orbc_ctx *c, int kernel) {
    return guarded(c, [&] {
        if (kernel < ORBC_CLEAR_FORCE || kernel > ORBC_BOUNCE_BACK) return fail(c, ORBC_ERR_INVALID, "bad kernel");
        ensure_live(c);
        IntegArgs a = integ_args(c);
        launch_integrate(kernel, a, c->mixed(), c->stream);
        if (kernel == ORBC_NH_INITIAL_BOUNCE_CLEAR || kernel == ORBC_NH_POST_TORQUE_FINAL)
            launch_zeta_update(c->red, reduce_blocks(c->n), c->scal, c->cfg.dt, nh_Q(c), 1.5 * (double)c->n * c->cfg.kBT,
                               c->stream);
        if (kernel == ORBC_VERLET_LANGEVIN) c->step++;
        if (kernel == ORBC_VERLET_LANGEVIN || kernel == ORBC_NH_INITIAL_BOUNCE_CLEAR || kernel == ORBC_BOUNCE_BACK)
            c->moved = true;
        return (int)ORBC_OK;
    });
}

int orbc_compute_temperature(orbc_ctx *c, double *T) {
    return guarded(c, [&] {
        if (!T) return fail(c, ORBC_ERR_INVALID, "null T");
        ensure_live(c);
        launch_reduce_ke(c->v, c->type, c->n, c->red, c->stream);
        launch_sum_partials(c->red, reduce_blocks(c->n), c->scal + 1, c->stream);
        double ke;
        ORBC_HIP_CHECK(hipMemcpyAsync(&ke, c->scal + 1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        sync(c);
        *T = 2.0 * ke / (3.0 * (double)c->n);
        return (int)ORBC_OK;
    });
}

static StepArgs step_args(orbc_ctx *c) {
    StepArgs a{};
    const int in = c->cur, out = 1 - c->cur;
    a.n = c->n; a.stride = c->stride;
    a.x_in = c->x[in]; a.n_in = c->nrm[in];
    a.x_out = c->x[out]; a.n_out = c->nrm[out];
    a.v = c->v; a.o = c->o;
    a.pq_in = c->pq[in];
    a.pq_out = c->pq[out];
    a.type = c->type; a.id = c->id; a.anc = c->anc; a.rb = c->rb;
    a.nbr = c->nbr; a.nnbr = c->nnbr;
    if (c->halo_ok) {
        a.halo_ext = c->halo_ext; a.halo_cnt = c->halo_cnt; a.nbr16 = c->nbr16; a.hord = c->hord;
        a.halo_cap = c->halo_cap_built;
    }
    a.bnd = c->nbonds ? c->bnd : nullptr;
    a.nbnd = c->nbonds ? c->nbnd : nullptr;
    a.has_bonds = c->nbonds > 0;
    a.f_ext = nullptr;
    a.anc_grid = c->ag;
    a.anc_dev = c->anc_dev;
    bool per = false;
    for (int d = 0; d < 3; ++d) per = per || c->cfg.periodic[d];
    a.bud2 = per ? nullptr : c->bud2;
    a.dt = c->cfg.dt;
    langevin_coeffs(c->cfg, a.gamma, a.sigma);
    a.grp = c->grp; a.sig_group = c->sig_group;
    a.seed = c->cfg.seed;
    a.step = (unsigned long long)c->step;
    a.dslot = c->flags + 4;
    a.half_skin2 = (float)(0.25 * c->cfg.skin * c->cfg.skin);
    a.stop = c->iflags;
    a.done = c->iflags + 1;
    a.bad = c->flags + 1;
    a.swizzle = 1;
    return a;
}

// time a kernel sequence with HIP events on the context stream
static void ev_pair(orbc_ctx *c, size_t k) {
    while (c->ev.size() < 2 * (k + 1)) {
        hipEvent_t e;
        ORBC_HIP_CHECK(hipEventCreate(&e));
        c->ev.push_back(e);
    }
}

static bool volume_on(const orbc_ctx *c) { return c->cfg.volume_target != 0.0 && c->cfg.volume_strength != 0.0; }

// Voronoi update every freq_voronoi steps (src/openrbc.cpp:200-205) and, with
// the volume constraint on, the volume evaluation of constrain_volume
// (src/constrain_volume.h:26-83) before the force evaluation of step `step`.
// Inside a batch every kernel of it is gated on the step's own skip
// condition, and the fused step kernel applies the push; otherwise (to_f) the
// push is added to the force accumulators.
static void cell_prestep(orbc_ctx *c, long long step, const Gate &gate, bool to_f, bool update_cells = true) {
    c->voro.gate = gate;
    if (update_cells && step % c->cfg.freq_voronoi == 0) voronoi_update_dev(c);
    if (volume_on(c)) {
        if (to_f)
            volume_push(c, c->vol_target, c->cfg.volume_strength, nullptr, true, false);
        else
            volume_eval(c, c->vol_target, c->cfg.volume_strength);
    }
    c->voro.gate = Gate{};
}

// orbc_run keeps Voronoi cells current when they exist and something uses
// them (the volume constraint, the stray-lipid cleanup)
static bool cells_in_run(const orbc_ctx *c) {
    return c->voro.ncell > 0 && (volume_on(c) || c->cfg.freq_cleanup > 0);
}

// delete_lipid runs on Voronoi-update steps that are multiples of
// freq_cleanup, before the update (src/openrbc.cpp:200-202)
static long long cleanup_period(const orbc_ctx *c) {
    if (!cells_in_run(c) || c->cfg.freq_cleanup <= 0) return 0;
    return std::lcm((long long)c->cfg.freq_voronoi, (long long)c->cfg.freq_cleanup);
}

static void cleanup_if_due(orbc_ctx *c) {
    if (c->tuning) return;
    const long long L = cleanup_period(c);
    if (L <= 0 || c->step % L != 0 || c->last_cleanup == c->step) return;
    delete_stray_lipids(c, c->cfg.stray_tolerance);
    c->last_cleanup = c->step;
}

static void volume_begin(orbc_ctx *c) {
    if (!volume_on(c) || c->vol_target_set) return;
    if (c->cfg.volume_target > 0) {
        c->vol_target = c->cfg.volume_target;
    } else {
        c->vol_target = volume_push(c, 1.0, 0.0, nullptr, false, true);
        if (!(c->vol_target > 0)) throw std::runtime_error("enclosed volume is not positive");
    }
    c->vol_target_set = true;
}

// One step of a batch: the gated Voronoi / volume work and the fused step
// kernel of step `step`.  step0 null: numbered on the host (k = step);
// otherwise the step is *step0 + k on the device (captured graphs).
static void enqueue_step(orbc_ctx *c, long long step, unsigned long long k, const unsigned long long *step0,
                         hipEvent_t e0, hipEvent_t e1) {
    const Gate gate{c->iflags, c->flags + 4, step0, k};
    if (cells_in_run(c)) cell_prestep(c, step, gate, false);
    StepArgs a = step_args(c);
    a.step = k;
    a.step0 = step0;
    if (volume_on(c)) {
        a.push = c->scal + 9;
        a.cell_normal = c->voro.normals;
        a.aff = c->voro.aff;
    }
    if (e0) ORBC_HIP_CHECK(hipEventRecord(e0, c->stream));
    launch_step_langevin(a, c->mixed(), c->stream);
    if (e1) ORBC_HIP_CHECK(hipEventRecord(e1, c->stream));
    c->cur = 1 - c->cur;
}

// Everything a captured batch of `len` steps bakes into its kernel arguments
// besides the device step counter: the step and Voronoi arguments of its
// first step (pointers, constants, parity of the double buffers), the volume
// target, the Voronoi phase of its first step and the timing events.
static std::vector<unsigned char> batch_signature(orbc_ctx *c, int len, long long first_step) {
    std::vector<unsigned char> sig;
    auto put = [&](const void *p, size_t nb) {
        const unsigned char *b = static_cast<const unsigned char *>(p);
        sig.insert(sig.end(), b, b + nb);
    };
    StepArgs a = step_args(c);
    a.anc_grid = Anchor{};          // read from anc_dev by the kernels
    a.step = 0;                     // numbered from the device counter
    put(&a, sizeof(a));
    const int head[8] = {len, c->cur, (int)(first_step % c->cfg.freq_voronoi), (int)cells_in_run(c),
                         (int)volume_on(c), (int)c->voro.aff_valid, (int)c->timing, c->cfg.freq_voronoi};
    put(head, sizeof(head));
    const double vol[2] = {c->vol_target, c->cfg.volume_strength};
    put(vol, sizeof(vol));
    if (cells_in_run(c)) {
        const auto &V = c->voro;
        const void *vp[] = {V.aff, V.aff_alt, V.c, V.sums, V.normals, V.partial, V.cg_start, V.cg_end, V.cg_idx,
                            V.cref, V.cdisp, c->x[0], c->x[1], c->nrm[0], c->nrm[1]};
        put(vp, sizeof(vp));
        put(&V.ncell, sizeof(V.ncell));
    }
    return sig;
}


// The captured batch of `len` steps starting at host step `first_step` with
// the current parity, from the cache or captured now; nullptr when capture is
// not available (the caller launches the steps one by one).
static orbc_ctx::BatchGraph *batch_graph(orbc_ctx *c, int len, long long first_step) {
    if (!c->graphs_ok) return nullptr;
    std::vector<unsigned char> sig = batch_signature(c, len, first_step);
    for (auto &g : c->graphs)
        if (g.sig == sig) { g.last_use = ++c->graph_clock; return &g; }
    if (c->graphs.size() >= 48) {
        auto it = std::min_element(c->graphs.begin(), c->graphs.end(),
                                   [](const orbc_ctx::BatchGraph &x, const orbc_ctx::BatchGraph &y) { return x.last_use < y.last_use; });
        if (it->exec) (void)hipGraphExecDestroy(it->exec);
        if (it->graph) (void)hipGraphDestroy(it->graph);
        c->graphs.erase(it);
    }
    orbc_ctx::BatchGraph g;
    g.sig = std::move(sig);
    const int cur0 = c->cur;
    bool ok = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
    c->capturing = ok;
    if (ok) {
        try {
            for (int k = 0; k < len; ++k)
                enqueue_step(c, first_step + k, (unsigned long long)k, c->step_dev, nullptr, nullptr);
            launch_add_u64(c->step_dev, (unsigned long long)len, c->stream);
        } catch (const HipError &) {
            ok = false;
        }
        c->capturing = false;
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &graph);
        ok = ok && e == hipSuccess && graph;
        if (ok && hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0) != hipSuccess) ok = false;
        g.graph = graph;
    }
    c->cur = cur0;
    if (!ok) {
        (void)hipGetLastError();
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
        c->graphs_ok = false;       // launched step by step from now on
        return nullptr;
    }
    g.last_use = ++c->graph_clock;
    c->graph_captures++;
    c->graphs.push_back(std::move(g));
    return &c->graphs.back();
}

static void run_langevin(orbc_ctx *c, long long nsteps) {
    long long remaining = nsteps;
    // the fused step keeps forces in registers: rebuilds inside the run leave
    // the accumulators alone, and they are cleared when it ends (as
    // integrate(verlet_langevin) leaves f = t = 0, src/integrate_langevin.h:148)
    struct Scratch {
        orbc_ctx *c;
        ~Scratch() {
            c->ft_scratch = false;
            try { clear_ft(c, false); } catch (...) {}      // an error of the run is already on its way out
        }
    } scratch{c};
    c->ft_scratch = true;
    const bool vol = volume_on(c);
    const long long cperiod = cleanup_period(c);
    while (remaining > 0) {
        cleanup_if_due(c);
        long long m = std::min<long long>(c->batch, remaining);
        // a batch ends before the next cleanup step (it needs the host)
        if (cperiod > 0) m = std::min<long long>(m, (c->step / cperiod + 1) * cperiod - c->step);
        if (c->cfg.rebuild_period > 0) {
            m = std::min<long long>(m, std::max<long long>(1, c->cfg.rebuild_period - c->since_build));
        } else if (c->list_life > 0) {
            // launch about as many steps as the list is expected to last: steps
            // after a skin trigger are skipped, but each still costs its launches
            // (and the Voronoi/volume work queued before it)
            const long long margin = vol ? 1 : 4;
            const long long left = (long long)c->list_life - c->since_build + margin;
            m = std::min<long long>(m, std::max<long long>(1, left));
        }
        launch_fill_u32(reinterpret_cast<uint32_t *>(c->iflags + 1), 0u, 1, c->stream);
        const int cur0 = c->cur;
        // captured graphs: the batch as segments of 32, 16, ..., 1 steps;
        // otherwise (or with kernel timing on: events recorded inside a graph
        // cannot be read back here) the steps launched one by one
        struct Seg { orbc_ctx::BatchGraph *g; long long off; int len; };
        std::vector<Seg> segs;
        if (c->cfg.graphs && c->graphs_ok && !c->timing) {
            // capture every length for this parity and the next step's, for
            // the store as it is and as it will be after the next rebuild, at
            // once: a run needs them all soon, and capturing costs host time
            // while the GPU waits, so it is paid once instead of inside later
            // batches (a rebuild swaps the permuted arrays' buffers)
            const long long miss_before = c->graph_captures;
            (void)batch_graph(c, 32, c->step);
            if (c->graph_captures != miss_before && c->graphs_ok) {
                for (int rb = 0; rb < 2 && c->graphs_ok; ++rb) {
                    if (rb) toggle_store_parity(c);
                    const int base = c->cur;
                    for (int par = 0; par < 2 && c->graphs_ok; ++par) {
                        c->cur = par ? 1 - base : base;
                        for (int len = 32; len >= 1 && c->graphs_ok; len >>= 1)
                            (void)batch_graph(c, len, c->step + par);
                    }
                    c->cur = base;
                    if (rb) toggle_store_parity(c);
                }
                c->cur = cur0;
            }
            launch_set_u64(c->step_dev, (unsigned long long)c->step, c->stream);
            long long off = 0;
            for (int len = 32; len >= 1 && c->graphs_ok; len >>= 1) {
                if (m - off < len) continue;
                orbc_ctx::BatchGraph *g = batch_graph(c, len, c->step + off);   // at this segment's parity
                if (!g) break;
                segs.push_back({g, off, len});
                off += len;
                if (len & 1) c->cur = 1 - c->cur;
            }
            c->cur = cur0;
            if (c->graphs_ok) {
                for (const Seg &sg : segs) {
                    ORBC_HIP_CHECK(hipGraphLaunch(sg.g->exec, c->stream));
                    c->graph_replays++;
                    if (sg.len & 1) c->cur = 1 - c->cur;
                }
            } else {
                segs.clear();
            }
        }
        if (segs.empty()) {
            if (c->timing) ev_pair(c, (size_t)m);
            for (long long s = 0; s < m; ++s)
                enqueue_step(c, c->step + s, (unsigned long long)(c->step + s), nullptr,
                             c->timing ? c->ev[2 * s] : nullptr, c->timing ? c->ev[2 * s + 1] : nullptr);
        }
        int fl[2];
        ORBC_HIP_CHECK(hipMemcpyAsync(fl, c->iflags, sizeof(fl), hipMemcpyDeviceToHost, c->stream));
        unsigned bad = 0;
        ORBC_HIP_CHECK(hipMemcpyAsync(&bad, c->flags + 1, sizeof(bad), hipMemcpyDeviceToHost, c->stream));
        sync(c);
        const long long done = fl[1];
        if (c->timing) {
            auto add = [&](hipEvent_t e0, hipEvent_t e1) {
                float ms = 0;
                ORBC_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                c->t_step_ms += ms;
                c->n_step_launches++;
            };
            for (long long s = 0; s < done; ++s) add(c->ev[2 * s], c->ev[2 * s + 1]);
        }
        c->cur = (cur0 + (int)(done & 1)) & 1;
        c->step += done;
        c->since_build += done;
        remaining -= done;
        if (bad) throw std::runtime_error("non-finite state in orbc_run");
        const bool period = c->cfg.rebuild_period > 0 && c->since_build >= c->cfg.rebuild_period;
        if (fl[0]) {
            // running estimate of how many steps a list lasts (skin criterion)
            c->list_life = c->list_life > 0 ? 0.75 * c->list_life + 0.25 * (double)c->since_build
                                            : (double)c->since_build;
        }
        if (fl[0] || period) {
            if (c->timing) { ev_pair(c, 0); ORBC_HIP_CHECK(hipEventRecord(c->ev[0], c->stream)); }
            rebuild(c);
            if (c->timing) {
                ORBC_HIP_CHECK(hipEventRecord(c->ev[1], c->stream));
                ORBC_HIP_CHECK(hipEventSynchronize(c->ev[1]));
                float ms = 0;
                ORBC_HIP_CHECK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
                c->t_rebuild_ms += ms;
                c->n_timed_rebuilds++;
            }
        }
    }
    c->moved = true;
}

// Nose-Hoover (src/openrbc.cpp:176-243 with FUSED_INTEGRATOR) in batches:
// step k = first half (which also flags slot k % 3 when a particle moved past
// its skin budget) + zeta update + fused forces and second half + zeta update,
// every kernel skipping once the run has stopped.  The second half of a step
// whose first half set the flag stops the run instead; the host then rebuilds
// between that step's halves, as the per-step loop does, finishes the step
// and launches the next batch.
static void run_nh_batched(orbc_ctx *c, long long nsteps, double Q) {
    const double nkT = 1.5 * (double)c->n * c->cfg.kBT;    // no cleanup here (no cells): N is fixed
    const long long nb = reduce_blocks(c->n);
    const float half2 = (float)(0.25 * c->cfg.skin * c->cfg.skin);
    bool per = false;
    for (int d = 0; d < 3; ++d) per = per || c->cfg.periodic[d];
    launch_fill_u32(c->flags + 4, 0u, 3, c->stream);
    long long remaining = nsteps;
    while (remaining > 0) {
        long long m = std::min<long long>(c->batch, remaining);
        if (c->list_life > 0)
            m = std::min<long long>(m, std::max<long long>(1, (long long)c->list_life - c->since_build + 4));
        launch_fill_u32(reinterpret_cast<uint32_t *>(c->iflags), 0u, 2, c->stream);
        for (long long s = 0; s < m; ++s) {
            const long long k = c->step + s;
            IntegArgs a = integ_args(c);
            a.rb = c->rb;
            a.stop = c->iflags;
            a.skin_slot = c->flags + 4 + (int)(k % 3);
            a.skin_clear = c->flags + 4 + (int)((k + 1) % 3);
            a.bud2 = per ? nullptr : c->bud2;
            a.nnbr = c->nnbr;
            a.half_skin2 = half2;
            launch_integrate(ORBC_NH_INITIAL_BOUNCE_CLEAR, a, c->mixed(), c->stream);
            launch_zeta_update(c->red, nb, c->scal, c->cfg.dt, Q, nkT, c->stream, c->iflags);
            StepArgs sa = step_args(c);
            sa.step = (unsigned long long)k;
            sa.nh_gate = 1;
            launch_step_nh_second(sa, c->mixed(), c->f32, c->t32, c->f64, c->t64, c->red, c->scal, c->stream);
            launch_zeta_update(c->red, nb, c->scal, c->cfg.dt, Q, nkT, c->stream, c->iflags);
        }
        int fl[2];
        ORBC_HIP_CHECK(hipMemcpyAsync(fl, c->iflags, sizeof(fl), hipMemcpyDeviceToHost, c->stream));
        sync(c);
        const long long done = fl[1];
        c->step += done;
        c->since_build += done;
        remaining -= done;
        if (fl[0]) {
            // the first half of step c->step ran and found the list spent
            c->since_build++;
            c->list_life = c->list_life > 0 ? 0.75 * c->list_life + 0.25 * (double)c->since_build
                                            : (double)c->since_build;
            rebuild(c);
            launch_step_nh_second(step_args(c), c->mixed(), c->f32, c->t32, c->f64, c->t64, c->red, c->scal,
                                  c->stream);
            launch_zeta_update(c->red, nb, c->scal, c->cfg.dt, Q, nkT, c->stream);
            c->step++;
            remaining--;
        }
    }
    c->moved = true;
}

int orbc_run(orbc_ctx *c, long long nsteps) {
    return guarded(c, [&] {
        if (nsteps < 0) return fail(c, ORBC_ERR_INVALID, "nsteps < 0");
        if (volume_on(c) && c->voro.ncell <= 0)
            return fail(c, ORBC_ERR_STATE, "the volume constraint needs Voronoi cells: orbc_voronoi_init first");
        ensure_list(c);
        try {
            volume_begin(c);
        } catch (const std::runtime_error &e) {
            return fail(c, ORBC_ERR_INVALID, e.what());
        }
        if (c->cfg.thermostat == ORBC_RUN_LANGEVIN) {
            try {
                run_langevin(c, nsteps);
            } catch (const std::runtime_error &e) {
                return fail(c, ORBC_ERR_NAN, e.what());
            }
        } else {
            if (!c->group_kbt.empty())
                return fail(c, ORBC_ERR_INVALID, "ensemble members with their own kBT need the Langevin thermostat");
            // Nose-Hoover, src/openrbc.cpp:176-243 with FUSED_INTEGRATOR
            IntegArgs a;
            // N can shrink inside the loop (stray-lipid cleanup): 1.5 N kT and the
            // reduction width follow it
            const double Q = nh_Q(c);
            double nkT = 1.5 * (double)c->n * c->cfg.kBT;
            long long nb = reduce_blocks(c->n);
            const bool cells = cells_in_run(c);
            auto forces = [&](bool update) {
                clear_ft(c, false);
                ForceArgs fa = force_args(c);
                launch_forces(fa, c->mixed(), false, true, c->nbonds > 0, c->stream);
                if (cells) cell_prestep(c, c->step, Gate{}, true, update);
            };
            if (!c->nh_started) {
                forces(false);
                a = integ_args(c);
                launch_integrate(ORBC_NH_POST_TORQUE_FINAL, a, c->mixed(), c->stream);
                launch_zeta_update(c->red, nb, c->scal, c->cfg.dt, Q, nkT, c->stream);
                c->nh_started = true;
            }
            // batched: the skin check runs on the device between the halves and
            // the host synchronises once per batch (the per-step loop below is
            // kept for runs with Voronoi cells, whose volume push joins the
            // unfused forces, and for fixed rebuild periods)
            if (!cells && c->cfg.rebuild_period <= 0 && !c->nh_per_step) {
                run_nh_batched(c, nsteps, Q);
                nsteps = 0;
            }
            for (long long s = 0; s < nsteps; ++s) {
                a = integ_args(c);
                // the first half also measures the displacement since the build
                // (the skin check ensure_list reads), so no separate pass runs
                zero_flag(c, 0);
                a.rb = c->rb; a.disp_max = c->flags;
                launch_integrate(ORBC_NH_INITIAL_BOUNCE_CLEAR, a, c->mixed(), c->stream);
                launch_zeta_update(c->red, nb, c->scal, c->cfg.dt, Q, nkT, c->stream);
                c->moved = false;
                c->since_build++;
                bool rb = c->cfg.rebuild_period > 0 && c->since_build >= c->cfg.rebuild_period;
                if (rb) { rebuild(c); } else { ensure_list(c); }
                cleanup_if_due(c);
                nkT = 1.5 * (double)c->n * c->cfg.kBT;
                nb = reduce_blocks(c->n);
                if (cells) {
                    // the volume push joins the forces before the second half
                    forces(true);
                    a = integ_args(c);
                    launch_integrate(ORBC_NH_POST_TORQUE_FINAL, a, c->mixed(), c->stream);
                } else {
                    // forces and the second half in one pass, same arithmetic
                    launch_step_nh_second(step_args(c), c->mixed(), c->f32, c->t32, c->f64, c->t64, c->red, c->scal,
                                          c->stream);
                }
                launch_zeta_update(c->red, nb, c->scal, c->cfg.dt, Q, nkT, c->stream);
                c->step++;
            }
        }
        sync(c);
        return (int)ORBC_OK;
    });
}

int orbc_minimize(orbc_ctx *c, long long nsteps, double dr_opt, double dn_opt) {
    return guarded(c, [&] {
        for (long long s = 0; s < nsteps; ++s) {
            ensure_list(c);
            clear_ft(c, false);
            launch_forces(force_args(c), c->mixed(), false, true, c->nbonds > 0, c->stream);
            IntegArgs a = integ_args(c);
            a.dr_opt = dr_opt; a.dn_opt = dn_opt;
            launch_integrate(ORBC_POST_TORQUE, a, c->mixed(), c->stream);
            launch_minimize(a, c->mixed(), c->stream);
            launch_integrate(ORBC_BOUNCE_BACK, a, c->mixed(), c->stream);
            c->moved = true;
        }
        clear_ft(c, false);
        sync(c);
        return (int)ORBC_OK;
    });
}

int orbc_get_observables(orbc_ctx *c, orbc_observables *obs) {
    return guarded(c, [&] {
        if (!obs) return fail(c, ORBC_ERR_INVALID, "null obs");
        ensure_list(c);
        std::memset(obs, 0, sizeof(*obs));
        const long long nb = reduce_blocks(c->n);
        launch_reduce_ke(c->v, c->type, c->n, c->red, c->stream);
        launch_sum_partials(c->red, nb, c->scal + 1, c->stream);
        // energy and virial into scratch force buffers (f,t untouched)
        ForceArgs fa = force_args(c);
        if (c->mixed()) {
            ORBC_HIP_CHECK(hipMemsetAsync(c->fe32, 0, 12 * c->n, c->stream));
            ORBC_HIP_CHECK(hipMemsetAsync(c->te32, 0, 12 * c->n, c->stream));
            fa.f32 = c->fe32; fa.t32 = c->te32;
        } else {
            ORBC_HIP_CHECK(hipMemsetAsync(c->fe, 0, 24 * c->n, c->stream));
            ORBC_HIP_CHECK(hipMemsetAsync(c->te, 0, 24 * c->n, c->stream));
            fa.f64 = c->fe; fa.t64 = c->te;
        }
        ORBC_HIP_CHECK(hipMemsetAsync(c->e, 0, 8 * c->n, c->stream));
        ORBC_HIP_CHECK(hipMemsetAsync(c->w, 0, 8 * c->n, c->stream));
        launch_forces(fa, c->mixed(), true, true, c->nbonds > 0, c->stream);
        launch_sum_array(c->e, c->n, c->red, c->stream);
        launch_sum_partials(c->red, nb, c->scal + 2, c->stream);
        launch_sum_array(c->w, c->n, c->red, c->stream);
        launch_sum_partials(c->red, nb, c->scal + 3, c->stream);
        double s[4];
        ORBC_HIP_CHECK(hipMemcpyAsync(s, c->scal, sizeof(s), hipMemcpyDeviceToHost, c->stream));
        sync(c);
        obs->zeta = s[0];
        obs->kinetic = s[1];
        obs->temperature = 2.0 * s[1] / (3.0 * (double)c->n);
        obs->potential = s[2];
        obs->virial = s[3];
        obs->n = c->n;
        obs->n_lipid = c->live ? c->nl : c->sp[ORBC_LIPID].n;
        obs->n_protein = c->live ? c->np : c->sp[ORBC_PROTEIN].n;
        obs->lipids_deleted = c->deleted;
        obs->step = c->step;
        obs->rebuilds = c->rebuilds;
        obs->graph_captures = c->graph_captures;
        obs->graph_replays = c->graph_replays;
        if (c->halo_ok && c->n > 0) {
            const long long nb = (c->n + BLOCK - 1) / BLOCK;
            std::vector<uint32_t> hc(nb);
            ORBC_HIP_CHECK(hipMemcpyAsync(hc.data(), c->halo_cnt, 4 * nb, hipMemcpyDeviceToHost, c->stream));
            sync(c);
            obs->halo_blocks = nb;
            for (uint32_t v : hc) obs->halo_overflow += v == HALO_OVERFLOW;
        }
        // directed list entries
        std::vector<uint32_t> cnt(c->n);
        ORBC_HIP_CHECK(hipMemcpyAsync(cnt.data(), c->nnbr, 4 * c->n, hipMemcpyDeviceToHost, c->stream));
        sync(c);
        long long tot = 0;
        for (auto v : cnt) tot += v;
        obs->pairs = tot;
        obs->anchor_h = 0.0;
        for (int d = 0; d < 3; ++d) {
            if (!c->cfg.periodic[d]) obs->anchor_h = c->ag.hd[d];
            obs->anchor_lo[d] = c->ag.lo[d];
            obs->anchor_spacing[d] = c->ag.hd[d];
            obs->anchor_origin[d] = (double)c->ag.A0[d];
            obs->anchor_bits[d] = c->cfg.periodic[d] ? c->anchor_bits[d] : 0;
        }
        if (obs->anchor_h == 0.0) obs->anchor_h = c->ag.hd[0];
        obs->skin = c->cfg.skin;
        return (int)ORBC_OK;
    });
}

int orbc_get_pairs(orbc_ctx *c, long long cap, unsigned *id_i, unsigned *id_j, long long *count) {
    return guarded(c, [&] {
        if (!count || cap < 0 || (cap > 0 && (!id_i || !id_j))) return fail(c, ORBC_ERR_INVALID, "bad arguments");
        ensure_list(c);
        unsigned long long *cnt_d = nullptr;
        unsigned *di = nullptr, *dj = nullptr;
        cnt_d = c->mem.alloc<unsigned long long>(1);
        di = c->mem.alloc<unsigned>(cap > 0 ? cap : 1);
        dj = c->mem.alloc<unsigned>(cap > 0 ? cap : 1);
        ORBC_HIP_CHECK(hipMemsetAsync(cnt_d, 0, 8, c->stream));
        launch_pairs_export(force_args(c), c->mixed(), di, dj, cnt_d, cap, c->stream);
        unsigned long long tot = 0;
        ORBC_HIP_CHECK(hipMemcpyAsync(&tot, cnt_d, 8, hipMemcpyDeviceToHost, c->stream));
        sync(c);
        const long long m = std::min<long long>((long long)tot, cap);
        if (m > 0) {
            ORBC_HIP_CHECK(hipMemcpyAsync(id_i, di, 4 * m, hipMemcpyDeviceToHost, c->stream));
            ORBC_HIP_CHECK(hipMemcpyAsync(id_j, dj, 4 * m, hipMemcpyDeviceToHost, c->stream));
        }
        sync(c);
        c->mem.release(cnt_d); c->mem.release(di); c->mem.release(dj);
        *count = (long long)tot;
        return (int)ORBC_OK;
    });
}

int orbc_set_timing(orbc_ctx *c, int on) {
    return guarded(c, [&] {
        c->timing = on != 0;
        c->t_step_ms = c->t_rebuild_ms = 0;
        c->n_step_launches = c->n_timed_rebuilds = 0;
        return (int)ORBC_OK;
    });
}

int orbc_get_timing(orbc_ctx *c, double *step_kernel_ms, long long *step_launches, double *rebuild_ms,
                    long long *rebuilds) {
    if (!c) return ORBC_ERR_INVALID;
    if (step_kernel_ms) *step_kernel_ms = c->t_step_ms;
    if (step_launches) *step_launches = c->n_step_launches;
    if (rebuild_ms) *rebuild_ms = c->t_rebuild_ms;
    if (rebuilds) *rebuilds = c->n_timed_rebuilds;
    return ORBC_OK;
}

int orbc_synchronize(orbc_ctx *c) {
    return guarded(c, [&] {
        sync(c);
        return (int)ORBC_OK;
    });
}

}  // extern "C"


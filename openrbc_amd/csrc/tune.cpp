// tune.cpp -- the run-time autotuner of orbc_run (no reference counterpart).
//
// The knobs that set how fast the main loop runs on a given system -- the
// Verlet skin (list length against rebuild frequency), block halos of the
// fused step (LDS staging of each particle block's neighbours, or 32-bit
// gathers), the edge of the centroid grid behind the nearest-centroid
// partition, and the number of steps launched per host synchronisation --
// are measured on the uploaded system itself: the state is saved, every
// trial restores it, changes one knob and times orbc_run over a fixed number
// of steps (wall clock, rebuilds and Voronoi work included), knobs are taken
// one after another (coordinate descent), and the state is restored at the
// end with the best configuration applied.  The trials can append their table
// to a JSON-lines file.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.h"

using namespace orbc;

#define guarded orbc_guarded

namespace {

struct Snapshot {
    std::vector<double> x[2], v[2], n[2], o[2], f[2], t[2];
    std::vector<int> type[2];
    std::vector<unsigned> id[2];
    std::vector<float> cen;
    std::vector<int> aff[2];
    long long step = 0, since_build = 0;
    double list_life = 0, zeta = 0;
    bool nh_started = false, have_ft = false;
};

void check(orbc_ctx *c, int rc) {
    if (rc != ORBC_OK) throw std::runtime_error("autotune: " + c->err);
}

Snapshot save(orbc_ctx *c) {
    Snapshot s;
    for (int sp = 0; sp < 2; ++sp) {
        const long long m = c->sp[sp].n;
        for (auto *a : {&s.x[sp], &s.v[sp], &s.n[sp], &s.o[sp]}) a->assign(3 * m, 0.0);
        check(c, orbc_get_state(c, sp, s.x[sp].data(), s.v[sp].data(), s.n[sp].data(), s.o[sp].data()));
        s.type[sp] = c->sp[sp].type;
        s.id[sp].assign(c->sp[sp].id.begin(), c->sp[sp].id.end());
        if (c->cfg.thermostat == ORBC_RUN_NOSE_HOOVER && c->nh_started) {
            s.f[sp].assign(3 * m, 0.0);
            s.t[sp].assign(3 * m, 0.0);
            check(c, orbc_get_forces(c, sp, s.f[sp].data(), s.t[sp].data(), nullptr, nullptr));
            s.have_ft = true;
        }
    }
    if (c->voro.ncell > 0) {
        s.cen.assign(3 * c->voro.ncell, 0.f);
        s.aff[0].assign(c->sp[0].n, 0);
        s.aff[1].assign(c->sp[1].n, 0);
        check(c, orbc_voronoi_get(c, s.cen.data(), s.aff[0].data(), s.aff[1].data()));
    }
    ORBC_HIP_CHECK(hipMemcpyAsync(&s.zeta, c->scal, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    sync(c);
    s.step = c->step; s.since_build = c->since_build; s.list_life = c->list_life; s.nh_started = c->nh_started;
    return s;
}

void restore(orbc_ctx *c, const Snapshot &s) {
    for (int sp = 0; sp < 2; ++sp) {
        const int *ty = sp == ORBC_PROTEIN ? s.type[sp].data() : nullptr;
        check(c, orbc_set_state(c, sp, c->sp[sp].n, s.x[sp].data(), s.v[sp].data(), s.n[sp].data(), s.o[sp].data(),
                                ty, s.id[sp].data()));
    }
    ensure_live(c);
    if (!s.cen.empty()) {
        check(c, orbc_voronoi_set_centroids(c, (long long)s.cen.size() / 3, s.cen.data()));
        voronoi_set_affiliation(c, s.aff[0].data(), s.aff[1].data());
    }
    if (s.have_ft)
        for (int sp = 0; sp < 2; ++sp) check(c, orbc_set_forces(c, sp, s.f[sp].data(), s.t[sp].data()));
    ORBC_HIP_CHECK(hipMemcpyAsync(c->scal, &s.zeta, sizeof(double), hipMemcpyHostToDevice, c->stream));
    sync(c);
    c->step = s.step; c->since_build = s.since_build; c->list_life = s.list_life; c->nh_started = s.nh_started;
    c->need_rebuild = true;
}

struct Knobs {
    double skin;
    int block_halo;
    double cg_edge;
    int batch;
    int halo_cap;
};

Knobs current(const orbc_ctx *c) {
    return Knobs{c->cfg.skin, c->cfg.block_halo, c->cg_edge, c->batch, c->halo_cap};
}

void apply(orbc_ctx *c, const Knobs &k) {
    c->cfg.skin = k.skin;
    c->cfg.block_halo = k.block_halo;
    c->cg_edge = k.cg_edge;
    c->batch = k.batch;
    c->halo_cap = k.halo_cap;
    c->need_rebuild = true;
}

}  // namespace

extern "C" {

int orbc_autotune(orbc_ctx *c, int warm_steps, int trial_steps, const char *table_path, orbc_tuning *out) {
    return guarded(c, [&] {
        if (trial_steps < 4) return fail(c, ORBC_ERR_INVALID, "trial_steps must be at least 4");
        if (warm_steps < 0) warm_steps = trial_steps / 4;
        ensure_live(c);
        const bool cells = c->voro.ncell > 0 && ((c->cfg.volume_target != 0.0 && c->cfg.volume_strength != 0.0) ||
                                                c->cfg.freq_cleanup > 0);
        const Snapshot snap = save(c);
        const Knobs start = current(c);
        c->tuning = true;               // no stray-lipid cleanup inside the trials (the count must stay)
        struct Trial { Knobs k; double ms; };
        std::vector<Trial> trials;
        auto measure = [&](const Knobs &k) {
            restore(c, snap);
            apply(c, k);
            check(c, orbc_run(c, warm_steps));
            const auto t0 = std::chrono::steady_clock::now();
            check(c, orbc_run(c, trial_steps));
            const auto t1 = std::chrono::steady_clock::now();
            const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / trial_steps;
            trials.push_back({k, ms});
            return ms;
        };
        Knobs best = start;
        double best_ms = 0;
        try {
            best_ms = measure(best);
            const double t_default = best_ms;
            auto sweep = [&](auto set, const std::vector<double> &values) {
                Knobs base = best;
                for (double v : values) {
                    Knobs k = base;
                    set(k, v);
                    if (std::memcmp(&k, &base, sizeof(Knobs)) == 0) continue;
                    const double ms = measure(k);
                    if (ms < best_ms) { best_ms = ms; best = k; }
                }
            };
            sweep([](Knobs &k, double v) { k.skin = v; }, {0.4, 0.5, 0.6, 0.7, 0.8});
            sweep([](Knobs &k, double v) { k.block_halo = (int)v; }, {0, 1, 2});
            // the tile shape of the halo kernels: 896 local indices at 5 blocks
            // per CU, or 1024 at 4 (fewer blocks overflow into global gathers)
            if (best.block_halo != 0)
                sweep([](Knobs &k, double v) { k.halo_cap = (int)v; }, {(double)HALO_CAP, (double)HALO_CAP_WIDE});
            if (cells) sweep([](Knobs &k, double v) { k.cg_edge = v; }, {6.0, 8.0, 10.0});
            sweep([](Knobs &k, double v) { k.batch = (int)v; }, {8, 16, 32});
            c->tuning = false;
            restore(c, snap);
            apply(c, best);
            if (out) {
                out->skin = best.skin; out->block_halo = best.block_halo; out->centroid_cell = best.cg_edge;
                out->batch = best.batch; out->ms_per_step = best_ms; out->ms_per_step_start = t_default;
                out->trials = (int)trials.size();
                out->halo_capacity = best.halo_cap;
            }
        } catch (...) {
            c->tuning = false;
            throw;
        }
        if (table_path && *table_path) {
            std::FILE *fp = std::fopen(table_path, "a");
            if (!fp) return fail(c, ORBC_ERR_INVALID, std::string("cannot open ") + table_path);
            hipDeviceProp_t prop{};
            (void)hipGetDeviceProperties(&prop, c->cfg.device);
            std::fprintf(fp, "{\"device\": \"%s %s\", \"n_lipid\": %lld, \"n_protein\": %lld, \"n_bonds\": %lld, "
                             "\"precision\": \"%s\", \"thermostat\": \"%s\", \"cells\": %d, \"volume\": %d, "
                             "\"warm_steps\": %d, \"trial_steps\": %d, \"trials\": [",
                         prop.name, prop.gcnArchName, c->nl, c->np, c->nbonds, c->mixed() ? "mixed" : "fp64",
                         c->cfg.thermostat == ORBC_RUN_NOSE_HOOVER ? "nose-hoover" : "langevin", (int)cells,
                         (int)(c->cfg.volume_target != 0.0 && c->cfg.volume_strength != 0.0), warm_steps,
                         trial_steps);
            for (size_t k = 0; k < trials.size(); ++k)
                std::fprintf(fp, "%s{\"skin\": %g, \"block_halo\": %d, \"centroid_cell\": %g, \"batch\": %d, "
                                 "\"halo_capacity\": %d, \"ms_per_step\": %.5f}",
                             k ? ", " : "", trials[k].k.skin, trials[k].k.block_halo, trials[k].k.cg_edge,
                             trials[k].k.batch, trials[k].k.halo_cap, trials[k].ms);
            std::fprintf(fp, "], \"best\": {\"skin\": %g, \"block_halo\": %d, \"centroid_cell\": %g, \"batch\": %d, "
                             "\"halo_capacity\": %d, \"ms_per_step\": %.5f}}\n",
                         best.skin, best.block_halo, best.cg_edge, best.batch, best.halo_cap, best_ms);
            std::fclose(fp);
        }
        return (int)ORBC_OK;
    });
}

int orbc_set_tuning(orbc_ctx *c, const orbc_tuning *t) {
    return guarded(c, [&] {
        if (!t || !(t->skin > 0) || t->block_halo < 0 || t->block_halo > 2 || !(t->centroid_cell > 0) ||
            t->batch < 1 || t->batch > 32 ||
            (t->halo_capacity != 0 && t->halo_capacity != HALO_CAP && t->halo_capacity != HALO_CAP_WIDE))
            return fail(c, ORBC_ERR_INVALID, "bad tuning");
        apply(c, Knobs{t->skin, t->block_halo, t->centroid_cell, t->batch,
                       t->halo_capacity ? t->halo_capacity : c->halo_cap});
        return (int)ORBC_OK;
    });
}

int orbc_get_tuning(const orbc_ctx *c, orbc_tuning *t) {
    if (!c || !t) return ORBC_ERR_INVALID;
    std::memset(t, 0, sizeof(*t));
    t->skin = c->cfg.skin; t->block_halo = c->cfg.block_halo; t->centroid_cell = c->cg_edge; t->batch = c->batch;
    t->halo_capacity = c->halo_cap;
    return ORBC_OK;
}

}  // extern "C"

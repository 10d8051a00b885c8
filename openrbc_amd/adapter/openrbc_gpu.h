// openrbc_gpu.h -- the reference-side binding: OpenRBC's own driver
// (src/openrbc.cpp) runs its main loop on liborbc_hip through the C ABI of
// include/orbc.h.  Header-only C++11, to be included next to the reference's
// headers (it uses LipidContainer / ProteContainer, src/container.h:65-157,
// and RTParameter, src/runtime_parameter.h:38-80); link with -lorbc_hip.
//
//   openrbc::GpuLoop gpu(lipid, protein, param);     // upload once
//   gpu.run(param.freq_dump);                        // src/openrbc.cpp:189-256 between dumps
//   gpu.download(lipid, protein);                    // containers in upload order again
//   save_frame(traj, lipid, protein, cell_lipid, cell_protein, param);
//
// Every status is checked: a failing call throws std::runtime_error carrying
// orbc_last_error, instead of the reference's rt_assert / exit.
#ifndef OPENRBC_GPU_H_
#define OPENRBC_GPU_H_

#include <stdexcept>
#include <string>
#include <vector>

#include "orbc.h"

namespace openrbc {

class GpuLoop {
public:
    template <class LC, class PC, class P>
    GpuLoop(LC &lipid, PC &protein, P const &param, int precision = ORBC_MIXED) {
        orbc_config cfg;
        orbc_config_default(&cfg);
        cfg.precision = precision;
        cfg.kBT = param.kBT;                 // src/runtime_parameter.h:49-55
        cfg.eta = param.eta;
        cfg.dt = param.dt;
        cfg.rho = param.rho;
        cfg.voronoi_cell_size = param.voronoi_cell_size;
        cfg.stray_tolerance = param.stray_tolerance;
        cfg.freq_voronoi = param.freq_voronoi;
        cfg.freq_sort_ctrd = param.freq_sort_ctrd;
        cfg.freq_sort_bond = param.freq_sort_bond;
        cfg.freq_cleanup = param.freq_cleanup;
        cfg.seed = (unsigned long long)(unsigned)param.rseed;
        for (int d = 0; d < 3; ++d) {
            cfg.box_lo[d] = param.box[d][0];
            cfg.box_hi[d] = param.box[d][1];
        }
        const int rc = orbc_create(&cfg, &ctx_);
        if (rc != ORBC_OK) throw std::runtime_error("orbc_create failed with status " + std::to_string(rc));
        try {
            upload(lipid, protein);
        } catch (...) {
            orbc_destroy(ctx_);
            ctx_ = nullptr;
            throw;
        }
    }
    ~GpuLoop() {
        if (ctx_) orbc_destroy(ctx_);
    }
    GpuLoop(const GpuLoop &) = delete;
    GpuLoop &operator=(const GpuLoop &) = delete;

    orbc_ctx *context() const { return ctx_; }

    // both containers and the bonds by tag (src/container.h:30-34,117-157)
    template <class LC, class PC>
    void upload(LC &lipid, PC &protein) {
        std::vector<double> x, v, n, o;
        std::vector<unsigned> tag(lipid.size());
        for (std::size_t i = 0; i < lipid.size(); ++i) tag[i] = (unsigned)lipid.tag[(int)i];
        pack(lipid, x, v, n, o);
        check(orbc_set_state(ctx_, ORBC_LIPID, (long long)lipid.size(), x.data(), v.data(), n.data(), o.data(),
                             nullptr, tag.data()));
        pack(protein, x, v, n, o);
        std::vector<int> type(protein.size());
        tag.resize(protein.size());
        for (std::size_t i = 0; i < protein.size(); ++i) {
            type[i] = protein.type[i];
            tag[i] = (unsigned)protein.tag[i];
        }
        check(orbc_set_state(ctx_, ORBC_PROTEIN, (long long)protein.size(), x.data(), v.data(), n.data(), o.data(),
                             type.data(), tag.data()));
        std::vector<int> bt;
        std::vector<unsigned> bi, bj;
        for (std::size_t k = 0; k < protein.bonds.size(); ++k) {
            bt.push_back(protein.bonds[k].type);
            bi.push_back((unsigned)protein.bonds[k].i);
            bj.push_back((unsigned)protein.bonds[k].j);
        }
        check(orbc_set_bonds(ctx_, (long long)bt.size(), bt.data(), bi.data(), bj.data()));
    }

    // the reference's while loop for n steps (rebuilds on the skin criterion)
    void run(long long nsteps) { check(orbc_run(ctx_, nsteps)); }

    // the reference's per-call sequence, for a loop that keeps its own order
    void compute_pairwise_fused() { check(orbc_compute_pairwise(ctx_, 0u)); }
    void compute_bonded() { check(orbc_compute_bonded(ctx_, 0u)); }
    void verlet_langevin() { check(orbc_integrate(ctx_, ORBC_VERLET_LANGEVIN)); }

    // x, v, n, o back into the containers (each species in upload order; the
    // lipid count shrinks when the stray-lipid cleanup removed some)
    template <class LC, class PC>
    void download(LC &lipid, PC &protein) {
        orbc_observables obs;
        check(orbc_get_observables(ctx_, &obs));
        if ((std::size_t)obs.n_lipid != lipid.size()) lipid.resize((std::size_t)obs.n_lipid);
        unpack(ORBC_LIPID, lipid);
        unpack(ORBC_PROTEIN, protein);
    }

private:
    orbc_ctx *ctx_ = nullptr;

    void check(int rc) const {
        if (rc != ORBC_OK) throw std::runtime_error(std::string("liborbc_hip: ") + orbc_last_error(ctx_));
    }

    template <class C>
    static void pack(C const &c, std::vector<double> &x, std::vector<double> &v, std::vector<double> &n,
                     std::vector<double> &o) {
        const std::size_t m = c.size();
        x.resize(3 * m); v.resize(3 * m); n.resize(3 * m); o.resize(3 * m);
        for (std::size_t i = 0; i < m; ++i)
            for (int d = 0; d < 3; ++d) {
                x[3 * i + d] = c.x[i][d]; v[3 * i + d] = c.v[i][d];
                n[3 * i + d] = c.n[i][d]; o[3 * i + d] = c.o[i][d];
            }
    }

    template <class C>
    void unpack(int species, C &c) {
        std::vector<double> x(3 * c.size()), v(x.size()), n(x.size()), o(x.size());
        if (c.size()) check(orbc_get_state(ctx_, species, x.data(), v.data(), n.data(), o.data()));
        for (std::size_t i = 0; i < c.size(); ++i)
            for (int d = 0; d < 3; ++d) {
                c.x[i][d] = x[3 * i + d]; c.v[i][d] = v[3 * i + d];
                c.n[i][d] = n[3 * i + d]; c.o[i][d] = o[3 * i + d];
            }
    }
};

}  // namespace openrbc

#endif

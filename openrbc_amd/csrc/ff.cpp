// ff.cpp -- force-field tables (host), src/forcefield_canonical.h:24-156.
//
// The reference keeps separate lipid-lipid, lipid-protein (indexed by the
// protein type) and protein-protein (indexed type1 + 6*type2) tables and a
// 6x6 LJ table whose lipid-protein entries are read from row 0
// (src/compute_pairwise_fused.h:264-276).  The kernels index one symmetric
// 6x6 table by (owner type, neighbour type); this file folds the reference's
// tables into it.  FP32 values are the FP64 values rounded once.
#include <cmath>
#include <cstring>

#include "orbc_internal.h"

namespace orbc {

namespace {

struct Raw {
    double cut_poly[36], req[36], eps[36], alpha[36];
    int aniso[36];
    double lj_cut[36], lj_eps[36], lj_sigma[36];
};

Raw make_raw() {
    Raw r;
    std::memset(&r, 0, sizeof(r));
    // lipid-protein (src/forcefield_canonical.h:51-54); index 0 is lipid-lipid (:46-49)
    const double cutlp[NT] = {kCutLL, 2.6, 2.6, 2.6, 0, 0};
    const double reqlp[NT] = {kReqLL, 1.685, 1.685, 1.1225, 0, 0};
    const double epslp[NT] = {kEpsLL, 1.4, 2.8, 2.8, 0, 0};
    const double alplp[NT] = {kAlphaLL, 5, 5, 5, 0, 0};
    // protein-protein among band-3 (1,2) and glycophorin (3) (:56-73)
    auto pp_req = [](int a, int b) {
        if (a == 3 && b == 3) return 1.1225;
        if (a == 3 || b == 3) return 1.685;
        return 2.245;
    };
    // WCA table (:76-99): epsilon 1 where a cut is set
    const double c34 = 3.4 * 1.1225;
    const double lj_cut[36] = {0,      0,   0,   0, 1.1225, 1.1225, 0,      0,   0,      0, c34, c34,
                               0,      0,   0,   0, c34,    1.1225, 0,      0,   0,      0, 0,   0,
                               1.1225, c34, c34, 0, 0,      0,      1.1225, c34, 1.1225, 0, 0,   0};
    const double lj_sig[36] = {0, 0,   0,   0, 1,   1,   0, 0,   0,   0, 3.4, 3.4, 0, 0,   0, 0, 3.4, 1,
                               0, 0,   0,   0, 0,   0,   1, 3.4, 3.4, 0, 3,   1.8, 1, 3.4, 1, 0, 1.8, 1};
    const double lj_eps[36] = {0, 0, 0, 0, 1, 1, 0, 0, 0, 0, 1, 1, 0, 0, 0, 0, 1, 1,
                               0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0, 1, 1, 1, 0, 0, 0};
    for (int a = 0; a < NT; ++a)
        for (int b = 0; b < NT; ++b) {
            const int k = a * NT + b;
            int li;
            if (a == 0 || b == 0) {
                const int p = a == 0 ? b : a;
                r.cut_poly[k] = cutlp[p]; r.req[k] = reqlp[p]; r.eps[k] = epslp[p]; r.alpha[k] = alplp[p];
                r.aniso[k] = 1;
                li = p;
            } else {
                const bool band_gly = a <= 3 && b <= 3;
                r.cut_poly[k] = band_gly ? 2.6 : 0.0;
                r.req[k] = band_gly ? pp_req(a, b) : 0.0;
                r.eps[k] = band_gly ? 1.0 : 0.0;
                r.alpha[k] = 0.0;
                r.aniso[k] = 0;
                li = a + NT * b;
            }
            r.lj_cut[k] = lj_cut[li];
            r.lj_eps[k] = lj_eps[li];
            r.lj_sigma[k] = lj_sig[li];
        }
    return r;
}

}  // namespace

void ff_fill(FFTab<double> &d, FFTab<float> &f) {
    const Raw r = make_raw();
    std::memset(&d, 0, sizeof(d));
    std::memset(&f, 0, sizeof(f));
    for (int k = 0; k < 36; ++k) {
        const double c = r.cut_poly[k];
        d.cut_poly[k] = c;
        d.cut2_poly[k] = c * c;
        d.rep[k] = c > 0 ? r.eps[k] / ipow_c(c - r.req[k], 8) : 0.0;
        d.att[k] = c > 0 ? -2.0 * r.eps[k] / ipow_c(c - r.req[k], 4) : 0.0;
        d.alpha[k] = r.alpha[k];
        d.aniso[k] = r.aniso[k];
        d.cut2_lj[k] = r.lj_cut[k] * r.lj_cut[k];
        d.lj1[k] = 48.0 * r.lj_eps[k] * std::pow(r.lj_sigma[k], 12);
        d.lj2[k] = 24.0 * r.lj_eps[k] * std::pow(r.lj_sigma[k], 6);
    }
    const double mass[NT] = {1, 4, 4, 1, 10, 10};                         // :40
    const double r0[NBT] = {2.25, 1.1225, 2.25, 2.24};                      // :108
    for (int t = 0; t < NT; ++t) { d.mass[t] = mass[t]; d.inv_mass[t] = 1.0 / mass[t]; }
    for (int b = 0; b < NBT; ++b) { d.bond_r0[b] = r0[b]; d.bond_K[b] = 57.0; }   // :109
    for (int k = 0; k < 36; ++k) {
        f.cut_poly[k] = (float)d.cut_poly[k];
        f.cut2_poly[k] = (float)d.cut2_poly[k];
        f.rep[k] = (float)d.rep[k];
        f.att[k] = (float)d.att[k];
        f.alpha[k] = (float)d.alpha[k];
        f.aniso[k] = d.aniso[k];
        f.cut2_lj[k] = (float)d.cut2_lj[k];
        f.lj1[k] = (float)d.lj1[k];
        f.lj2[k] = (float)d.lj2[k];
    }
    for (int t = 0; t < NT; ++t) { f.mass[t] = (float)d.mass[t]; f.inv_mass[t] = (float)d.inv_mass[t]; }
    for (int b = 0; b < NBT; ++b) { f.bond_r0[b] = (float)d.bond_r0[b]; f.bond_K[b] = (float)d.bond_K[b]; }
}

double ff_max_cut(int a, int b) {
    static FFTab<double> d;
    static FFTab<float> f;
    static bool ready = false;
    if (!ready) { ff_fill(d, f); ready = true; }
    const int k = a * NT + b;
    const double c1 = d.cut_poly[k], c2 = std::sqrt(d.cut2_lj[k]);
    return c1 > c2 ? c1 : c2;
}

}  // namespace orbc

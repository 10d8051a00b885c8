/*
 * orbc_oracle.c -- CPU restatement of OpenRBC's per-timestep MD loop.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load this library; the product never does.
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math, so every
 * rounding below is the one written).
 *
 * Citations are OpenRBC source paths (src/...:line).  Pinning of each
 * function is listed in orbc_oracle.h.
 */
#include "orbc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Force field: src/forcefield_canonical.h:24-28 (rep/att), :40-109 (tables),
 * :143-155 (LJ coefficients).  The reference's 6x6 tables are folded into one
 * table per (owner type a, neighbour type b):
 *   lipid-lipid      -> cutll/reqll/epsilonll/alphall (anisotropic)
 *   lipid-protein p  -> cutlp[p]... (anisotropic), LJ row 0 column p
 *   protein-protein  -> cutpp[a+6b]... (isotropic repulsive), LJ[a+6b]        */
static orc_ff FF;
static int ff_ready = 0;

static double ipow(double b, int e) {
    double r = 1.0;
    for (int i = 0; i < e; ++i) r *= b;
    return r;
}

static void ff_init(void) {
    static const double mass[6] = {1, 4, 4, 1, 10, 10};
    static const double radius[6] = {0.56125, 1.12375, 1.12375, 0.56125, 1.5, 0.5};
    const double cutll = 2.6, reqll = 1.1225, epsll = 1.2, alphall = 1.55;
    const double cutlp[6] = {cutll, 2.6, 2.6, 2.6, 0, 0};
    const double reqlp[6] = {reqll, 1.685, 1.685, 1.1225, 0, 0};
    const double epslp[6] = {epsll, 1.4, 2.8, 2.8, 0, 0};
    const double alphalp[6] = {alphall, 5, 5, 5, 0, 0};
    const double cutpp[36] = {
        cutlp[0], cutlp[1], cutlp[2], cutlp[3], 0, 0,
        cutlp[1], 2.6, 2.6, 2.6, 0, 0,
        cutlp[2], 2.6, 2.6, 2.6, 0, 0,
        cutlp[3], 2.6, 2.6, 2.6, 0, 0,
        0, 0, 0, 0, 0, 0,
        0, 0, 0, 0, 0, 0};
    const double reqpp[36] = {
        reqlp[0], reqlp[1], reqlp[2], reqlp[3], 0, 0,
        reqlp[1], 2.245, 2.245, 1.685, 0, 0,
        reqlp[2], 2.245, 2.245, 1.685, 0, 0,
        reqlp[3], 1.685, 1.685, 1.1225, 0, 0,
        0, 0, 0, 0, 0, 0,
        0, 0, 0, 0, 0, 0};
    const double epspp[36] = {
        epslp[0], epslp[1], epslp[2], epslp[3], 0, 0,
        epslp[1], 1, 1, 1, 0, 0,
        epslp[2], 1, 1, 1, 0, 0,
        epslp[3], 1, 1, 1, 0, 0,
        0, 0, 0, 0, 0, 0,
        0, 0, 0, 0, 0, 0};
    const double lj_eps[36] = {
        0, 0, 0, 0, 1, 1,
        0, 0, 0, 0, 1, 1,
        0, 0, 0, 0, 1, 1,
        0, 0, 0, 0, 0, 0,
        1, 1, 1, 0, 0, 0,
        1, 1, 1, 0, 0, 0};
    const double lj_sigma[36] = {
        0, 0, 0, 0, 1, 1,
        0, 0, 0, 0, 3.4, 3.4,
        0, 0, 0, 0, 3.4, 1,
        0, 0, 0, 0, 0, 0,
        1, 3.4, 3.4, 0, 3, 1.8,
        1, 3.4, 1, 0, 1.8, 1};
    const double c34 = 3.4 * 1.1225;
    const double lj_cut[36] = {
        0, 0, 0, 0, 1.1225, 1.1225,
        0, 0, 0, 0, c34, c34,
        0, 0, 0, 0, c34, 1.1225,
        0, 0, 0, 0, 0, 0,
        1.1225, c34, c34, 0, 0, 0,
        1.1225, c34, 1.1225, 0, 0, 0};
    memset(&FF, 0, sizeof(FF));
    for (int a = 0; a < 6; ++a) {
        FF.mass[a] = mass[a];
        FF.radius[a] = radius[a];
        for (int b = 0; b < 6; ++b) {
            const int k = a * 6 + b;
            double cut, req, eps, alpha;
            int li;
            if (a == 0 || b == 0) {
                const int p = (a == 0) ? b : a;
                cut = cutlp[p]; req = reqlp[p]; eps = epslp[p]; alpha = alphalp[p];
                FF.aniso[k] = 1;
                li = p;                       /* lj_cutsq[type]: row 0 */
            } else {
                const int idx = a + 6 * b;    /* type12 = type1 + type2 * n_type */
                cut = cutpp[idx]; req = reqpp[idx]; eps = epspp[idx]; alpha = 0;
                FF.aniso[k] = 0;
                li = idx;
            }
            FF.cut_poly[k] = cut;
            FF.cut2_poly[k] = cut * cut;
            FF.rep[k] = (cut > 0) ? eps / ipow(cut - req, 8) : 0.0;
            FF.att[k] = (cut > 0) ? -2.0 * eps / ipow(cut - req, 4) : 0.0;
            FF.alpha[k] = alpha;
            FF.cut2_lj[k] = lj_cut[li] * lj_cut[li];
            FF.lj1[k] = 48.0 * lj_eps[li] * pow(lj_sigma[li], 12);
            FF.lj2[k] = 24.0 * lj_eps[li] * pow(lj_sigma[li], 6);
        }
    }
    const double r0[4] = {2.25, 1.1225, 2.25, 2.24};
    for (int b = 0; b < 4; ++b) { FF.bond_r0[b] = r0[b]; FF.bond_K[b] = 57.0; }
    ff_ready = 1;
}

const orc_ff *orc_ff_get(void) {
    if (!ff_ready) ff_init();
    return &FF;
}

static inline double dot3(const double *a, const double *b) {
    return fma(a[2], b[2], fma(a[1], b[1], a[0] * b[0]));
}

/* ------------------------------------------------------------------------ */
int orc_pair(int ti, int tj, const double dx[3], const double ni[3], const double nj[3],
             double fi[3], double tq[3], double *u, double *w) {
    const orc_ff *ff = orc_ff_get();
    const double r2 = dot3(dx, dx);
    if (!(r2 > 1e-5)) return 0;
    const int k = ti * 6 + tj;
    double f[3];
    double uu;
    if (r2 < ff->cut2_poly[k]) {
        const double r = sqrt(r2);
        const double rinv = 1.0 / r;
        const double e[3] = {dx[0] * rinv, dx[1] * rinv, dx[2] * rinv};
        const double s = ff->cut_poly[k] - r;
        const double s3 = s * s * s, s4 = s * s3, s7 = s3 * s4;
        if (ff->aniso[k]) {
            /* src/pairwise_kernel.h:37-62 */
            const double alpha = ff->alpha[k], rep = ff->rep[k], att = ff->att[k];
            const double ninj = dot3(ni, nj), nie = dot3(ni, e), nje = dot3(nj, e);
            const double a = ninj - nie * nje;
            const double A = 1.0 + alpha * (a - 1.0);
            double pi[3], pj[3];
            for (int d = 0; d < 3; ++d) { pi[d] = ni[d] - nie * e[d]; pj[d] = nj[d] - nje * e[d]; }
            const double ua = att * s4;
            const double aua = alpha * ua;
            const double auar = aua / r;
            const double fra = 8.0 * rep * s7 + A * 4.0 * att * s3;
            for (int d = 0; d < 3; ++d) {
                f[d] = fra * e[d] + auar * (nje * pi[d] + nie * pj[d]);
                tq[d] -= aua * pj[d];
            }
            uu = rep * s4 * s4 + A * att * s4;
        } else {
            /* src/pairwise_kernel_fused.h:79-97 */
            const double fr = 8.0 * ff->rep[k] * s7;
            for (int d = 0; d < 3; ++d) f[d] = fr * e[d];
            uu = ff->rep[k] * s4 * s4;
        }
    } else if (r2 < ff->cut2_lj[k]) {
        /* src/pairwise_kernel_fused.h:63-77 (WCA, unshifted) */
        const double r2inv = 1.0 / r2;
        const double r6inv = r2inv * r2inv * r2inv;
        const double fpair = r6inv * (ff->lj1[k] * r6inv - ff->lj2[k]) * r2inv;
        for (int d = 0; d < 3; ++d) f[d] = dx[d] * fpair;
        uu = r6inv * (ff->lj1[k] / 12.0 * r6inv - ff->lj2[k] / 6.0);
    } else {
        return 0;
    }
    for (int d = 0; d < 3; ++d) fi[d] += f[d];
    if (u) *u = uu;
    if (w) *w = dot3(dx, f);
    return 1;
}

/* ------------------------------------------------------------------------ */
static struct { double lo[3], L[3]; int per[3]; } BOX;

void orc_set_box(const double lo[3], const double L[3], const int per[3]) {
    for (int d = 0; d < 3; ++d) {
        BOX.lo[d] = lo ? lo[d] : 0.0;
        BOX.L[d] = L ? L[d] : 0.0;
        BOX.per[d] = per ? (per[d] != 0) : 0;
    }
}

static inline void min_image(double dx[3]) {
    for (int d = 0; d < 3; ++d)
        if (BOX.per[d]) dx[d] -= BOX.L[d] * rint(dx[d] / BOX.L[d]);
}

void orc_wrap(int64_t n, double *x) {
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d)
            if (BOX.per[d]) {
                const double u = x[3 * i + d] - BOX.lo[d];
                x[3 * i + d] = BOX.lo[d] + (u - BOX.L[d] * floor(u / BOX.L[d]));
            }
}

static inline void pair_dx(const double *x, int64_t a, int64_t b, double dx[3]) {
    for (int d = 0; d < 3; ++d) dx[d] = x[3 * a + d] - x[3 * b + d];
    min_image(dx);
}

static inline int interacts_fp64(int ti, int tj, double r2) {
    const orc_ff *ff = orc_ff_get();
    const int k = ti * 6 + tj;
    return (r2 > 1e-5) && (r2 < ff->cut2_poly[k] || r2 < ff->cut2_lj[k]);
}

/* ------------------------------------------------------------------------ */
int64_t orc_pairs_brute_fp64(int64_t n, const double *x, const int32_t *type,
                             int64_t cap, int64_t *out) {
    int64_t cnt = 0;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = i + 1; j < n; ++j) {
            double dx[3];
            pair_dx(x, i, j, dx);
            if (interacts_fp64(type[i], type[j], dot3(dx, dx))) {
                if (cnt < cap) { out[2 * cnt] = i; out[2 * cnt + 1] = j; }
                ++cnt;
            }
        }
    return cnt;
}

static int cmp_pair(const void *a, const void *b) {
    const int64_t *p = (const int64_t *)a, *q = (const int64_t *)b;
    if (p[0] != q[0]) return p[0] < q[0] ? -1 : 1;
    if (p[1] != q[1]) return p[1] < q[1] ? -1 : 1;
    return 0;
}

int64_t orc_pairs_cell_fp64(int64_t n, const double *x, const int32_t *type,
                            int64_t cap, int64_t *out) {
    if (n <= 0) return 0;
    const orc_ff *ff = orc_ff_get();
    double cmax2 = 0;
    for (int k = 0; k < 36; ++k) {
        if (ff->cut2_poly[k] > cmax2) cmax2 = ff->cut2_poly[k];
        if (ff->cut2_lj[k] > cmax2) cmax2 = ff->cut2_lj[k];
    }
    const double cmax = sqrt(cmax2) * 1.0001;
    double lo[3], hi[3], cs[3];
    for (int d = 0; d < 3; ++d) { lo[d] = x[d]; hi[d] = x[d]; }
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            if (x[3 * i + d] < lo[d]) lo[d] = x[3 * i + d];
            if (x[3 * i + d] > hi[d]) hi[d] = x[3 * i + d];
        }
    int64_t dim[3];
    for (int d = 0; d < 3; ++d) {
        if (BOX.per[d]) {
            /* periodic: cells tile [lo, lo+L) exactly (at least 3, else brute force) */
            lo[d] = BOX.lo[d];
            dim[d] = (int64_t)floor(BOX.L[d] / cmax);
            if (dim[d] < 3) return orc_pairs_brute_fp64(n, x, type, cap, out);
            cs[d] = BOX.L[d] / (double)dim[d];
        } else {
            cs[d] = cmax;
            dim[d] = (int64_t)((hi[d] - lo[d]) / cs[d]) + 1;
        }
    }
    const int64_t ncell = dim[0] * dim[1] * dim[2];
    int64_t *cell = malloc(sizeof(int64_t) * n), *start = calloc(ncell + 1, sizeof(int64_t));
    int64_t *order = malloc(sizeof(int64_t) * n), *fill = calloc(ncell, sizeof(int64_t));
    int64_t (*cc)[3] = malloc(sizeof(*cc) * n);
    for (int64_t i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) {
            double u = x[3 * i + d] - lo[d];
            if (BOX.per[d]) u -= BOX.L[d] * floor(u / BOX.L[d]);
            int64_t c = (int64_t)(u / cs[d]);
            if (c < 0) c = 0;
            if (c >= dim[d]) c = dim[d] - 1;
            cc[i][d] = c;
        }
        cell[i] = (cc[i][2] * dim[1] + cc[i][1]) * dim[0] + cc[i][0];
        start[cell[i] + 1]++;
    }
    for (int64_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
    for (int64_t i = 0; i < n; ++i) order[start[cell[i]] + fill[cell[i]]++] = i;
    int64_t cnt = 0;
    for (int64_t i = 0; i < n; ++i) {
        for (int64_t dz = -1; dz <= 1; ++dz)
            for (int64_t dy = -1; dy <= 1; ++dy)
                for (int64_t dxc = -1; dxc <= 1; ++dxc) {
                    int64_t c3[3] = {cc[i][0] + dxc, cc[i][1] + dy, cc[i][2] + dz};
                    int out_of = 0;
                    for (int d = 0; d < 3; ++d) {
                        if (BOX.per[d]) c3[d] = (c3[d] + dim[d]) % dim[d];
                        else if (c3[d] < 0 || c3[d] >= dim[d]) out_of = 1;
                    }
                    if (out_of) continue;
                    const int64_t c = (c3[2] * dim[1] + c3[1]) * dim[0] + c3[0];
                    for (int64_t q = start[c]; q < start[c + 1]; ++q) {
                        const int64_t j = order[q];
                        if (j <= i) continue;
                        double dx[3];
                        pair_dx(x, i, j, dx);
                        if (interacts_fp64(type[i], type[j], dot3(dx, dx))) {
                            if (cnt < cap) { out[2 * cnt] = i; out[2 * cnt + 1] = j; }
                            ++cnt;
                        }
                    }
                }
    }
    if (cnt <= cap) qsort(out, (size_t)cnt, 2 * sizeof(int64_t), cmp_pair);
    free(cell); free(start); free(order); free(fill); free(cc);
    return cnt;
}

int64_t orc_pairs_brute_mixed_anchor(int64_t n, const double *x, const double *xa, const int32_t *type,
                                     const double lo[3], const double h[3], const double A0[3],
                                     const int bits[3], int64_t cap, int64_t *out) {
    const orc_ff *ff = orc_ff_get();
    float cut2p[36], cut2l[36];
    for (int k = 0; k < 36; ++k) { cut2p[k] = (float)ff->cut2_poly[k]; cut2l[k] = (float)ff->cut2_lj[k]; }
    int64_t *A = malloc(sizeof(int64_t) * 3 * n);
    float *r = malloc(sizeof(float) * 3 * n);
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            int64_t a = (int64_t)floor((xa[3 * i + d] - lo[d]) / h[d]) - (int64_t)A0[d];
            if (bits[d] > 0) {
                const int64_t M = (int64_t)1 << bits[d];
                if (a < 0) a = 0;
                if (a > M - 1) a = M - 1;
            }
            A[3 * i + d] = a;
            r[3 * i + d] = (float)(x[3 * i + d] - (lo[d] + h[d] * (double)(a + (int64_t)A0[d])));
        }
    int64_t cnt = 0;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = i + 1; j < n; ++j) {
            float dd[3];
            for (int k = 0; k < 3; ++k) {
                int32_t dA = (int32_t)(A[3 * i + k] - A[3 * j + k]);
                if (bits[k] > 0) dA = (int32_t)((uint32_t)dA << (32 - bits[k])) >> (32 - bits[k]);
                const float a = r[3 * i + k] - r[3 * j + k];
                dd[k] = fmaf((float)h[k], (float)dA, a);
            }
            const float r2 = fmaf(dd[2], dd[2], fmaf(dd[1], dd[1], dd[0] * dd[0]));
            const int k = type[i] * 6 + type[j];
            if (r2 > 1e-5f && (r2 < cut2p[k] || r2 < cut2l[k])) {
                if (cnt < cap) { out[2 * cnt] = i; out[2 * cnt + 1] = j; }
                ++cnt;
            }
        }
    free(A); free(r);
    return cnt;
}

int64_t orc_pairs_brute_mixed(int64_t n, const double *x, const double *xa, const int32_t *type,
                              double h, int64_t cap, int64_t *out) {
    const orc_ff *ff = orc_ff_get();
    float cut2p[36], cut2l[36];
    for (int k = 0; k < 36; ++k) { cut2p[k] = (float)ff->cut2_poly[k]; cut2l[k] = (float)ff->cut2_lj[k]; }
    int64_t *A = malloc(sizeof(int64_t) * 3 * n);
    float *r = malloc(sizeof(float) * 3 * n);
    for (int64_t i = 0; i < 3 * n; ++i) {
        A[i] = (int64_t)floor(xa[i] / h);
        r[i] = (float)(x[i] - h * (double)A[i]);
    }
    int64_t cnt = 0;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = i + 1; j < n; ++j) {
            float d[3];
            for (int k = 0; k < 3; ++k) {
                const float a = r[3 * i + k] - r[3 * j + k];
                const float s = (float)(h * (double)(A[3 * i + k] - A[3 * j + k]));
                d[k] = a + s;
            }
            const float r2 = fmaf(d[2], d[2], fmaf(d[1], d[1], d[0] * d[0]));
            const int k = type[i] * 6 + type[j];
            if (r2 > 1e-5f && (r2 < cut2p[k] || r2 < cut2l[k])) {
                if (cnt < cap) { out[2 * cnt] = i; out[2 * cnt + 1] = j; }
                ++cnt;
            }
        }
    free(A); free(r);
    return cnt;
}

/* ------------------------------------------------------------------------ */
void orc_forces_pairs(int64_t n, const double *x, const double *nrm, const int32_t *type,
                      int64_t npairs, const int64_t *pairs,
                      double *f, double *t, double *e, double *w) {
    (void)n;
    for (int64_t p = 0; p < npairs; ++p) {
        const int64_t i = pairs[2 * p], j = pairs[2 * p + 1];
        /* owner-computes both sides: src/compute_pairwise_fused.h:249-316 */
        for (int side = 0; side < 2; ++side) {
            const int64_t a = side ? j : i, b = side ? i : j;
            double dx[3];
            pair_dx(x, a, b, dx);
            double u = 0, ww = 0;
            if (orc_pair(type[a], type[b], dx, nrm + 3 * a, nrm + 3 * b, f + 3 * a, t + 3 * a, &u, &ww)) {
                if (e) e[a] += 0.5 * u;
                if (w) w[a] += 0.5 * ww;
            }
        }
    }
}

void orc_bonded(int64_t nb, const int32_t *bt, const int64_t *bi, const int64_t *bj,
                const double *x, double *f, double *e, double *w) {
    const orc_ff *ff = orc_ff_get();
    for (int64_t k = 0; k < nb; ++k) {
        const int64_t i = bi[k], j = bj[k];
        double dx[3];
        pair_dx(x, j, i, dx);
        const double r = sqrt(dot3(dx, dx));
        const double K = ff->bond_K[bt[k]], r0 = ff->bond_r0[bt[k]];
        const double c = K * (1.0 - r0 / r);
        for (int d = 0; d < 3; ++d) { f[3 * i + d] += c * dx[d]; f[3 * j + d] -= c * dx[d]; }
        if (e) { const double u = 0.5 * K * (r - r0) * (r - r0); e[i] += 0.5 * u; e[j] += 0.5 * u; }
        if (w) { const double ww = -c * r * r; w[i] += 0.5 * ww; w[j] += 0.5 * ww; }
    }
}

/* ------------------------------------------------------------------------ */
static inline uint32_t mulhilo(uint32_t a, uint32_t b, uint32_t *hi) {
    const uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    return (uint32_t)p;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint32_t hi0, hi1;
        const uint32_t lo0 = mulhilo(0xD2511F53u, c0, &hi0);
        const uint32_t lo1 = mulhilo(0xCD9E8D57u, c2, &hi1);
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void orc_langevin_noise(uint64_t seed, uint32_t id, uint64_t step, double r[3]) {
    const uint32_t ctr[4] = {id, (uint32_t)step, (uint32_t)(step >> 32), 0u};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    /* uint2u11: u * 2^-31 - 1, uniform in [-1,1) (src/math_vector_integer.h:62-66) */
    for (int d = 0; d < 3; ++d) r[d] = (double)o[d] * 4.656612873077392578125e-10 - 1.0;
}

static inline void cross3(const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

static inline void normalize3(double *a) {
    const double s = 1.0 / sqrt(dot3(a, a));
    a[0] *= s; a[1] *= s; a[2] *= s;
}

void orc_langevin(int64_t n, double *x, double *v, double *nrm, double *o,
                  double *f, double *t, const int32_t *type, const uint32_t *id,
                  double kT, double eta, double dt, uint64_t seed, uint64_t step) {
    const orc_ff *ff = orc_ff_get();
    double gamma[6], sigma[6];
    for (int k = 0; k < 6; ++k) {
        gamma[k] = 6.0 * M_PI * eta * ff->radius[k];
        sigma[k] = sqrt(2.0 * kT * gamma[k]) * sqrt(3.0 / dt);
    }
    for (int64_t i = 0; i < n; ++i) {
        double *xi = x + 3 * i, *vi = v + 3 * i, *ni = nrm + 3 * i, *oi = o + 3 * i;
        double *fi = f + 3 * i, *ti = t + 3 * i;
        double tq[3], g[3], on[3];
        cross3(ni, ti, tq);                                  /* t = n x t */
        for (int d = 0; d < 3; ++d) oi[d] += dt * tq[d];     /* o += dt/I t */
        cross3(oi, ni, on);
        for (int d = 0; d < 3; ++d) g[d] = ni[d] + on[d] * dt;
        normalize3(g);
        for (int d = 0; d < 3; ++d) { ni[d] = g[d]; ti[d] = 0.0; }
        double r[3];
        orc_langevin_noise(seed, id[i], step, r);
        const int ty = type[i];
        const double dtm = dt / ff->mass[ty];
        for (int d = 0; d < 3; ++d) {
            fi[d] -= gamma[ty] * vi[d] + sigma[ty] * r[d];
            vi[d] += fi[d] * dtm;
            xi[d] += vi[d] * dt;
            fi[d] = 0.0;
        }
    }
}

double orc_nh_first(int64_t n, double *x, double *v, double *nrm, double *o,
                    double *f, double *t, const int32_t *type,
                    double dt, double zeta, const double box_lo[3], const double box_hi[3]) {
    const orc_ff *ff = orc_ff_get();
    const double g = 1.0 / (1.0 + 0.5 * dt * zeta);
    double ke = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double *xi = x + 3 * i, *vi = v + 3 * i, *ni = nrm + 3 * i, *oi = o + 3 * i;
        const double m = ff->mass[type[i]];
        for (int d = 0; d < 3; ++d) {
            vi[d] = (vi[d] + 0.5 / m * dt * f[3 * i + d]) * g;
            xi[d] += vi[d] * dt;
        }
        for (int d = 0; d < 3; ++d) {           /* src/integrate_nh.h:200-209 */
            if (xi[d] < box_lo[d]) { xi[d] = box_lo[d] + (box_lo[d] - xi[d]); vi[d] = -vi[d]; }
            else if (xi[d] > box_hi[d]) { xi[d] = box_hi[d] - (xi[d] - box_hi[d]); vi[d] = -vi[d]; }
        }
        ke += 0.5 * m * dot3(vi, vi);
        double on[3], gg[3];
        for (int d = 0; d < 3; ++d) oi[d] += 0.5 * dt * t[3 * i + d];
        cross3(oi, ni, on);
        for (int d = 0; d < 3; ++d) gg[d] = ni[d] + on[d] * dt;
        normalize3(gg);
        for (int d = 0; d < 3; ++d) { ni[d] = gg[d]; f[3 * i + d] = 0; t[3 * i + d] = 0; }
    }
    return ke;
}

double orc_nh_second(int64_t n, const double *x, double *v, const double *nrm, double *o,
                     const double *f, double *t, const int32_t *type, double dt, double zeta) {
    (void)x;
    const orc_ff *ff = orc_ff_get();
    double ke = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double *vi = v + 3 * i, *oi = o + 3 * i, *ti = t + 3 * i;
        const double m = ff->mass[type[i]];
        double tq[3];
        cross3(nrm + 3 * i, ti, tq);
        for (int d = 0; d < 3; ++d) {
            ti[d] = tq[d];
            vi[d] += 0.5 * dt * (f[3 * i + d] / m - zeta * vi[d]);
            oi[d] += 0.5 * dt * ti[d];
        }
        ke += 0.5 * m * dot3(vi, vi);
    }
    return ke;
}

double orc_nh_zeta(double zeta, double dt, double Q, double ke, int64_t n, double kT) {
    return zeta + 0.5 * dt / Q * (ke - 0.5 * 3.0 * (double)n * kT);
}

double orc_temperature(int64_t n, const double *v, const int32_t *type) {
    const orc_ff *ff = orc_ff_get();
    double ek = 0.0;
    for (int64_t i = 0; i < n; ++i) ek += ff->mass[type[i]] * dot3(v + 3 * i, v + 3 * i);
    return n > 0 ? ek / (3.0 * (double)n) : 0.0;
}

void orc_minimize_move(int64_t n, double *x, double *nrm, const double *f, const double *t,
                       const int32_t *type, double dr_opt, double dn_opt, double dt) {
    const orc_ff *ff = orc_ff_get();
    for (int64_t i = 0; i < n; ++i) {
        const double m = ff->mass[type[i]];
        double dxv[3], dn[3];
        for (int d = 0; d < 3; ++d) dxv[d] = f[3 * i + d] / m;
        cross3(t + 3 * i, nrm + 3 * i, dn);
        const double ndx = sqrt(dot3(dxv, dxv)), ndn = sqrt(dot3(dn, dn));
        double h = dt;
        if (ndx > dr_opt || ndn > dn_opt) {
            const double a = dr_opt / ndx, b = dn_opt / ndn;
            h = a < b ? a : b;
        }
        for (int d = 0; d < 3; ++d) {
            x[3 * i + d] += f[3 * i + d] * (h / m);
            nrm[3 * i + d] += dn[d] * h;
        }
        normalize3(nrm + 3 * i);
    }
}

/* ------------------------------------------------------------------------ */
static unsigned bit_space3(unsigned x) {
    x = (x | (x << 12)) & 0x00FC003Fu;
    x = (x | (x << 6)) & 0x381C0E07u;
    x = (x | (x << 4)) & 0x190C8643u;
    x = (x | (x << 2)) & 0x49249249u;
    return x;
}

uint32_t orc_morton_ref(float x, float y, float z, double bsize) {
    const unsigned i = (unsigned)(2.0f * x + bsize);
    const unsigned j = (unsigned)(2.0f * y + bsize);
    const unsigned k = (unsigned)(2.0f * z + bsize);
    return bit_space3(i) | (bit_space3(j) << 1) | (bit_space3(k) << 2);
}

static inline float normsq_f(float a, float b, float c) {
    /* sequential, unfused: a*a + b*b + c*c */
    float s = a * a;
    s = s + b * b;
    s = s + c * c;
    return s;
}

void orc_partition(int64_t n, const double *x, int64_t ncell, const float *c, int32_t *aff) {
    for (int64_t i = 0; i < n; ++i) {
        const float p0 = (float)x[3 * i], p1 = (float)x[3 * i + 1], p2 = (float)x[3 * i + 2];
        float best = INFINITY;
        int32_t bi = -1;
        for (int64_t k = 0; k < ncell; ++k) {
            const float d = normsq_f(p0 - c[3 * k], p1 - c[3 * k + 1], p2 - c[3 * k + 2]);
            if (d < best) { best = d; bi = (int32_t)k; }
        }
        aff[i] = bi;
    }
}

void orc_centroids(int64_t n, const double *x, const int32_t *aff, int64_t ncell, float *c) {
    double *s = calloc((size_t)(3 * ncell), sizeof(double));
    int64_t *cnt = calloc((size_t)ncell, sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) {
        const int32_t a = aff[i];
        for (int d = 0; d < 3; ++d) s[3 * a + d] += x[3 * i + d];
        cnt[a]++;
    }
    for (int64_t k = 0; k < ncell; ++k)
        for (int d = 0; d < 3; ++d)
            c[3 * k + d] = cnt[k] ? (float)(s[3 * k + d] / (double)cnt[k]) : c[3 * k + d];
    free(s); free(cnt);
}

int64_t orc_stencil(int64_t ncell, const float *c, int64_t cell, float rmax, int64_t cap, int64_t *out) {
    int64_t cnt = 0;
    const float r2 = rmax * rmax;
    for (int64_t k = 0; k < ncell; ++k) {
        const float d = normsq_f(c[3 * k] - c[3 * cell], c[3 * k + 1] - c[3 * cell + 1], c[3 * k + 2] - c[3 * cell + 2]);
        if (d < r2) { if (cnt < cap) out[cnt] = k; ++cnt; }
    }
    return cnt;
}

int64_t orc_pairs_voronoi_fp64(int64_t n, const double *x, const int32_t *type,
                               const int32_t *aff, int64_t ncell, const float *c,
                               int64_t cap, int64_t *out) {
    (void)ncell;
    int64_t cnt = 0;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = i + 1; j < n; ++j) {
            const double dx[3] = {x[3 * i] - x[3 * j], x[3 * i + 1] - x[3 * j + 1], x[3 * i + 2] - x[3 * j + 2]};
            if (!interacts_fp64(type[i], type[j], dot3(dx, dx))) continue;
            const int32_t a = aff[i], b = aff[j];
            const float d = normsq_f(c[3 * b] - c[3 * a], c[3 * b + 1] - c[3 * a + 1], c[3 * b + 2] - c[3 * a + 2]);
            const int nl = (type[i] == 0) + (type[j] == 0);
            const float R = nl == 2 ? 6.0f : (nl == 1 ? 8.0f : 9.0f);
            if (d < R * R) {
                if (cnt < cap) { out[2 * cnt] = i; out[2 * cnt + 1] = j; }
                ++cnt;
            }
        }
    return cnt;
}

double orc_constrain_volume(int64_t n, const double *nrm, const int32_t *type,
                            const int32_t *aff, int64_t ncell, const float *c,
                            double target, double strength, double *f) {
    const orc_ff *ff = orc_ff_get();
    double center[3] = {0, 0, 0};
    for (int64_t k = 0; k < ncell; ++k)
        for (int d = 0; d < 3; ++d) center[d] += c[3 * k + d];
    for (int d = 0; d < 3; ++d) center[d] /= (double)ncell;
    double *nm = calloc((size_t)(3 * ncell), sizeof(double));
    int64_t *cnt = calloc((size_t)ncell, sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) {
        if (type[i] != 0) continue;
        for (int d = 0; d < 3; ++d) nm[3 * aff[i] + d] += nrm[3 * i + d];
        cnt[aff[i]]++;
    }
    double vol = 0.0;
    for (int64_t k = 0; k < ncell; ++k) {
        double *m = nm + 3 * k;
        if (cnt[k] == 0) { m[0] = m[1] = m[2] = 0; continue; }
        normalize3(m);
        const double dist[3] = {c[3 * k] - center[0], c[3 * k + 1] - center[1], c[3 * k + 2] - center[2]};
        if (dot3(m, dist) < 0) { m[0] = -m[0]; m[1] = -m[1]; m[2] = -m[2]; }
        const double h = dot3(dist, m);
        vol += h * (double)cnt[k] * 3.1415926 * 1.26 / 4.0 / 3.0 * 1e-6;
    }
    const double fp = strength * (target - vol) / target;
    for (int64_t i = 0; i < n; ++i) {
        const double *m = nm + 3 * aff[i];
        for (int d = 0; d < 3; ++d) f[3 * i + d] += fp * m[d] * ff->mass[type[i]];
    }
    free(nm); free(cnt);
    return vol;
}

static int cmp_double(const void *a, const void *b) {
    const double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

int64_t orc_stray_delete(int64_t n, const double *x, const int32_t *type, const int32_t *aff, int64_t ncell,
                         const float *c, double tol, uint8_t *del) {
    double *d2 = malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    int64_t *cnt = calloc((size_t)ncell + 1, sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) {
        del[i] = 0;
        const int32_t k = aff[i];
        if (type[i] != 0 || k < 0 || k >= ncell) { d2[i] = 0.0; continue; }
        const double dx = x[3 * i] - (double)c[3 * k], dy = x[3 * i + 1] - (double)c[3 * k + 1],
                     dz = x[3 * i + 2] - (double)c[3 * k + 2];
        d2[i] = dx * dx + dy * dy + dz * dz;
        cnt[k + 1]++;
    }
    for (int64_t k = 0; k < ncell; ++k) cnt[k + 1] += cnt[k];          /* cell k: [cnt[k], cnt[k+1]) */
    double *sorted = malloc((size_t)(cnt[ncell] > 0 ? cnt[ncell] : 1) * sizeof(double));
    int64_t *fill = calloc((size_t)ncell, sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) {
        const int32_t k = aff[i];
        if (type[i] != 0 || k < 0 || k >= ncell) continue;
        sorted[cnt[k] + fill[k]++] = d2[i];
    }
    double *thr = malloc((size_t)(ncell > 0 ? ncell : 1) * sizeof(double));
    for (int64_t k = 0; k < ncell; ++k) {
        const int64_t m = cnt[k + 1] - cnt[k];
        thr[k] = 0.0;
        if (m == 0) continue;
        qsort(sorted + cnt[k], (size_t)m, sizeof(double), cmp_double);
        thr[k] = sorted[cnt[k] + m / 2] * tol * tol;       /* src/cleanup.h:52-54 */
    }
    int64_t nd = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int32_t k = aff[i];
        if (type[i] != 0 || k < 0 || k >= ncell) continue;
        if (!(d2[i] < thr[k])) { del[i] = 1; ++nd; }       /* keep iff d2 < threshold, :55-58 */
    }
    free(d2); free(cnt); free(sorted); free(fill); free(thr);
    return nd;
}

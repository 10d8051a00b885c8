// ref_harness.cpp -- drives the OpenRBC reference's own functions, compiled
// in place from its source tree (REF=<openrbc>/src, see oracle/Makefile), into
// oracle/_ref/libref_openrbc.so.
//
// TEST INFRASTRUCTURE ONLY: used to pin the oracle (tests/golden/make_golden.py,
// CPU tests) and as bench.py's cpu_baseline ("reference" kind).  Nothing here
// is product code; no reference source is copied, only #included.
//
// The reference state is FP32 (src/config_static.h:34) and is built with its
// own Makefile flags (src/Makefile:9) except -march=x86-64-v3 instead of
// -march=native, so the library also runs on the GPU host's CPU.
#include "runtime_parameter.h"
#include "container.h"
#include "forcefield.h"
#include "voronoi.h"
#include "compute_pairwise_fused.h"
#include "compute_bonded.h"
#include "integrate_nh.h"
#include "integrate_langevin.h"
#include "compute_temperature.h"
#include "reorder_morton.h"
#include "kdtree.h"
#include "cleanup.h"
#include "assign_temperature.h"
#include "trajectory.h"
#include "topology.h"
#include "init_rbc.h"

#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

using namespace openrbc;

namespace {

struct QuietCout {
    std::streambuf *old;
    std::ostringstream sink;
    QuietCout() : old(std::cout.rdbuf(sink.rdbuf())) {}
    ~QuietCout() { std::cout.rdbuf(old); }
};

RTParameter *make_param(double kT, double eta, double dt, int seed) {
    std::vector<std::string> a = {"ref", "-i", "lipid", "-T", std::to_string(kT), "-e", std::to_string(eta),
                                  "-k", std::to_string(dt), "--random-seed", std::to_string(seed)};
    std::vector<char *> argv;
    for (auto &s : a) argv.push_back(&s[0]);
    QuietCout q;
    RTParameter *p = new RTParameter((int)argv.size(), argv.data());
    p->kBT = (float)kT;   // to_string rounds to 6 decimals; set exactly
    p->eta = (float)eta;
    p->dt = dt;
    p->Q = 0;
    p->zeta = 0;
    return p;
}

template<class C>
void load(C &c, int n, const float *x, const float *v, const float *nrm, const float *o,
          const float *f, const float *t) {
    c.resize(n);
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            c.x[i][d] = x ? x[3 * i + d] : 0.f;
            c.v[i][d] = v ? v[3 * i + d] : 0.f;
            c.n[i][d] = nrm ? nrm[3 * i + d] : 0.f;
            c.o[i][d] = o ? o[3 * i + d] : 0.f;
            c.f[i][d] = f ? f[3 * i + d] : 0.f;
            c.t[i][d] = t ? t[3 * i + d] : 0.f;
        }
}

template<class C>
void store(C const &c, float *x, float *v, float *nrm, float *o, float *f, float *t) {
    for (std::size_t i = 0; i < c.size(); ++i)
        for (int d = 0; d < 3; ++d) {
            if (x) x[3 * i + d] = c.x[i][d];
            if (v) v[3 * i + d] = c.v[i][d];
            if (nrm) nrm[3 * i + d] = c.n[i][d];
            if (o) o[3 * i + d] = c.o[i][d];
            if (f) f[3 * i + d] = c.f[i][d];
            if (t) t[3 * i + d] = c.t[i][d];
        }
}

void load_protein_meta(ProteContainer &p, int np, const int *ptype, const int *ptag) {
    for (int i = 0; i < np; ++i) {
        p.type[i] = ptype[i];
        p.tag[i] = ptag ? ptag[i] : i + 1;
    }
}

}  // namespace

extern "C" {

// Force-field tables as the reference holds them (FP32).
// out: mass[6] radius[6] cutsqlp[6] replp[6] attlp[6] alphalp[6] cutlp[6]
//      cutsqpp[36] reppp[36] cutpp[36] lj_cutsq[36] lj_lj1[36] lj_lj2[36] r0[4] K[4]
void ref_ff_tables(float *out) {
    int k = 0;
    for (int i = 0; i < 6; ++i) out[k++] = ForceField::mass[i];
    for (int i = 0; i < 6; ++i) out[k++] = ForceField::radius[i];
    for (int i = 0; i < 6; ++i) out[k++] = ForceField::cutsqlp[i];
    for (int i = 0; i < 6; ++i) out[k++] = ForceField::replp[i];
    for (int i = 0; i < 6; ++i) out[k++] = ForceField::attlp[i];
    for (int i = 0; i < 6; ++i) out[k++] = ForceField::alphalp[i];
    for (int i = 0; i < 6; ++i) out[k++] = ForceField::cutlp[i];
    for (int i = 0; i < 36; ++i) out[k++] = ForceField::cutsqpp[i];
    for (int i = 0; i < 36; ++i) out[k++] = ForceField::reppp[i];
    for (int i = 0; i < 36; ++i) out[k++] = ForceField::cutpp[i];
    for (int i = 0; i < 36; ++i) out[k++] = ForceField::lj_cutsq[i];
    for (int i = 0; i < 36; ++i) out[k++] = ForceField::lj_lj1[i];
    for (int i = 0; i < 36; ++i) out[k++] = ForceField::lj_lj2[i];
    for (int i = 0; i < 4; ++i) out[k++] = ForceField::r0[i];
    for (int i = 0; i < 4; ++i) out[k++] = ForceField::K[i];
}

// compute_pairwise_fused (src/compute_pairwise_fused.h:238) over ONE Voronoi
// cell holding every particle, so the sweep is the exact all-pairs set, then
// compute_bonded (src/compute_bonded.h:90).  Outputs accumulate from zero.
int ref_forces(int nl, const float *lx, const float *ln, int np, const float *px, const float *pn,
               const int *ptype, const int *ptag, int nb, const int *bt, const int *bi_tag,
               const int *bj_tag, int do_pair, int do_bond, float *lf, float *lt, float *pf,
               float *pt) {
    omp_set_num_threads(1);
    LipidContainer lipid("lipid");
    ProteContainer protein("protein");
    load(lipid, nl, lx, nullptr, ln, nullptr, nullptr, nullptr);
    load(protein, np, px, nullptr, pn, nullptr, nullptr, nullptr);
    load_protein_meta(protein, np, ptype, ptag);
    lipid.tag.set_base(np + 1);
    protein.bonds.resize(0);
    for (int k = 0; k < nb; ++k) protein.bonds.emplace_back(bt[k], bi_tag[k], bj_tag[k]);
    protein.tag2idx.build_map(protein);

    VoronoiDiagram voronoi(1);
    voronoi.centroids.resize(1);
    voronoi.centroids[0] = vector<real, 3>(0, 0, 0);
    voronoi.tree.build(voronoi.centroids);
    VCellList cl, cp;
    cl.n_cells = cp.n_cells = 1;
    cl.cell_start.resize(2);
    cp.cell_start.resize(2);
    cl.cell_start[0] = 0; cl.cell_start[1] = nl;
    cp.cell_start[0] = 0; cp.cell_start[1] = np;

    if (do_pair) compute_pairwise_fused(voronoi, lipid, protein, cl, cp);
    if (do_bond && nb > 0) {
        Service<BalancerMap>::call()["protein-bonds"].set_range(nb);
        compute_bonded(protein);
    }
    store(lipid, nullptr, nullptr, nullptr, nullptr, lf, lt);
    store(protein, nullptr, nullptr, nullptr, nullptr, pf, pt);
    return 0;
}

// integrate(verlet_langevin) (src/integrate_langevin.h:99-149) on both
// containers.  With kT = 0 the noise term vanishes and the step is
// deterministic.  All arrays are in/out.
void ref_langevin(int nl, float *lx, float *lv, float *ln, float *lo, float *lf, float *lt,
                  int np, float *px, float *pv, float *pn, float *po, float *pf, float *pt,
                  const int *ptype, double kT, double eta, double dt, int seed) {
    omp_set_num_threads(1);
    RTParameter *param = make_param(kT, eta, dt, seed);
    LipidContainer lipid("lipid");
    ProteContainer protein("protein");
    load(lipid, nl, lx, lv, ln, lo, lf, lt);
    load(protein, np, px, pv, pn, po, pf, pt);
    load_protein_meta(protein, np, ptype, nullptr);
    integrate(verlet_langevin(*param), lipid, protein);
    store(lipid, lx, lv, ln, lo, lf, lt);
    store(protein, px, pv, pn, po, pf, pt);
    delete param;
}

// Nose-Hoover fused halves, src/integrate_nh.h:178-235 (which=0) and
// :237-273 (which=1).  zeta/Q in, updated zeta returned.
double ref_nh(int which, int nl, float *lx, float *lv, float *ln, float *lo, float *lf, float *lt,
              int np, float *px, float *pv, float *pn, float *po, float *pf, float *pt,
              const int *ptype, double kT, double dt, double zeta, double Q) {
    omp_set_num_threads(1);
    RTParameter *param = make_param(kT, 0.01, dt, 1);
    param->zeta = (float)zeta;
    param->Q = (float)Q;
    LipidContainer lipid("lipid");
    ProteContainer protein("protein");
    load(lipid, nl, lx, lv, ln, lo, lf, lt);
    load(protein, np, px, pv, pn, po, pf, pt);
    load_protein_meta(protein, np, ptype, nullptr);
    if (which == 0)
        integrate(verlet_initial_bounce_clearforce_update(*param), lipid, protein);
    else
        integrate(post_toque_final_update(*param), lipid, protein);
    store(lipid, lx, lv, ln, lo, lf, lt);
    store(protein, px, pv, pn, po, pf, pt);
    const double z = param->zeta;
    delete param;
    return z;
}

// compute_temperature, src/compute_temperature.h:23-29
double ref_temperature(int nl, const float *lv, int np, const float *pv, const int *ptype) {
    omp_set_num_threads(1);
    RTParameter *param = make_param(0.22, 0.01, 0.01, 1);
    LipidContainer lipid("lipid");
    ProteContainer protein("protein");
    load(lipid, nl, nullptr, lv, nullptr, nullptr, nullptr, nullptr);
    load(protein, np, nullptr, pv, nullptr, nullptr, nullptr, nullptr);
    load_protein_meta(protein, np, ptype, nullptr);
    const double T = compute_temperature(lipid, protein, *param);
    delete param;
    return T;
}

// morton_encode, src/reorder_morton.h:37-42 (bsize = 2000, the default box)
void ref_morton(int n, const float *pts, unsigned *keys) {
    RTParameter *param = make_param(0.22, 0.01, 0.01, 1);
    for (int i = 0; i < n; ++i) keys[i] = morton_encode(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], *param);
    delete param;
}

// reorder_morton of a point set, src/reorder_morton.h:44-122: returns the
// permuted points (in place).
void ref_reorder_morton(int n, float *pts) {
    omp_set_num_threads(1);
    RTParameter *param = make_param(0.22, 0.01, 0.01, 1);
    AlignedArray<vector<real, 3>, true> p(n);
    for (int i = 0; i < n; ++i) p[i] = vector<real, 3>(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
    reorder_morton(p, *param);
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) pts[3 * i + d] = p[i][d];
    delete param;
}

// KDTree find_nearest / find_within, src/kdtree.h:206-285
void ref_kdtree_nearest(int n, const float *pts, int nq, const float *q, int *idx) {
    omp_set_num_threads(1);
    KDTree<real, 3, int> tree;
    AlignedArray<vector<real, 3>, true> p(n);
    for (int i = 0; i < n; ++i) p[i] = vector<real, 3>(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
    tree.build(p);
    for (int k = 0; k < nq; ++k) {
        vector<real, 3> qq(q[3 * k], q[3 * k + 1], q[3 * k + 2]);
        idx[k] = std::get<1>(tree.find_nearest(qq));
    }
}

// counts[k] = number found; out[k*cap ..] = indices (first cap)
void ref_kdtree_within(int n, const float *pts, int nq, const float *q, float rmax, int cap,
                       int *counts, int *out) {
    omp_set_num_threads(1);
    KDTree<real, 3, int> tree;
    AlignedArray<vector<real, 3>, true> p(n);
    for (int i = 0; i < n; ++i) p[i] = vector<real, 3>(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
    tree.build(p);
    AlignedArray<int, true> res;
    for (int k = 0; k < nq; ++k) {
        vector<real, 3> qq(q[3 * k], q[3 * k + 1], q[3 * k + 2]);
        const int c = tree.find_within(qq, rmax, res);
        counts[k] = c;
        for (int i = 0; i < c && i < cap; ++i) out[(long)k * cap + i] = res[i];
    }
}

// VCellList::partition, src/voronoi.h:179-237: nearest-centroid affiliation
// of every point, given fixed centroids.
void ref_partition(int n, const float *x, int ncell, const float *c, int *aff) {
    omp_set_num_threads(1);
    VoronoiDiagram voronoi(ncell);
    voronoi.centroids.resize(ncell);
    for (int k = 0; k < ncell; ++k) voronoi.centroids[k] = vector<real, 3>(c[3 * k], c[3 * k + 1], c[3 * k + 2]);
    voronoi.tree.build(voronoi.centroids);
    LipidContainer cont("lipid");
    load(cont, n, x, nullptr, nullptr, nullptr, nullptr, nullptr);
    VCellList cl;
    cl.affiliation.resize(n);
    for (int i = 0; i < n; ++i) cl.affiliation[i] = -1;
    cl.partition(cont, voronoi);
    for (int i = 0; i < n; ++i) aff[i] = cl.affiliation[i];
}

// delete_lipid, src/cleanup.h:29-91, on n lipids and fixed centroids: the
// cells are the reference's own partition and reorder (VCellList::update);
// aff[i] is the cell of input lipid i and keep[i] whether it survives
// (survivors are recognised by their positions, which delete_lipid copies
// unchanged; the inputs must be distinct points).
void ref_stray(int n, const float *x, int ncell, const float *c, double tol, int *aff, unsigned char *keep) {
    omp_set_num_threads(1);
    RTParameter *param = make_param(0.22, 0.01, 0.01, 1);
    param->stray_tolerance = (float)tol;
    VoronoiDiagram voronoi(ncell);
    voronoi.centroids.resize(ncell);
    for (int k = 0; k < ncell; ++k) voronoi.centroids[k] = vector<real, 3>(c[3 * k], c[3 * k + 1], c[3 * k + 2]);
    voronoi.tree.build(voronoi.centroids);
    LipidContainer lipid("lipid");
    load(lipid, n, x, nullptr, nullptr, nullptr, nullptr, nullptr);
    lipid.tag2idx.build_map(lipid);
    VCellList cl;
    cl.affiliation.resize(n);
    for (int i = 0; i < n; ++i) cl.affiliation[i] = -1;
    QuietCout q;
    cl.update(lipid, voronoi, *param);
    std::map<std::vector<float>, int> where;
    for (int i = 0; i < n; ++i) where[{x[3 * i], x[3 * i + 1], x[3 * i + 2]}] = i;
    for (int k = 0; k < ncell; ++k)
        for (int j = cl.cell_start[k]; j < cl.cell_start[k + 1]; ++j)
            aff[where[{lipid.x[j][0], lipid.x[j][1], lipid.x[j][2]}]] = k;
    delete_lipid(lipid, voronoi, cl, *param);
    for (int i = 0; i < n; ++i) keep[i] = 0;
    for (std::size_t j = 0; j < lipid.size(); ++j) keep[where[{lipid.x[j][0], lipid.x[j][1], lipid.x[j][2]}]] = 1;
    delete param;
}

// save_frame (src/trajectory.h:61-105) and save_topology (src/topology.h:23-52)
// of containers loaded from the arrays (lipid tags are base + i, as the
// reference's LipidContainer numbers them).
void ref_save_frame(const char *path, int nl, const float *lx, const float *lv, const float *ln, const int *laff,
                    int lbase, int np, const float *px, const float *pv, const float *pn, const int *paff,
                    const int *ptype, const int *ptag, int nstep, int dump_field, int append) {
    RTParameter *param = make_param(0.22, 0.01, 0.01, 1);
    param->nstep = nstep;
    param->dump_field = dump_field;
    LipidContainer lipid("lipid");
    ProteContainer protein("protein");
    load(lipid, nl, lx, lv, ln, nullptr, nullptr, nullptr);
    load(protein, np, px, pv, pn, nullptr, nullptr, nullptr);
    load_protein_meta(protein, np, ptype, ptag);
    lipid.tag.set_base(lbase);
    VCellList cl, cp;
    cl.affiliation.resize(nl);
    cp.affiliation.resize(np);
    for (int i = 0; i < nl; ++i) cl.affiliation[i] = laff ? laff[i] : -1;
    for (int i = 0; i < np; ++i) cp.affiliation[i] = paff ? paff[i] : -1;
    {
        std::ofstream f(path, append ? (std::ios::binary | std::ios::app) : std::ios::binary);
        save_frame(f, lipid, protein, cl, cp, *param);
    }
    delete param;
}

void ref_save_topology(const char *path, int nl, const float *lx, int lbase, int np, const float *px,
                       const int *ptype, const int *ptag, int nb, const int *bt, const int *bi, const int *bj) {
    RTParameter *param = make_param(0.22, 0.01, 0.01, 1);
    LipidContainer lipid("lipid");
    ProteContainer protein("protein");
    load(lipid, nl, lx, nullptr, nullptr, nullptr, nullptr, nullptr);
    load(protein, np, px, nullptr, nullptr, nullptr, nullptr, nullptr);
    load_protein_meta(protein, np, ptype, ptag);
    lipid.tag.set_base(lbase);
    protein.bonds.resize(0);
    for (int k = 0; k < nb; ++k) protein.bonds.emplace_back(bt[k], bi[k], bj[k]);
    save_topology(std::ofstream(path), *param, protein, lipid);
    delete param;
}

// compute_pairwise_fused (src/compute_pairwise_fused.h:238-320) over the given
// Voronoi cells: the containers arrive sorted by cell (lcell_start /
// pcell_start: n_cells + 1 offsets each), the centroids fix the stencils
// (radii 9 / 8 / 6 for protein-protein / protein-lipid / lipid-lipid), then
// compute_bonded.  One thread: every stencil cell is inside the thread's
// range (newton_on).  Outputs accumulate from zero, in the input order.
int ref_forces_cells(int nl, const float *lx, const float *ln, int np, const float *px, const float *pn,
                     const int *ptype, const int *ptag, int nb, const int *bt, const int *bi_tag,
                     const int *bj_tag, int ncell, const float *centroids, const int *lcell_start,
                     const int *pcell_start, float *lf, float *lt, float *pf, float *pt) {
    omp_set_num_threads(1);
    LipidContainer lipid("lipid");
    ProteContainer protein("protein");
    load(lipid, nl, lx, nullptr, ln, nullptr, nullptr, nullptr);
    load(protein, np, px, nullptr, pn, nullptr, nullptr, nullptr);
    load_protein_meta(protein, np, ptype, ptag);
    lipid.tag.set_base(np + 1);
    protein.bonds.resize(0);
    for (int k = 0; k < nb; ++k) protein.bonds.emplace_back(bt[k], bi_tag[k], bj_tag[k]);
    protein.tag2idx.build_map(protein);

    VoronoiDiagram voronoi(ncell);
    voronoi.centroids.resize(ncell);
    for (int k = 0; k < ncell; ++k)
        voronoi.centroids[k] = vector<real, 3>(centroids[3 * k], centroids[3 * k + 1], centroids[3 * k + 2]);
    voronoi.tree.build(voronoi.centroids);
    VCellList cl, cp;
    cl.n_cells = cp.n_cells = ncell;
    cl.cell_start.resize(ncell + 1);
    cp.cell_start.resize(ncell + 1);
    for (int k = 0; k <= ncell; ++k) { cl.cell_start[k] = lcell_start[k]; cp.cell_start[k] = pcell_start[k]; }

    compute_pairwise_fused(voronoi, lipid, protein, cl, cp);
    if (nb > 0) {
        Service<BalancerMap>::call()["protein-bonds"].set_range(nb);
        compute_bonded(protein);
    }
    store(lipid, nullptr, nullptr, nullptr, nullptr, lf, lt);
    store(protein, nullptr, nullptr, nullptr, nullptr, pf, pt);
    return 0;
}

// init_rbc (src/init_rbc.h:30-337) on the mesh files prefix.{vert,bond,face}.txt
// with --random-seed `seed`.  Called with null arrays it only returns the
// counts (nl, np, nb); the second call fills the arrays (lipids in generation
// order, face by face; proteins in tag order; bonds as generated).  Returns
// the total mesh area.
double ref_init_rbc(const char *prefix, int seed, double rho, long *counts, float *lx, float *ln, float *px,
                    float *pn, int *ptype, int *ptag, int *bt, int *bi, int *bj) {
    omp_set_num_threads(1);
    RTParameter *param = make_param(0.22, 0.01, 0.01, seed);
    param->init = "trimesh";
    param->mesh = prefix;
    param->rho = (float)rho;
    LipidContainer lipid("lipid");
    ProteContainer protein("protein");
    double area;
    {
        QuietCout q;
        area = init_rbc(lipid, protein, *param, 0);    // its progress lines go to stdout
        std::fflush(stdout);
    }
    counts[0] = (long)lipid.size();
    counts[1] = (long)protein.size();
    counts[2] = (long)protein.bonds.size();
    if (lx) {
        store(lipid, lx, nullptr, ln, nullptr, nullptr, nullptr);
        store(protein, px, nullptr, pn, nullptr, nullptr, nullptr);
        for (std::size_t i = 0; i < protein.size(); ++i) { ptype[i] = protein.type[i]; ptag[i] = protein.tag[i]; }
        for (std::size_t k = 0; k < protein.bonds.size(); ++k) {
            bt[k] = protein.bonds[k].type; bi[k] = protein.bonds[k].i; bj[k] = protein.bonds[k].j;
        }
    }
    delete param;
    return area;
}

// Langevin statistics of the reference: nsteps of compute_pairwise_fused over
// one Voronoi cell holding every lipid (the exact pair set) and
// integrate(verlet_langevin) (src/integrate_langevin.h:99-149) from the given
// lipid state, on one thread.  Every `every` steps the state is copied out:
// xs, ns, vs hold (nsteps / every) snapshots of 3 nl floats each.
void ref_langevin_lipids(int nl, const float *lx, const float *ln, const float *lv, double kT, double eta, double dt,
                         int seed, int nsteps, int every, float *xs, float *ns, float *vs) {
    omp_set_num_threads(1);
    RTParameter *param = make_param(kT, eta, dt, seed);
    LipidContainer lipid("lipid");
    ProteContainer protein("protein");
    load(lipid, nl, lx, lv, ln, nullptr, nullptr, nullptr);
    protein.resize(0);
    VoronoiDiagram voronoi(1);
    voronoi.centroids.resize(1);
    voronoi.centroids[0] = vector<real, 3>(0, 0, 0);
    voronoi.tree.build(voronoi.centroids);
    VCellList cl, cp;
    cl.n_cells = cp.n_cells = 1;
    cl.cell_start.resize(2);
    cp.cell_start.resize(2);
    cl.cell_start[0] = 0; cl.cell_start[1] = nl;
    cp.cell_start[0] = 0; cp.cell_start[1] = 0;
    integrate(clear_force(), lipid, protein);
    long k = 0;
    for (int step = 0; step < nsteps; ++step) {
        param->nstep = step;
        compute_pairwise_fused(voronoi, lipid, protein, cl, cp);
        integrate(verlet_langevin(*param), lipid, protein);
        if ((step + 1) % every == 0) {
            store(lipid, xs + 3L * nl * k, vs + 3L * nl * k, ns + 3L * nl * k, nullptr, nullptr, nullptr);
            ++k;
        }
    }
    delete param;
}

// The reference main loop (src/openrbc.cpp:155-256, default build: fused
// pairwise + Langevin), timed over `nsteps` steps after Voronoi init with
// `n_lloyd` iterations.  Returns loop seconds; *n_final = particles at the end.
// Run it in a process of its own (function-local statics size to the thread
// count of the first call).
double ref_run(int nl, const float *lx, const float *ln, int np, const float *px, const float *pn,
               const int *ptype, const int *ptag, int nb, const int *bt, const int *bi_tag,
               const int *bj_tag, double kT, double eta, double dt, int seed, int nsteps,
               int n_lloyd, int threads, long *n_final) {
    if (threads > 0) omp_set_num_threads(threads);
    RTParameter *param = make_param(kT, eta, dt, seed);
    LipidContainer lipid("lipid");
    ProteContainer protein("protein");
    load(lipid, nl, lx, nullptr, ln, nullptr, nullptr, nullptr);
    load(protein, np, px, nullptr, pn, nullptr, nullptr, nullptr);
    load_protein_meta(protein, np, ptype, ptag);
    lipid.tag.set_base(np + 1);
    protein.bonds.resize(0);
    for (int k = 0; k < nb; ++k) protein.bonds.emplace_back(bt[k], bi_tag[k], bj_tag[k]);
    protein.tag2idx.build_map(protein);
    lipid.tag2idx.build_map(lipid);

    QuietCout q;
    VoronoiDiagram voronoi(std::max<std::size_t>(1, lipid.size() / param->voronoi_cell_size));
    VCellList cell_lipid, cell_protein;
    voronoi.init(lipid, cell_lipid, *param, n_lloyd);
    cell_lipid.update_particle_affiliation(lipid);
    cell_protein.update(protein, voronoi, *param);
    cell_protein.update_particle_affiliation(protein);

    voronoi.update(lipid, cell_lipid, *param);
    cell_lipid.update(lipid, voronoi, *param);
    cell_protein.update(protein, voronoi, *param);
    integrate(clear_force(), lipid, protein);
    integrate(assign_temperature(*param), lipid, protein);
    compute_pairwise_fused(voronoi, lipid, protein, cell_lipid, cell_protein);
    compute_bonded(protein);
    integrate(verlet_langevin(*param), lipid, protein);

    param->nstep = 0;
    const double t0 = omp_get_wtime();
    while (param->nstep < nsteps) {
        if (param->nstep % param->freq_voronoi == 0) {
            if (param->nstep % param->freq_cleanup == 0) delete_lipid(lipid, voronoi, cell_lipid, *param);
            voronoi.update(lipid, cell_lipid, *param);
            cell_lipid.update(lipid, voronoi, *param);
            cell_protein.update(protein, voronoi, *param);
        }
        compute_pairwise_fused(voronoi, lipid, protein, cell_lipid, cell_protein);
        compute_bonded(protein);
        integrate(verlet_langevin(*param), lipid, protein);
        ++param->nstep;
    }
    const double t1 = omp_get_wtime();
    if (n_final) *n_final = (long)(lipid.size() + protein.size());
    delete param;
    return t1 - t0;
}

}  // extern "C"

// systems.cpp -- mesh -> particle initialiser (host C++), following the rules
// of init_rbc (src/init_rbc.h:30-337):
//   * scale the mesh so the mean edge is 80 nm (1 unit = 4.4545 nm), :70-83
//   * face normals (v0-v1) x (v2-v1), vertex/edge normals = normalised sums, :85-118
//   * glycophorin (type 3) on vertices, actin (type 4) 2.25 below, bonded (type 0), :122-157
//   * immobile band-3 (type 2) on edge midpoints, :159-176
//   * 19 spectrin beads (type 5) per edge between the two actins, shrunk by one
//     unit at each end, jittered by 0.05 u11; bonds: chain (3), ends to actin
//     (2), bead 9 to the edge's band-3 (1), :200-233
//   * int(area * rho) lipids per face, uniform on the triangle (Osada et al.),
//     rejected within 1.685 of any protein or 0.55 of a lipid of the same face;
//     director = face normal, :235-305
// The arithmetic of the counts (mesh in FP32, area in FP32, rho as FP32)
// follows the reference so that int(area*rho) per face is the same; the random
// placement uses this library's own generator (splitmix64), not MT19937.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "../../include/orbc.h"

namespace {

struct F3 {
    float x, y, z;
};
inline F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline F3 mul(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline F3 cross(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float norm(F3 a) {
    volatile float s = a.x * a.x;
    s = s + a.y * a.y;
    s = s + a.z * a.z;
    return std::sqrt((float)s);
}
inline F3 normalize(F3 a) {
    const float r = 1.0f / norm(a);
    return mul(a, r);
}

struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double u01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double u11() { return 2.0 * u01() - 1.0; }
};

// uniform grid hashing of protein positions for the clutter test
struct PointGrid {
    double cs;
    std::unordered_map<uint64_t, std::vector<int>> cells;
    const std::vector<F3> *pts;
    static uint64_t key(long long x, long long y, long long z) {
        return ((uint64_t)(x + (1 << 20)) << 42) | ((uint64_t)(y + (1 << 20)) << 21) | (uint64_t)(z + (1 << 20));
    }
    void build(const std::vector<F3> &p, double cell) {
        cs = cell;
        pts = &p;
        for (int i = 0; i < (int)p.size(); ++i)
            cells[key((long long)std::floor(p[i].x / cs), (long long)std::floor(p[i].y / cs),
                      (long long)std::floor(p[i].z / cs))].push_back(i);
    }
    bool any_within(double x, double y, double z, double r) const {
        const long long cx = (long long)std::floor(x / cs), cy = (long long)std::floor(y / cs),
                        cz = (long long)std::floor(z / cs);
        const double r2 = r * r;
        for (long long a = -1; a <= 1; ++a)
            for (long long b = -1; b <= 1; ++b)
                for (long long c = -1; c <= 1; ++c) {
                    auto it = cells.find(key(cx + a, cy + b, cz + c));
                    if (it == cells.end()) continue;
                    for (int i : it->second) {
                        const F3 &p = (*pts)[i];
                        const double dx = p.x - x, dy = p.y - y, dz = p.z - z;
                        if (dx * dx + dy * dy + dz * dz < r2) return true;
                    }
                }
        return false;
    }
};

template <class T>
T *dup(const std::vector<T> &v) {
    T *p = static_cast<T *>(std::malloc(sizeof(T) * (v.size() ? v.size() : 1)));
    if (p && !v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

}  // namespace

extern "C" {

int orbc_init_rbc(const double *vert, long long nv, const int *edge, long long ne, const int *face, long long nf,
                  double rho, unsigned long long seed, orbc_system *out) {
    if (!vert || !edge || !face || !out || nv <= 0 || ne <= 0 || nf <= 0) return ORBC_ERR_INVALID;
    std::memset(out, 0, sizeof(*out));
    for (long long k = 0; k < 2 * ne; ++k)
        if (edge[k] < 0 || edge[k] >= nv) return ORBC_ERR_INVALID;
    for (long long k = 0; k < 3 * nf; ++k)
        if (face[k] < 0 || face[k] >= nv) return ORBC_ERR_INVALID;
    Rng rng{seed};
    std::vector<F3> v(nv);
    for (long long i = 0; i < nv; ++i) v[i] = {(float)vert[3 * i], (float)vert[3 * i + 1], (float)vert[3 * i + 2]};
    double avg = 0;
    for (long long k = 0; k < ne; ++k) avg += norm(sub(v[edge[2 * k]], v[edge[2 * k + 1]]));
    avg /= (double)ne;
    const double unit2reduced = (80.0 / 4.4545) / avg;
    const float layer = 2.25f;   // ForceField::r0[0]
    for (auto &p : v) p = mul(p, (float)unit2reduced);

    std::vector<F3> nface(nf), nvert(nv, F3{0, 0, 0}), nedge(ne, F3{0, 0, 0});
    std::vector<std::vector<long long>> v2f(nv);
    for (long long i = 0; i < nf; ++i) {
        const F3 a = v[face[3 * i]], b = v[face[3 * i + 1]], c = v[face[3 * i + 2]];
        nface[i] = normalize(cross(sub(a, b), sub(c, b)));
        for (int d = 0; d < 3; ++d) v2f[face[3 * i + d]].push_back(i);
    }
    for (long long k = 0; k < ne; ++k) {
        int shared = 0;
        for (long long f1 : v2f[edge[2 * k]])
            for (long long f2 : v2f[edge[2 * k + 1]])
                if (f1 == f2) { nedge[k] = add(nedge[k], nface[f1]); ++shared; }
        if (shared != 2) return ORBC_ERR_INVALID;   // src/init_rbc.h:103
        nedge[k] = normalize(nedge[k]);
    }
    for (long long i = 0; i < nv; ++i) {
        for (long long f : v2f[i]) nvert[i] = add(nvert[i], nface[f]);
        nvert[i] = normalize(nvert[i]);
    }

    const int spc = 19;
    const long long np = 2 * nv + ne + ne * spc;
    std::vector<F3> px((size_t)np), pn((size_t)np);
    std::vector<int> pt((size_t)np);
    std::vector<int> btype;
    std::vector<unsigned> bi, bj;
    for (long long i = 0; i < nv; ++i) {            // glycophorin
        px[i] = v[i]; pn[i] = nvert[i]; pt[i] = 3;
    }
    for (long long i = 0; i < nv; ++i) {            // actin
        const long long k = nv + i;
        px[k] = sub(v[i], mul(nvert[i], layer)); pn[k] = nvert[i]; pt[k] = 4;
        btype.push_back(0); bi.push_back((unsigned)(i + 1)); bj.push_back((unsigned)(k + 1));
    }
    for (long long e = 0; e < ne; ++e) {            // band-3
        const long long k = 2 * nv + e;
        px[k] = mul(add(v[edge[2 * e]], v[edge[2 * e + 1]]), 0.5f); pn[k] = nedge[e]; pt[k] = 2;
    }
    for (long long e = 0; e < ne; ++e) {            // spectrin
        const int a0 = edge[2 * e], a1 = edge[2 * e + 1];
        F3 beg = px[nv + a0], end = px[nv + a1];
        F3 inc = mul(sub(end, beg), 1.0f / (spc + 1));
        const F3 u = mul(inc, 1.0f / norm(inc));
        beg = add(beg, u);
        end = sub(end, u);
        inc = mul(sub(end, beg), 1.0f / (spc + 1));
        for (int j = 0; j < spc; ++j) {
            const long long k = 2 * nv + ne + e * spc + j;
            const F3 x = add(beg, mul(inc, (float)(j + 1)));
            const F3 noise{(float)rng.u11(), (float)rng.u11(), (float)rng.u11()};
            px[k] = add(x, mul(noise, 0.05f));
            pn[k] = x;
            pt[k] = 5;
            const unsigned tag = (unsigned)(k + 1);
            if (j > 0) { btype.push_back(3); bi.push_back(tag - 1); bj.push_back(tag); }
            if (j == 0) { btype.push_back(2); bi.push_back((unsigned)(nv + a0 + 1)); bj.push_back(tag); }
            if (j == spc / 2) { btype.push_back(1); bi.push_back((unsigned)(2 * nv + e + 1)); bj.push_back(tag); }
            if (j == spc - 1) { btype.push_back(2); bi.push_back((unsigned)(nv + a1 + 1)); bj.push_back(tag); }
        }
    }

    PointGrid grid;
    grid.build(px, 1.685);
    const double min_dist = 0.55, min_dist_lp = 1.685;   // :243-244
    const double rho_f = (double)(float)rho;
    std::vector<double> lx, ln;
    double total_area = 0;
    long long clutter_cap = 0;
    for (long long i = 0; i < nf; ++i) {
        const F3 v1 = v[face[3 * i]], v2 = v[face[3 * i + 1]], v3 = v[face[3 * i + 2]];
        const double area = 0.5 * (double)norm(cross(sub(v1, v2), sub(v3, v2)));
        total_area += area;
        const int n = (int)(area * rho_f);
        std::vector<double> pts;
        for (int j = 0; j < n; ++j) {
            double r[3];
            long long retry = 0;
            for (;;) {
                const double r1 = rng.u01(), r2 = rng.u01();
                const double s1 = std::sqrt(r1);
                const double w1 = 1.0 - s1, w2 = s1 * (1.0 - r2), w3 = s1 * r2;
                r[0] = v1.x * w1 + v2.x * w2 + v3.x * w3;
                r[1] = v1.y * w1 + v2.y * w2 + v3.y * w3;
                r[2] = v1.z * w1 + v2.z * w2 + v3.z * w3;
                bool clutter = grid.any_within(r[0], r[1], r[2], min_dist_lp);
                for (size_t k = 0; !clutter && k < pts.size(); k += 3) {
                    const double dx = pts[k] - r[0], dy = pts[k + 1] - r[1], dz = pts[k + 2] - r[2];
                    if (dx * dx + dy * dy + dz * dz < min_dist * min_dist) clutter = true;
                }
                if (!clutter) break;
                if (++retry > 1000LL * (n + 1)) { ++clutter_cap; break; }   // accept after many retries
            }
            for (int d = 0; d < 3; ++d) pts.push_back(r[d]);
            lx.insert(lx.end(), r, r + 3);
            ln.push_back(nface[i].x); ln.push_back(nface[i].y); ln.push_back(nface[i].z);
        }
    }
    const long long nl = (long long)lx.size() / 3;
    std::vector<double> pxd(3 * np), pnd(3 * np);
    std::vector<unsigned> pid(np), lid(nl);
    for (long long k = 0; k < np; ++k) {
        pxd[3 * k] = px[k].x; pxd[3 * k + 1] = px[k].y; pxd[3 * k + 2] = px[k].z;
        pnd[3 * k] = pn[k].x; pnd[3 * k + 1] = pn[k].y; pnd[3 * k + 2] = pn[k].z;
        pid[k] = (unsigned)(k + 1);
    }
    for (long long k = 0; k < nl; ++k) lid[k] = (unsigned)(np + 1 + k);   // lipid tag base, :306
    out->n_lipid = nl;
    out->n_protein = np;
    out->n_bond = (long long)btype.size();
    out->lipid_x = dup(lx);
    out->lipid_n = dup(ln);
    out->lipid_id = dup(lid);
    out->protein_x = dup(pxd);
    out->protein_n = dup(pnd);
    out->protein_type = dup(pt);
    out->protein_id = dup(pid);
    out->bond_type = dup(btype);
    out->bond_i = dup(bi);
    out->bond_j = dup(bj);
    out->area = total_area;
    out->clutter_capped = clutter_cap;
    return ORBC_OK;
}

void orbc_system_free(orbc_system *s) {
    if (!s) return;
    void *ps[] = {s->lipid_x, s->lipid_n, s->lipid_id, s->protein_x, s->protein_n, s->protein_type, s->protein_id,
                  s->bond_type, s->bond_i, s->bond_j};
    for (void *p : ps) std::free(p);
    std::memset(s, 0, sizeof(*s));
}

}  // extern "C"

// io.cpp -- trajectory frames, LAMMPS topology (host side).
//
// save_frame / read_frame (src/trajectory.h:26-150) and save_topology
// (src/topology.h:23-52) of the reference, over the device state: lipids
// first, then proteins, each species in upload order.  The frame layout is
// the reference's byte for byte: 8-byte zero-padded section titles, the step
// as a 32-bit int, the atom count as a 64-bit size, per particle a 32-bit
// tag and type, FP32 vectors and 32-bit Voronoi cells.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "ctx.h"

using namespace orbc;

#define guarded orbc_guarded

namespace {

// the reference's DumpField bits (src/runtime_parameter.h:27-33)
enum { kPosition = 1, kRotation = 2, kVoronoi = 4, kVelocity = 8, kForce = 16 };

void put_title(std::FILE *f, const char *t) {
    char b[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    std::memcpy(b, t, std::min<size_t>(8, std::strlen(t)));
    std::fwrite(b, 1, 8, f);
}

template <typename T>
void put(std::FILE *f, const T &v) { std::fwrite(&v, sizeof(T), 1, f); }

void put_vec(std::FILE *f, const std::vector<double> &a) {
    std::vector<float> b(a.size());
    for (size_t k = 0; k < a.size(); ++k) b[k] = (float)a[k];
    std::fwrite(b.data(), sizeof(float), b.size(), f);
}

struct Reader {
    std::FILE *f;
    bool ok = true;
    std::string title() {
        char b[9] = {0};
        if (std::fread(b, 1, 8, f) != 8) ok = false;
        return std::string(b);
    }
    template <typename T> T get() {
        T v{};
        if (std::fread(&v, sizeof(T), 1, f) != 1) ok = false;
        return v;
    }
    void read(void *p, size_t bytes) {
        if (p) { if (std::fread(p, 1, bytes, f) != bytes) ok = false; }
        else if (std::fseek(f, (long)bytes, SEEK_CUR) != 0) ok = false;
    }
};

}  // namespace

extern "C" {

int orbc_save_frame(orbc_ctx *c, const char *path, int dump_field, int append) {
    return guarded(c, [&] {
        if (!path) return fail(c, ORBC_ERR_INVALID, "null path");
        ensure_live(c);
        const long long nsp[2] = {c->nl, c->np};
        // per species, upload order
        std::vector<double> x[2], v[2], n[2], f[2];
        std::vector<int> aff[2];
        for (int s = 0; s < 2; ++s) {
            const long long m = nsp[s];
            x[s].resize(3 * m); v[s].resize(3 * m); n[s].resize(3 * m);
            if (m > 0) {
                int rc = orbc_get_state(c, s, x[s].data(), v[s].data(), n[s].data(), nullptr);
                if (rc != ORBC_OK) return rc;
                if (dump_field & kForce) {
                    f[s].resize(3 * m);
                    rc = orbc_get_forces(c, s, f[s].data(), nullptr, nullptr, nullptr);
                    if (rc != ORBC_OK) return rc;
                }
            }
            aff[s].assign(m, -1);
        }
        if ((dump_field & kVoronoi) && c->voro.ncell > 0) {
            int rc = orbc_voronoi_get(c, nullptr, aff[0].data(), aff[1].data());
            if (rc != ORBC_OK) return rc;
        }
        std::FILE *fp = std::fopen(path, append ? "ab" : "wb");
        if (!fp) return fail(c, ORBC_ERR_INVALID, std::string("cannot open ") + path);
        put_title(fp, "FRAMEBEG");
        put(fp, (int)c->step);
        put_title(fp, "NATOM");
        put(fp, (size_t)(nsp[0] + nsp[1]));
        put_title(fp, "IDENTITY");
        for (int s = 0; s < 2; ++s) {
            const Species &S = c->sp[s];
            for (long long i = 0; i < nsp[s]; ++i) {
                put(fp, (int)S.id[i]);
                put(fp, s == 0 ? 0 : (int)S.type[i]);
            }
        }
        const struct { int bit; const char *title; std::vector<double> *a; } vec[] = {
            {kPosition, "POSITION", x}, {kVelocity, "VELOCITY", v}, {kRotation, "ROTATION", n}};
        for (const auto &e : vec) {
            if (!(dump_field & e.bit)) continue;
            put_title(fp, e.title);
            put_vec(fp, e.a[0]); put_vec(fp, e.a[1]);
        }
        if (dump_field & kVoronoi) {
            put_title(fp, "VORONOI");
            for (int s = 0; s < 2; ++s) std::fwrite(aff[s].data(), sizeof(int), aff[s].size(), fp);
        }
        if (dump_field & kForce) {
            put_title(fp, "FORCE");
            put_vec(fp, f[0]); put_vec(fp, f[1]);
        }
        put_title(fp, "FRAMEEND");
        const bool ok = std::fclose(fp) == 0;
        return ok ? (int)ORBC_OK : fail(c, ORBC_ERR_INVALID, "write failed");
    });
}

int orbc_read_frame(const char *path, long long index, long long *step, long long *natom, int *fields,
                    unsigned *tag, int *type, float *x, float *v, float *nrm, int *aff, float *f) {
    if (!path || index < 0) return ORBC_ERR_INVALID;
    std::FILE *fp = std::fopen(path, "rb");
    if (!fp) return ORBC_ERR_INVALID;
    Reader r{fp};
    long long frame = -1, n = 0;
    int st = 0, got = 0;
    bool inside = false;
    while (r.ok) {
        const std::string t = r.title();
        if (!r.ok) break;
        const bool want = frame + (inside ? 0 : 1) == index;
        if (t == "FRAMEBEG") {
            ++frame; inside = true; got = 0;
            st = r.get<int>();
        } else if (t == "FRAMEEND") {
            inside = false;
            if (frame == index) break;
        } else if (t == "NATOM") {
            n = (long long)r.get<size_t>();
            if (n < 0 || n > (1LL << 34)) { r.ok = false; break; }
        } else if (t == "IDENTITY") {
            for (long long i = 0; i < n && r.ok; ++i) {
                const int tg = r.get<int>(), ty = r.get<int>();
                if (want && tag) tag[i] = (unsigned)tg;
                if (want && type) type[i] = ty;
            }
        } else if (t == "POSITION" || t == "VELOCITY" || t == "ROTATION" || t == "FORCE") {
            float *dst = t == "POSITION" ? x : t == "VELOCITY" ? v : t == "ROTATION" ? nrm : f;
            got |= t == "POSITION" ? kPosition : t == "VELOCITY" ? kVelocity : t == "ROTATION" ? kRotation : kForce;
            r.read(want ? dst : nullptr, 12 * (size_t)n);
        } else if (t == "VORONOI") {
            got |= kVoronoi;
            r.read(want ? aff : nullptr, 4 * (size_t)n);
        } else {
            r.ok = false;               // unknown section: its length is not known
        }
    }
    std::fclose(fp);
    if (frame < index || !r.ok) return ORBC_ERR_INVALID;
    if (step) *step = st;
    if (natom) *natom = n;
    if (fields) *fields = got;
    return ORBC_OK;
}

int orbc_save_topology(orbc_ctx *c, const char *path) {
    return guarded(c, [&] {
        if (!path) return fail(c, ORBC_ERR_INVALID, "null path");
        ensure_live(c);
        const long long nl = c->nl, np = c->np;
        std::vector<double> lx(3 * nl), px(3 * np);
        if (nl > 0 && orbc_get_state(c, ORBC_LIPID, lx.data(), nullptr, nullptr, nullptr) != ORBC_OK)
            return (int)ORBC_ERR_INVALID;
        if (np > 0 && orbc_get_state(c, ORBC_PROTEIN, px.data(), nullptr, nullptr, nullptr) != ORBC_OK)
            return (int)ORBC_ERR_INVALID;
        std::ofstream file(path);
        if (!file) return fail(c, ORBC_ERR_INVALID, std::string("cannot open ") + path);
        // src/topology.h:23-52 (proteins, then lipids); coordinates as the
        // reference's FP32 containers print them
        file << "WholeCell" << std::endl << std::endl;
        file << np + nl << " " << "atoms" << std::endl;
        file << c->b_type.size() << " " << "bonds" << std::endl;
        file << std::endl << std::endl;
        const char *ax[3] = {"xlo xhi", "ylo yhi", "zlo zhi"};
        for (int d = 0; d < 3; ++d)
            file << (float)c->cfg.box_lo[d] << " " << (float)c->cfg.box_hi[d] << " " << ax[d] << std::endl;
        file << std::endl << std::endl;
        file << NT << " " << "atom types" << std::endl;
        file << "1 " << " " << "bond types" << std::endl;
        file << std::endl << std::endl;
        file << "Atoms" << std::endl;
        const Species &P = c->sp[ORBC_PROTEIN], &L = c->sp[ORBC_LIPID];
        for (long long i = 0; i < np; ++i) {
            file << (int)P.id[i] << " " << "1 " << P.type[i] + 1 << " ";
            for (int d = 0; d < 3; ++d) file << (float)px[3 * i + d] << " ";
            file << std::endl;
        }
        for (long long i = 0; i < nl; ++i) {
            file << (int)L.id[i] << " " << "1 " << 0 + 1 << " ";
            for (int d = 0; d < 3; ++d) file << (float)lx[3 * i + d] << " ";
            file << std::endl;
        }
        file << std::endl << std::endl;
        file << "Bonds" << std::endl << std::endl;
        for (size_t b = 0; b < c->b_type.size(); ++b)
            file << (b + 1) << " " << c->b_type[b] << " " << (int)c->b_i[b] << " " << (int)c->b_j[b] << std::endl;
        file.close();
        return file ? (int)ORBC_OK : fail(c, ORBC_ERR_INVALID, "write failed");
    });
}

}  // extern "C"

// ---- restart files (no reference counterpart) --------------------------------
// The full FP64 state in upload order, so a run resumes where it stopped: the
// configuration (with the current Nose-Hoover friction and the resolved
// volume target), the step counter, both species (id, type, x, v, n, o and
// the force accumulators the Nose-Hoover first half-step reads), ensemble
// members, bonds, and the Voronoi centroids with every particle's cell.
namespace {

constexpr char kRestartMagic[9] = "ORBCRST1";
constexpr char kRestartEnd[9] = "ORBCEND1";

template <typename T>
void put_arr(std::FILE *f, const std::vector<T> &a) {
    const long long n = (long long)a.size();
    std::fwrite(&n, sizeof(n), 1, f);
    if (n) std::fwrite(a.data(), sizeof(T), a.size(), f);
}

template <typename T>
bool get_arr(std::FILE *f, std::vector<T> &a, long long max_n) {
    long long n = 0;
    if (std::fread(&n, sizeof(n), 1, f) != 1 || n < 0 || n > max_n) return false;
    a.resize((size_t)n);
    return n == 0 || std::fread(a.data(), sizeof(T), a.size(), f) == a.size();
}

}  // namespace

extern "C" {

int orbc_save_restart(orbc_ctx *c, const char *path) {
    return guarded(c, [&] {
        if (!path) return fail(c, ORBC_ERR_INVALID, "null path");
        ensure_live(c);
        orbc_config cfg = c->cfg;
        double zeta = 0.0;
        ORBC_HIP_CHECK(hipMemcpyAsync(&zeta, c->scal, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        sync(c);
        cfg.zeta = zeta;
        if (c->vol_target_set) cfg.volume_target = c->vol_target;
        if (!(cfg.Q > 0) && c->nh_q > 0) cfg.Q = c->nh_q;       // the Nose-Hoover mass in use
        std::FILE *fp = std::fopen(path, "wb");
        if (!fp) return fail(c, ORBC_ERR_INVALID, std::string("cannot open ") + path);
        std::fwrite(kRestartMagic, 1, 8, fp);
        const unsigned cs = (unsigned)sizeof(orbc_config);
        std::fwrite(&cs, sizeof(cs), 1, fp);
        std::fwrite(&cfg, sizeof(cfg), 1, fp);
        const long long head[3] = {c->step, c->nh_started ? 1LL : 0LL, c->deleted};
        std::fwrite(head, sizeof(head), 1, fp);
        for (int s = 0; s < 2; ++s) {
            const Species &S = c->sp[s];
            const long long m = s == 0 ? c->nl : c->np;
            std::vector<double> x(3 * m), v(3 * m), nn(3 * m), o(3 * m), f(3 * m), t(3 * m);
            if (m > 0) {
                int rc = orbc_get_state(c, s, x.data(), v.data(), nn.data(), o.data());
                if (rc == ORBC_OK) rc = orbc_get_forces(c, s, f.data(), t.data(), nullptr, nullptr);
                if (rc != ORBC_OK) { std::fclose(fp); return rc; }
            }
            put_arr(fp, S.id); put_arr(fp, S.type); put_arr(fp, S.group);
            put_arr(fp, x); put_arr(fp, v); put_arr(fp, nn); put_arr(fp, o); put_arr(fp, f); put_arr(fp, t);
        }
        put_arr(fp, c->group_kbt);
        put_arr(fp, c->b_type); put_arr(fp, c->b_i); put_arr(fp, c->b_j);
        std::vector<float> cen(3 * std::max(0LL, c->voro.ncell));
        std::vector<int> al(c->voro.ncell > 0 ? c->nl : 0), ap(c->voro.ncell > 0 ? c->np : 0);
        if (c->voro.ncell > 0) {
            const int rc = orbc_voronoi_get(c, cen.data(), al.data(), ap.data());
            if (rc != ORBC_OK) { std::fclose(fp); return rc; }
        }
        put_arr(fp, cen); put_arr(fp, al); put_arr(fp, ap);
        std::fwrite(kRestartEnd, 1, 8, fp);
        const bool ok = std::fclose(fp) == 0;
        return ok ? (int)ORBC_OK : fail(c, ORBC_ERR_INVALID, "write failed");
    });
}

int orbc_load_restart(orbc_ctx *c, const char *path) {
    return guarded(c, [&] {
        if (!path) return fail(c, ORBC_ERR_INVALID, "null path");
        std::FILE *fp = std::fopen(path, "rb");
        if (!fp) return fail(c, ORBC_ERR_INVALID, std::string("cannot open ") + path);
        struct Closer { std::FILE *f; ~Closer() { std::fclose(f); } } closer{fp};
        char magic[8];
        unsigned cs = 0;
        orbc_config cfg;
        long long head[3];
        const long long big = 1LL << 34;
        if (std::fread(magic, 1, 8, fp) != 8 || std::memcmp(magic, kRestartMagic, 8) != 0 ||
            std::fread(&cs, sizeof(cs), 1, fp) != 1 || cs != sizeof(orbc_config) ||
            std::fread(&cfg, sizeof(cfg), 1, fp) != 1 || std::fread(head, sizeof(head), 1, fp) != 1)
            return fail(c, ORBC_ERR_INVALID, "not a restart file of this build");
        std::vector<uint32_t> id[2];
        std::vector<int> type[2];
        std::vector<uint16_t> group[2];
        std::vector<double> x[2], v[2], nn[2], o[2], f[2], t[2];
        bool ok = true;
        for (int s = 0; s < 2 && ok; ++s)
            ok = get_arr(fp, id[s], big) && get_arr(fp, type[s], big) && get_arr(fp, group[s], big) &&
                 get_arr(fp, x[s], 3 * big) && get_arr(fp, v[s], 3 * big) && get_arr(fp, nn[s], 3 * big) &&
                 get_arr(fp, o[s], 3 * big) && get_arr(fp, f[s], 3 * big) && get_arr(fp, t[s], 3 * big);
        std::vector<double> gk;
        std::vector<int> bt;
        std::vector<uint32_t> bi, bj;
        std::vector<float> cen;
        std::vector<int> al, ap;
        ok = ok && get_arr(fp, gk, 65536) && get_arr(fp, bt, big) && get_arr(fp, bi, big) && get_arr(fp, bj, big) &&
             get_arr(fp, cen, 3 * big) && get_arr(fp, al, big) && get_arr(fp, ap, big);
        char end[8];
        ok = ok && std::fread(end, 1, 8, fp) == 8 && std::memcmp(end, kRestartEnd, 8) == 0;
        for (int s = 0; s < 2 && ok; ++s) {
            const size_t m = id[s].size();
            ok = type[s].size() == m && x[s].size() == 3 * m && v[s].size() == 3 * m && nn[s].size() == 3 * m &&
                 o[s].size() == 3 * m && f[s].size() == 3 * m && t[s].size() == 3 * m &&
                 (group[s].empty() || group[s].size() == m);
        }
        ok = ok && bi.size() == bt.size() && bj.size() == bt.size() && cen.size() % 3 == 0 &&
             (cen.empty() || (al.size() == id[0].size() && ap.size() == id[1].size()));
        if (!ok) return fail(c, ORBC_ERR_INVALID, "truncated or inconsistent restart file");
        // the configuration of the saved run (this context keeps its device)
        cfg.device = c->cfg.device;
        int rc = orbc_set_config_internal(c, cfg);
        if (rc != ORBC_OK) return rc;
        for (int s = 0; s < 2; ++s) {
            const long long m = (long long)id[s].size();
            rc = orbc_set_state(c, s, m, x[s].data(), v[s].data(), nn[s].data(), o[s].data(), type[s].data(),
                                id[s].data());
            if (rc != ORBC_OK) return rc;
            if (!gk.empty() && !group[s].empty()) {
                std::vector<int> g(group[s].begin(), group[s].end());
                rc = orbc_set_groups(c, s, m, g.data(), (int)gk.size(), gk.data());
                if (rc != ORBC_OK) return rc;
            }
        }
        rc = orbc_set_bonds(c, (long long)bt.size(), bt.data(), bi.data(), bj.data());
        if (rc != ORBC_OK) return rc;
        ensure_live(c);
        for (int s = 0; s < 2; ++s)
            if (!id[s].empty() && (rc = orbc_set_forces(c, s, f[s].data(), t[s].data())) != ORBC_OK) return rc;
        if (!cen.empty()) {
            rc = orbc_voronoi_set_centroids(c, (long long)cen.size() / 3, cen.data());
            if (rc != ORBC_OK) return rc;
            voronoi_set_affiliation(c, al.data(), ap.data());
        }
        c->step = head[0];
        c->nh_started = head[1] != 0;
        c->deleted = head[2];
        if (cfg.volume_target > 0) { c->vol_target = cfg.volume_target; c->vol_target_set = true; }
        sync(c);
        return (int)ORBC_OK;
    });
}

}  // extern "C"

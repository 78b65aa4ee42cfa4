// The case-file and restart-file readers and writers (csrc/case_io.cpp, host
// C++) under AddressSanitizer and UndefinedBehaviorSanitizer, without a GPU:
// the engine entry points that case_io.cpp also calls are replaced here by
// functions that return an error (this driver never reaches them).
//
//   case_io_san input.txt griddata.txt spdata.txt workdir
//
// 1. reads the case, writes it again and reads that back (counts must agree);
// 2. reads copies of griddata.txt / spdata.txt cut short at many lengths and
//    copies with one byte replaced, which must be rejected or read, never
//    touch memory out of bounds or throw across the C ABI;
// 3. saves restart files for PIC, TPIC and APIC at 7 and 17 digits, loads
//    them, compares, and loads truncated and corrupted copies.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "mpm/mpm_capi.h"
#include "mpm/mpm_case.h"

namespace mpm {
int set_error_msg(int code, const char *) { return code; }
int exception_status() { return MPM_ERR_INTERNAL; }
}  // namespace mpm

extern "C" {
int mpm_engine_create(const mpm_grid_desc *, const mpm_material_desc *, int32_t, const mpm_step_params *,
                      mpm_engine **) {
    return MPM_ERR_UNSUPPORTED;
}
int mpm_engine_destroy(mpm_engine *) { return 0; }
int mpm_get_engine_info(mpm_engine *, mpm_engine_info *) { return MPM_ERR_UNSUPPORTED; }
int mpm_set_bc(mpm_engine *, int32_t, const int32_t *, const double *, int32_t) { return MPM_ERR_UNSUPPORTED; }
int mpm_upload_particles(mpm_engine *, const mpm_particles_view *) { return MPM_ERR_UNSUPPORTED; }
int mpm_download_particles(mpm_engine *, mpm_particles_view *) { return MPM_ERR_UNSUPPORTED; }
int mpm_set_shifting(mpm_engine *, int32_t, const int32_t *) { return MPM_ERR_UNSUPPORTED; }
int mpm_download_affine(mpm_engine *, double *, double *) { return MPM_ERR_UNSUPPORTED; }
int mpm_upload_affine(mpm_engine *, const double *, const double *) { return MPM_ERR_UNSUPPORTED; }
}

static int g_fail = 0;
#define CHECK(c, msg)                                     \
    do {                                                  \
        if (!(c)) {                                       \
            std::printf("FAIL: %s\n", msg);               \
            g_fail++;                                     \
        }                                                 \
    } while (0)

static std::string slurp(const std::string &p) {
    std::ifstream f(p, std::ios::binary);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

static void spit(const std::string &p, const std::string &s) {
    std::ofstream f(p, std::ios::binary);
    f << s;
}

static uint64_t g_rng = 0x2545F4914F6CDD1Dull;
static uint64_t next() {
    g_rng ^= g_rng << 13;
    g_rng ^= g_rng >> 7;
    g_rng ^= g_rng << 17;
    return g_rng;
}

// reads a case from possibly damaged files; any status is acceptable
static int try_read(const char *in, const char *gd, const char *sp) {
    mpm_case *c = nullptr;
    const int rc = mpm_case_read(in, gd, sp, &c);
    if (rc == 0) {
        mpm_case_desc d;
        if (mpm_case_describe(c, &d) == 0) {
            mpm_particles_view v;
            if (mpm_case_particles(c, &v) == 0 && d.n > 0) {
                volatile double s = 0.0;
                for (int64_t i = 0; i < d.n; i++) s += v.coord[3 * i] + v.mass[i] + v.vol0[i] + v.matid[i];
            }
            for (int comp = 0; comp < 3; comp++) {
                const int32_t *nbc;
                const double *fbc;
                int32_t ibc;
                if (mpm_case_bc(c, comp, &nbc, &fbc, &ibc) == 0)
                    for (int32_t i = 0; i < ibc; i++) (void)(nbc[i] + fbc[i]);
            }
        }
        mpm_case_free(c);
    }
    return rc;
}

static void damage_sweep(const std::string &in, const std::string &src, const std::string &work, bool is_grid,
                         const std::string &other) {
    const std::string text = slurp(src);
    const std::string path = work + (is_grid ? "/damaged_grid.txt" : "/damaged_sp.txt");
    int accepted = 0, rejected = 0;
    const size_t n = text.size();
    for (int t = 0; t < 60; t++) {
        std::string s = text;
        if (t < 30) {
            s.resize((size_t)(next() % (n + 1)));  // cut short
        } else {
            const char repl[] = {'9', '-', ' ', '\n', 'x', '.', 'e', '0'};
            s[(size_t)(next() % n)] = repl[next() % sizeof repl];  // one byte replaced
        }
        spit(path, s);
        const int rc = is_grid ? try_read(in.c_str(), path.c_str(), other.c_str())
                               : try_read(in.c_str(), other.c_str(), path.c_str());
        (rc == 0 ? accepted : rejected)++;
    }
    std::printf("%s damaged 60 times: %d read, %d rejected\n", is_grid ? "griddata" : "spdata", accepted, rejected);
}

static void restart_round_trip(const std::string &work) {
    const int n = 9;
    std::vector<double> coord(3 * n), vel(3 * n), mass(n), vol(n), vol0(n), stress(6 * n), F(9 * n), J(n), cm(9 * n),
        di(9 * n);
    std::vector<int32_t> matid(n), id(n), surf(n);
    for (int i = 0; i < n; i++) {
        for (int a = 0; a < 3; a++) {
            coord[3 * i + a] = 0.1 * i + 0.01 * a;
            vel[3 * i + a] = 1e-3 * (i - a);
        }
        for (int q = 0; q < 6; q++) stress[6 * i + q] = 10.0 * i - q;
        for (int q = 0; q < 9; q++) {
            F[9 * i + q] = (q % 4 == 0 ? 1.0 : 0.0) + 1e-4 * (i + q);
            cm[9 * i + q] = 0.5 * q - i;
            di[9 * i + q] = 1.0 / (1.0 + q + i);
        }
        mass[i] = 1.0 + i;
        vol0[i] = 1e-3;
        vol[i] = 1.001e-3;
        J[i] = vol[i] / vol0[i];
        matid[i] = i % 2;
        id[i] = i;
        surf[i] = i % 3 == 0;
    }
    mpm_particles_view v{};
    v.n = n;
    v.coord = coord.data();
    v.vel = vel.data();
    v.mass = mass.data();
    v.vol = vol.data();
    v.vol0 = vol0.data();
    v.stress = stress.data();
    v.def_grad = F.data();
    v.det_def_grad = J.data();
    v.matid = matid.data();
    const int schemes[3] = {MPM_SCHEME_PIC, MPM_SCHEME_TPIC, MPM_SCHEME_APIC};
    for (int scheme : schemes)
        for (int digits : {0, 17}) {
            const std::string path = work + "/sp0_re.txt";
            const int rc = mpm_restart_save(path.c_str(), &v, scheme, scheme >= MPM_SCHEME_TPIC ? cm.data() : nullptr,
                                            scheme == MPM_SCHEME_APIC ? di.data() : nullptr, id.data(), surf.data(),
                                            digits);
            CHECK(rc == 0, "restart save");
            mpm_restart *r = nullptr;
            CHECK(mpm_restart_load(path.c_str(), scheme, &r) == 0, "restart load");
            if (r) {
                mpm_particles_view w;
                const double *c2 = nullptr, *d2 = nullptr;
                const int32_t *id2 = nullptr, *surf2 = nullptr;
                CHECK(mpm_restart_view(r, &w, &c2, &d2, &id2, &surf2) == 0, "restart view");
                CHECK(w.n == n, "restart count");
                if (digits == 17 && w.n == n) {
                    bool same = true;
                    for (int i = 0; i < 3 * n; i++) same = same && w.coord[i] == coord[i];
                    for (int i = 0; i < 6 * n; i++) same = same && w.stress[i] == stress[i];
                    if (scheme == MPM_SCHEME_APIC)
                        for (int i = 0; i < 9 * n; i++) same = same && c2[i] == cm[i] && d2[i] == di[i];
                    CHECK(same, "17-digit restart reads back bit for bit");
                }
                mpm_restart_free(r);
            }
            // damaged copies
            const std::string text = slurp(path);
            const std::string bad = work + "/bad_re.txt";
            int rejected = 0;
            for (int t = 0; t < 30; t++) {
                std::string s = text;
                if (t % 2)
                    s.resize((size_t)(next() % (s.size() + 1)));
                else
                    s[(size_t)(next() % s.size())] = "9- \nx.e0"[next() % 8];
                spit(bad, s);
                mpm_restart *q = nullptr;
                const int rq = mpm_restart_load(bad.c_str(), scheme, &q);
                if (rq == 0) {
                    mpm_particles_view w;
                    if (mpm_restart_view(q, &w, nullptr, nullptr, nullptr, nullptr) == 0)
                        for (int64_t i = 0; i < w.n; i++) (void)(w.coord[3 * i] + w.vol[i]);
                    mpm_restart_free(q);
                } else {
                    rejected++;
                }
            }
            std::printf("restart scheme %d digits %d: round trip, 30 damaged copies (%d rejected)\n", scheme, digits,
                        rejected);
        }
    spit(work + "/huge_re.txt", "2000000000\n");
    mpm_restart *h = nullptr;
    CHECK(mpm_restart_load((work + "/huge_re.txt").c_str(), MPM_SCHEME_PIC, &h) != 0, "huge restart count rejected");
}

int main(int argc, char **argv) {
    if (argc < 5) {
        std::printf("usage: %s input.txt griddata.txt spdata.txt workdir\n", argv[0]);
        return 2;
    }
    const std::string in = argv[1], gd = argv[2], sp = argv[3], work = argv[4];
    mpm_case *c = nullptr;
    CHECK(mpm_case_read(in.c_str(), gd.c_str(), sp.c_str(), &c) == 0, "case read");
    if (c) {
        mpm_case_desc d;
        CHECK(mpm_case_describe(c, &d) == 0, "describe");
        std::vector<mpm_material_desc> mats(d.nmat > 0 ? d.nmat : 1);
        CHECK(mpm_case_materials(c, mats.data(), (int32_t)mats.size()) >= 0, "materials");
        const std::string gd2 = work + "/griddata_w.txt", sp2 = work + "/spdata_w.txt";
        CHECK(mpm_case_write(c, gd2.c_str(), sp2.c_str()) == 0, "case write");
        mpm_case *c2 = nullptr;
        CHECK(mpm_case_read(in.c_str(), gd2.c_str(), sp2.c_str(), &c2) == 0, "written case read back");
        if (c2) {
            mpm_case_desc d2;
            mpm_case_describe(c2, &d2);
            CHECK(d2.n == d.n && d2.ibc[0] == d.ibc[0] && d2.ibc[1] == d.ibc[1] && d2.ibc[2] == d.ibc[2],
                  "counts after the round trip");
            std::printf("case: %lld points, BC lists %d %d %d; written and read back\n", (long long)d.n, d.ibc[0],
                        d.ibc[1], d.ibc[2]);
            mpm_case_free(c2);
        }
        mpm_engine *e = nullptr;
        CHECK(mpm_case_engine(c, MPM_PREC_FP32, 0, 0, &e) != 0, "engine creation reaches the stand-in");
        mpm_case_free(c);
    }
    damage_sweep(in, gd, work, true, sp);
    damage_sweep(in, sp, work, false, gd);
    restart_round_trip(work);
    std::printf(g_fail ? "FAILED (%d)\n" : "OK\n", g_fail);
    return g_fail ? 1 : 0;
}

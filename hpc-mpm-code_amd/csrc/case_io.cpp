// Reference case files (input.txt, griddata.txt, spdata.txt) for one rank:
// readers in the layouts the reference's own readers expect, writers in the
// layouts its generator writes (include/mpm/mpm_case.h).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/mpm/mpm_capi.h"
#include "../../include/mpm/mpm_case.h"

namespace mpm {
int set_error_msg(int code, const char *msg);
int exception_status();
}

struct mpm_case {
    mpm_case_desc d{};
    std::vector<mpm_material_desc> mats;
    std::vector<double> coord, mass, vol0;
    std::vector<int32_t> id, matid, surf;
    std::vector<int32_t> nbc[3];
    std::vector<double> fbc[3];
    std::vector<int32_t> node_count;  // nelem, node, nodec as written in griddata.txt
};

namespace {

// Line-oriented reader with the reference's conventions: a header line is
// skipped whole, values are read with >>, the rest of a value line is skipped.
struct Reader {
    std::ifstream in;
    std::string name;
    bool ok = true;
    explicit Reader(const char *path) : in(path), name(path) { ok = in.is_open(); }
    void skip_line() { in.ignore(1 << 20, '\n'); }
    // Bytes in the file: a declared count of n records needs at least
    // 2 n bytes (one digit and a separator each), so counts above half the
    // size are rejected before anything is allocated for them.
    long long size() {
        const auto here = in.tellg();
        in.seekg(0, std::ios::end);
        const long long sz = (long long)in.tellg();
        in.seekg(here);
        return sz;
    }
    template <typename T>
    void get(T &v) {
        if (!(in >> v)) ok = false;
    }
};

int fail(const std::string &what) { return mpm::set_error_msg(MPM_ERR_INVALID_ARGUMENT, what.c_str()); }

int read_input(mpm_case *c, const char *path) {
    Reader r(path);
    if (!r.ok) return fail(std::string("cannot open ") + path);
    mpm_case_desc &d = c->d;
    std::string sol;
    int nr_flag = 0;
    double gamma_nb = 0, beta_nb = 0;
    int npropmax = 0;
    r.skip_line();
    r.get(sol);
    r.get(nr_flag);
    r.get(d.rstflag);
    r.get(d.step.nlstep);
    r.skip_line();
    r.skip_line();
    r.get(d.ista);
    r.get(d.iend);
    r.get(d.iout);
    r.skip_line();
    r.skip_line();
    r.get(d.step.dt);
    r.get(d.step.mtol);
    r.get(gamma_nb);
    r.get(beta_nb);
    r.skip_line();
    r.skip_line();
    for (int i = 0; i < 3; i++) r.get(d.grid.xmin[i]);
    r.skip_line();
    r.skip_line();
    for (int i = 0; i < 3; i++) r.get(d.grid.xmax[i]);
    r.skip_line();
    r.skip_line();
    for (int i = 0; i < 3; i++) r.get(d.grid.nelem[i]);
    r.skip_line();
    r.skip_line();
    for (int i = 0; i < 3; i++) r.get(d.grid.idimc[i]);
    r.skip_line();
    r.skip_line();
    for (int i = 0; i < 3; i++) r.get(d.npxye[i]);
    r.skip_line();
    r.skip_line();
    r.get(d.rho);
    r.get(d.nmat);
    r.get(npropmax);
    r.skip_line();
    if (!r.ok || d.nmat < 1 || d.nmat > 1024 || npropmax < 0 || npropmax > 8)
        return fail(std::string("malformed parameters in ") + path);
    c->mats.assign(d.nmat, mpm_material_desc{});
    for (int m = 0; m < d.nmat; m++) {
        int iprop = 0, nprop = 0;
        r.skip_line();
        r.get(iprop);
        r.get(nprop);
        r.skip_line();
        r.skip_line();
        if (nprop < 0 || nprop > 8) return fail(std::string("bad nprop in ") + path);
        for (int i = 0; i < nprop; i++) r.get(c->mats[m].prop[i]);
        r.skip_line();
        c->mats[m].model = iprop;
        c->mats[m].nprop = nprop;
    }
    r.skip_line();
    for (int i = 0; i < 3; i++) r.get(d.step.bb[i]);
    if (!r.ok) return fail(std::string("malformed material or gravity block in ") + path);
    // ParseMapScheme (module/dataset.h:21-27)
    if (sol == "PIC")
        d.step.scheme = MPM_SCHEME_PIC;
    else if (sol == "FLIP")
        d.step.scheme = MPM_SCHEME_FLIP;
    else if (sol == "TPIC")
        d.step.scheme = MPM_SCHEME_TPIC;
    else if (sol == "APIC")
        d.step.scheme = MPM_SCHEME_APIC;
    else
        return fail("unknown mapping scheme '" + sol + "' in " + path);
    return 0;
}

int read_griddata(mpm_case *c, const char *path) {
    Reader r(path);
    if (!r.ok) return fail(std::string("cannot open ") + path);
    int idimc[3], nelem, node, nodec, xyelem[3], xynode[3], xynodec[3];
    double xmin[3], xmax[3];
    for (int i = 0; i < 3; i++) r.get(idimc[i]);
    r.skip_line();
    for (int i = 0; i < 3; i++) r.get(xmin[i]);
    r.skip_line();
    for (int i = 0; i < 3; i++) r.get(xmax[i]);
    r.skip_line();
    r.get(nelem);
    for (int i = 0; i < 3; i++) r.get(xyelem[i]);
    r.skip_line();
    r.get(node);
    for (int i = 0; i < 3; i++) r.get(xynode[i]);
    r.skip_line();
    r.get(nodec);
    for (int i = 0; i < 3; i++) r.get(xynodec[i]);
    r.skip_line();
    if (!r.ok) return fail(std::string("malformed mesh header in ") + path);
    for (int i = 0; i < 3; i++) {
        c->d.grid.idimc[i] = idimc[i];
        c->d.grid.xmin[i] = xmin[i];
        c->d.grid.xmax[i] = xmax[i];
        c->d.grid.nelem[i] = xyelem[i];
        if (xynode[i] != xyelem[i] + 1 || xynodec[i] != xyelem[i] + idimc[i])
            return fail(std::string("inconsistent node counts in ") + path);
    }
    c->node_count = {nelem, node, nodec};
    for (int comp = 0; comp < 3; comp++) {  // BCInput: ibc, then ibc lines "nbc fbc"
        int ibc = 0;
        r.get(ibc);
        r.skip_line();
        if (!r.ok || ibc < 0 || ibc > nodec || ibc > r.size() / 2)
            return fail(std::string("malformed BC list in ") + path);
        c->nbc[comp].resize(ibc);
        c->fbc[comp].resize(ibc);
        for (int i = 0; i < ibc; i++) {
            r.get(c->nbc[comp][i]);
            r.get(c->fbc[comp][i]);
            r.skip_line();
        }
        if (!r.ok) return fail(std::string("malformed BC entry in ") + path);
        c->d.ibc[comp] = ibc;
    }
    return 0;
}

int read_spdata(mpm_case *c, const char *path) {
    Reader r(path);
    if (!r.ok) return fail(std::string("cannot open ") + path);
    long long num = 0;
    r.get(num);
    r.skip_line();
    if (!r.ok || num < 0 || num >= (1ll << 31) || num > r.size() / 2)
        return fail(std::string("bad point count in ") + path);
    c->coord.resize(3 * num);
    for (long long i = 0; i < num; i++) {
        for (int a = 0; a < 3; a++) r.get(c->coord[3 * i + a]);
        r.skip_line();
    }
    c->id.resize(num);
    c->matid.resize(num);
    c->surf.assign(num, 0);
    c->mass.resize(num);
    c->vol0.resize(num);
    // Point lines: "id matid mass vol0" as the generator writes them
    // (solid_generator.cpp:203-206) or "id matid surf_point mass vol0" as the
    // solver's reader takes them (module/solid/solid_data_io.cpp:22). The
    // layout is fixed by the first line; a line with another column count
    // is an error.
    int cols = 0;
    std::string line;
    for (long long i = 0; r.ok && i < num; i++) {
        if (!std::getline(r.in, line)) {
            r.ok = false;
            break;
        }
        std::istringstream ls(line);
        std::vector<std::string> tok;
        for (std::string t; ls >> t;) tok.push_back(t);
        if (i == 0) cols = (int)tok.size();
        if ((cols != 4 && cols != 5) || (int)tok.size() != cols)
            return fail(std::string("point line with an unexpected column count in ") + path);
        std::istringstream vs(line);
        vs >> c->id[i] >> c->matid[i];
        if (cols == 5) vs >> c->surf[i];
        vs >> c->mass[i] >> c->vol0[i];
        if (vs.fail()) r.ok = false;
    }
    if (!r.ok) return fail(std::string("malformed point data in ") + path);
    c->d.n = num;
    return 0;
}

}  // namespace

extern "C" {

int mpm_case_read(const char *input_txt, const char *griddata_txt, const char *spdata_txt, mpm_case **out) try {
    if (!out) return fail("null output");
    *out = nullptr;
    auto *c = new mpm_case();
    c->d.step.mtol = 1e-12;
    c->d.step.nlstep = 1;
    int rc = 0;
    if (!rc && input_txt) rc = read_input(c, input_txt);
    if (!rc && griddata_txt) rc = read_griddata(c, griddata_txt);
    if (!rc && spdata_txt) rc = read_spdata(c, spdata_txt);
    if (!rc && !c->mats.empty())
        for (int32_t m : c->matid)
            if (m < 0 || m >= (int)c->mats.size()) {
                rc = fail("point material id out of range");
                break;
            }
    if (rc) {
        delete c;
        return rc;
    }
    *out = c;
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_case_describe(const mpm_case *c, mpm_case_desc *out) try {
    if (!c || !out) return fail("null argument");
    *out = c->d;
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_case_materials(const mpm_case *c, mpm_material_desc *mats, int32_t max) try {
    if (!c || (max > 0 && !mats)) return fail("null argument");
    for (int i = 0; i < (int)c->mats.size() && i < max; i++) mats[i] = c->mats[i];
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_case_particles(const mpm_case *c, mpm_particles_view *v) try {
    if (!c || !v) return fail("null argument");
    std::memset(v, 0, sizeof *v);
    v->n = c->d.n;
    auto *mc = const_cast<mpm_case *>(c);
    v->coord = mc->coord.data();
    v->mass = mc->mass.data();
    v->vol0 = mc->vol0.data();
    v->matid = mc->matid.data();
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_case_bc(const mpm_case *c, int32_t comp, const int32_t **nbc, const double **fbc, int32_t *ibc) try {
    if (!c || comp < 0 || comp > 2 || !nbc || !fbc || !ibc) return fail("bad arguments");
    *nbc = c->nbc[comp].data();
    *fbc = c->fbc[comp].data();
    *ibc = (int32_t)c->nbc[comp].size();
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_case_write(const mpm_case *c, const char *griddata_txt, const char *spdata_txt) try {
    if (!c) return fail("null case");
    const mpm_case_desc &d = c->d;
    if (griddata_txt) {
        std::ofstream o(griddata_txt);
        if (!o.is_open()) return fail(std::string("cannot write ") + griddata_txt);
        o.flags(std::ios::right | std::ios::scientific);
        int nel = 1, nod = 1, noc = 1;
        for (int i = 0; i < 3; i++) {
            nel *= d.grid.nelem[i];
            nod *= d.grid.nelem[i] + 1;
            noc *= d.grid.nelem[i] + d.grid.idimc[i];
        }
        for (int i = 0; i < 3; i++) o << std::setw(10) << d.grid.idimc[i];
        o << "\n";
        for (int i = 0; i < 3; i++) o << std::setw(15) << d.grid.xmin[i];
        o << "\n";
        for (int i = 0; i < 3; i++) o << std::setw(15) << d.grid.xmax[i];
        o << "\n";
        o << std::setw(10) << nel;
        for (int i = 0; i < 3; i++) o << std::setw(10) << d.grid.nelem[i];
        o << "\n";
        o << std::setw(10) << nod;
        for (int i = 0; i < 3; i++) o << std::setw(10) << d.grid.nelem[i] + 1;
        o << "\n";
        o << std::setw(10) << noc;
        for (int i = 0; i < 3; i++) o << std::setw(10) << d.grid.nelem[i] + d.grid.idimc[i];
        o << "\n";
        const char *names[3] = {"usbc", "vsbc", "wsbc"};
        for (int comp = 0; comp < 3; comp++) {
            o << std::setw(15) << c->nbc[comp].size() << std::setw(15) << "! ---- " << names[comp] << ".ibc ----"
              << "\n";
            for (size_t i = 0; i < c->nbc[comp].size(); i++)
                o << std::setw(15) << c->nbc[comp][i] << std::setw(15) << c->fbc[comp][i] << "\n";
        }
        if (!o.good()) return fail(std::string("write failed: ") + griddata_txt);
    }
    if (spdata_txt) {
        std::ofstream o(spdata_txt);
        if (!o.is_open()) return fail(std::string("cannot write ") + spdata_txt);
        o.flags(std::ios::right | std::ios::scientific);
        o << std::setw(10) << d.n << "\n";
        for (int64_t i = 0; i < d.n; i++) {
            for (int a = 0; a < 3; a++) o << std::setw(15) << c->coord[3 * i + a];
            o << "\n";
        }
        for (int64_t i = 0; i < d.n; i++)
            o << std::setw(10) << c->id[i] << std::setw(10) << c->matid[i] << std::setw(15) << c->mass[i]
              << std::setw(15) << c->vol0[i] << "\n";
        if (!o.good()) return fail(std::string("write failed: ") + spdata_txt);
    }
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_case_engine(const mpm_case *c, int32_t precision, int32_t relayout_every, int32_t shifting,
                    mpm_engine **out) try {
    if (!c || !out) return fail("null argument");
    if (c->mats.empty() || c->d.n <= 0) return fail("case has no materials or no points");
    mpm_step_params p = c->d.step;
    p.precision = precision;
    p.relayout_every = relayout_every;
    int rc = mpm_engine_create(&c->d.grid, c->mats.data(), (int32_t)c->mats.size(), &p, out);
    if (rc) return rc;
    for (int comp = 0; comp < 3 && !rc; comp++)
        rc = mpm_set_bc(*out, comp, c->nbc[comp].data(), c->fbc[comp].data(), (int32_t)c->nbc[comp].size());
    mpm_particles_view v;
    if (!rc) rc = mpm_case_particles(c, &v);
    if (!rc) rc = mpm_upload_particles(*out, &v);
    if (!rc && shifting) rc = mpm_set_shifting(*out, 1, c->d.npxye);
    if (rc) {
        mpm_engine_destroy(*out);
        *out = nullptr;
    }
    return rc;
} catch (...) {
    return mpm::exception_status();
}

void mpm_case_free(mpm_case *c) { delete c; }

}  // extern "C"

struct mpm_restart {
    int32_t scheme = 0;
    std::vector<double> coord, vel, mass, vol, vol0, stress, F, J, cmat, dinv;
    std::vector<int32_t> id, matid, surf;
};

namespace {

// OutputVector rows (module/data_io.h:49-60): width w per value
void put_rows(std::ofstream &o, const double *a, int64_t n, int per, int w) {
    for (int64_t i = 0; i < n; i++) {
        for (int q = 0; q < per; q++) o << std::setw(w) << (a ? a[per * i + q] : 0.0);
        o << "\n";
    }
}

// InputVector rows (module/data_io.h:18-38): per values, then the rest of the line
bool get_rows(Reader &r, std::vector<double> *a, int64_t n, int per) {
    if (a) a->resize((size_t)(per * n));
    for (int64_t i = 0; i < n; i++) {
        for (int q = 0; q < per; q++) {
            double v = 0.0;
            r.get(v);
            if (a) (*a)[per * i + q] = v;
        }
        r.skip_line();
    }
    return r.ok;
}

}  // namespace

extern "C" {

int mpm_restart_save(const char *path, const mpm_particles_view *p, int32_t scheme, const double *cmat,
                     const double *dinv, const int32_t *id, const int32_t *surf, int32_t digits) try {
    if (!path || !p || p->n < 0 || (p->n > 0 && (!p->coord || !p->mass || !p->vol0)))
        return fail("restart needs a path and a view with coord, mass, vol0");
    if (scheme < MPM_SCHEME_PIC || scheme > MPM_SCHEME_APIC) return fail("unknown transfer scheme");
    if (digits != 0 && (digits < 1 || digits > 17)) return fail("digits must be 0 or 1..17");
    std::ofstream o;
    o.flags(std::ios::right | std::ios::scientific);  // solid_data_io.cpp:112-113
    o.open(path);
    if (!o.is_open()) return fail(std::string("cannot write ") + path);
    const int w = digits ? digits + 9 : 15;
    if (digits) o << std::setprecision(digits - 1);
    const int64_t n = p->n;
    o << n << "     --- Number of solid material point --- " << "\n";
    if (n) {
        put_rows(o, p->coord, n, 3, w);
        for (int64_t i = 0; i < n; i++) {
            double vol = p->vol ? p->vol[i] : p->vol0[i];
            o << (id ? id[i] : (int32_t)i) << std::setw(15) << (p->matid ? p->matid[i] : 0) << std::setw(15)
              << (surf ? surf[i] : 0) << std::setw(w) << p->mass[i] << std::setw(w) << p->vol0[i] << std::setw(w)
              << vol << "\n";
        }
        put_rows(o, p->vel, n, 3, w);
        put_rows(o, nullptr, n, 3, w);  // accel
        put_rows(o, p->stress, n, 6, w);
        if (p->def_grad) {
            put_rows(o, p->def_grad, n, 9, w);
        } else {
            std::vector<double> eye(9 * (size_t)n, 0.0);
            for (int64_t i = 0; i < n; i++) eye[9 * i] = eye[9 * i + 4] = eye[9 * i + 8] = 1.0;
            put_rows(o, eye.data(), n, 9, w);
        }
        if (scheme == MPM_SCHEME_TPIC) {
            put_rows(o, cmat, n, 9, w);
            put_rows(o, nullptr, n, 9, w);  // accel_grad
        } else if (scheme == MPM_SCHEME_APIC) {
            put_rows(o, cmat, n, 9, w);
            put_rows(o, nullptr, n, 9, w);  // accel_Bmat
            put_rows(o, dinv, n, 9, w);
        }
    }
    if (!o.good()) return fail(std::string("write failed: ") + path);
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_restart_load(const char *path, int32_t scheme, mpm_restart **out) try {
    if (!path || !out) return fail("null argument");
    *out = nullptr;
    if (scheme < MPM_SCHEME_PIC || scheme > MPM_SCHEME_APIC) return fail("unknown transfer scheme");
    Reader r(path);
    if (!r.ok) return fail(std::string("cannot open ") + path);
    long long num = 0;
    r.get(num);
    r.skip_line();
    if (!r.ok || num < 0 || num >= (1ll << 31) || num > r.size() / 2)
        return fail(std::string("bad point count in ") + path);
    auto *R = new mpm_restart();
    R->scheme = scheme;
    const int64_t n = num;
    bool ok = get_rows(r, &R->coord, n, 3);
    R->id.resize(n);
    R->matid.resize(n);
    R->surf.resize(n);
    R->mass.resize(n);
    R->vol0.resize(n);
    R->vol.resize(n);
    R->J.resize(n);
    for (int64_t i = 0; ok && i < n; i++) {
        r.get(R->id[i]);
        r.get(R->matid[i]);
        r.get(R->surf[i]);
        r.get(R->mass[i]);
        r.get(R->vol0[i]);
        r.get(R->vol[i]);
        r.skip_line();
        R->J[i] = R->vol[i] / R->vol0[i];  // solid_data_io.cpp:79
        ok = r.ok;
    }
    ok = ok && get_rows(r, &R->vel, n, 3) && get_rows(r, nullptr, n, 3) && get_rows(r, &R->stress, n, 6) &&
         get_rows(r, &R->F, n, 9);
    if (ok && scheme == MPM_SCHEME_TPIC) ok = get_rows(r, &R->cmat, n, 9) && get_rows(r, nullptr, n, 9);
    if (ok && scheme == MPM_SCHEME_APIC)
        ok = get_rows(r, &R->cmat, n, 9) && get_rows(r, nullptr, n, 9) && get_rows(r, &R->dinv, n, 9);
    if (!ok) {
        delete R;
        return fail(std::string("malformed restart file ") + path);
    }
    *out = R;
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_restart_view(const mpm_restart *R, mpm_particles_view *v, const double **cmat, const double **dinv,
                     const int32_t **id, const int32_t **surf) try {
    if (!R || !v) return fail("null argument");
    auto *m = const_cast<mpm_restart *>(R);
    std::memset(v, 0, sizeof *v);
    v->n = (int64_t)R->mass.size();
    v->coord = m->coord.data();
    v->vel = m->vel.data();
    v->mass = m->mass.data();
    v->vol = m->vol.data();
    v->vol0 = m->vol0.data();
    v->stress = m->stress.data();
    v->def_grad = m->F.data();
    v->det_def_grad = m->J.data();
    v->matid = m->matid.data();
    if (cmat) *cmat = R->cmat.empty() ? nullptr : R->cmat.data();
    if (dinv) *dinv = R->dinv.empty() ? nullptr : R->dinv.data();
    if (id) *id = R->id.data();
    if (surf) *surf = R->surf.data();
    return 0;
} catch (...) {
    return mpm::exception_status();
}

void mpm_restart_free(mpm_restart *r) { delete r; }

int mpm_restart_write(mpm_engine *eng, const char *path, int32_t digits) try {
    mpm_engine_info info;
    int rc = mpm_get_engine_info(eng, &info);
    if (rc) return rc;
    if (info.n <= 0) return fail("no particles uploaded");
    const size_t n = (size_t)info.n;
    std::vector<double> x(3 * n), v(3 * n), m(n), V(n), V0(n), sig(6 * n), F(9 * n), J(n), cm, di;
    std::vector<int32_t> mid(n);
    mpm_particles_view p{};
    p.n = info.n;
    p.coord = x.data();
    p.vel = v.data();
    p.mass = m.data();
    p.vol = V.data();
    p.vol0 = V0.data();
    p.stress = sig.data();
    p.def_grad = F.data();
    p.det_def_grad = J.data();
    p.matid = mid.data();
    if ((rc = mpm_download_particles(eng, &p))) return rc;
    if (info.scheme >= MPM_SCHEME_TPIC) {
        cm.resize(9 * n);
        if (info.scheme == MPM_SCHEME_APIC) di.resize(9 * n);
        if ((rc = mpm_download_affine(eng, cm.data(), di.empty() ? nullptr : di.data()))) return rc;
    }
    return mpm_restart_save(path, &p, info.scheme, cm.empty() ? nullptr : cm.data(),
                            di.empty() ? nullptr : di.data(), nullptr, nullptr, digits);
} catch (...) {
    return mpm::exception_status();
}

int mpm_restart_read(mpm_engine *eng, const char *path) try {
    mpm_engine_info info;
    int rc = mpm_get_engine_info(eng, &info);
    if (rc) return rc;
    mpm_restart *R = nullptr;
    if ((rc = mpm_restart_load(path, info.scheme, &R))) return rc;
    mpm_particles_view p;
    const double *cm = nullptr, *di = nullptr;
    rc = mpm_restart_view(R, &p, &cm, &di, nullptr, nullptr);
    if (!rc) rc = mpm_upload_particles(eng, &p);
    if (!rc && info.scheme >= MPM_SCHEME_TPIC) rc = mpm_upload_affine(eng, cm, di);
    mpm_restart_free(R);
    return rc;
} catch (...) {
    return mpm::exception_status();
}

}  // extern "C"

// Drives the engine through the C++ host interface (include/mpm/mpm_explicit.hpp)
// with containers shaped like the reference's point container
// (module/material_point.h:113-153) and BoundaryCondition (module/bc.h:7-22),
// and checks it against the FP64 oracle (test infrastructure):
//   1. FP64-exact engine, reference pass sequence per step
//      (MeshPointLinklist, Particle2NodeS, SolveSolid, Node2ParticleS):
//      particle state and node mass bit-identical to the oracle after 40 steps;
//   2. linked lists after MeshPointLinklist equal the oracle's;
//   3. FP32 engine with Step(): positions within 1e-6 (domain length 1);
//   4. a particle outside the grid raises mpm::Error(MPM_ERR_OUT_OF_GRID);
//   5. implicit transfers (Particle2Node, Node2Particle for a given nodal
//      displacement) from the state after step 1: nodes, particles and
//      accelerations bit-identical to the oracle.
// Exit status 0 on success.
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "mpm/mpm_explicit.hpp"
#include "mpm_oracle.h"

struct Points {  // field names and types of the reference's point container
    int num = 0;
    std::vector<int> matid;
    std::vector<double> mass, vol, vol0, det_def_grad;
    std::vector<std::array<double, 3>> coord, vel, trac_force;
    std::vector<std::array<double, 6>> stress;
    std::vector<std::array<std::array<double, 3>, 3>> def_grad;
    std::vector<std::array<double, 3>> accel;
    std::vector<int> idepf, idp2p, numep;
    std::vector<double> nmass, nvof, nmome, nvel, naccel, nforce;
};

struct BC {  // BoundaryCondition
    int ibc = 0;
    std::vector<int> nbc;
    std::vector<double> fbc;
};

static const int NE = 8;
static const double H = 1.0 / NE, RHO = 1000.0, DT = 1e-5;

static Points make_points(int extra_outside) {
    Points p;
    uint64_t s = 0x5EED1234;
    auto rnd = [&]() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int k = 2; k < 5; k++)
        for (int j = 0; j < 3; j++)
            for (int i = 2; i < 5; i++)
                for (int c = 0; c < 8; c++) {
                    std::array<double, 3> x = {(i + 0.25 + 0.5 * (c & 1)) * H, (j + 0.25 + 0.5 * ((c >> 1) & 1)) * H,
                                               (k + 0.25 + 0.5 * (c >> 2)) * H};
                    p.coord.push_back(x);
                    p.vel.push_back({2e-3 * (rnd() - 0.5), 2e-3 * (rnd() - 0.5) + 0.05, 2e-3 * (rnd() - 0.5)});
                }
    if (extra_outside) p.coord.push_back({1.5, 0.5, 0.5}), p.vel.push_back({0, 0, 0});
    const int n = (int)p.coord.size();
    p.num = n;
    const double vp = H * H * H / 8.0;
    p.mass.assign(n, RHO * vp);
    p.vol.assign(n, vp);
    p.vol0.assign(n, vp);
    p.det_def_grad.assign(n, 1.0);
    p.stress.assign(n, {0, 0, 0, 0, 0, 0});
    p.def_grad.assign(n, {{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}});
    p.trac_force.assign(n, {0, 0, 0});
    p.matid.assign(n, 0);
    return p;
}

static BC floor_bc() {
    BC b;
    const int nx = NE + 1;
    for (int k = 0; k < nx; k++)
        for (int i = 0; i < nx; i++) b.nbc.push_back(i + nx * nx * k);  // j = 0
    b.ibc = (int)b.nbc.size();
    b.fbc.assign(b.ibc, 0.0);
    return b;
}

static mpm::Grid grid() {
    mpm::Grid g;
    g.xmin = {0, 0, 0};
    g.xmax = {1, 1, 1};
    g.nelem = {NE, NE, NE};
    return g;
}

static int fails = 0;
#define EXPECT(c, ...)                       \
    do {                                     \
        if (!(c)) {                          \
            std::printf("FAIL: " __VA_ARGS__); \
            std::printf("\n");               \
            fails++;                         \
        }                                    \
    } while (0)

template <class A>
static double maxdiff(const std::vector<A> &a, const double *b, size_t per) {
    const double *x = reinterpret_cast<const double *>(a.data());
    double m = 0;
    for (size_t i = 0; i < a.size() * per; i++) m = std::fmax(m, std::fabs(x[i] - b[i]));
    return m;
}

int main() {
    const int NSTEP = 40;
    const std::vector<mpm::Material> mats = {{MPM_MODEL_LINELAST, {1e6, 0.3}}};
    mpm::StepConfig cfg;
    cfg.dt = DT;
    cfg.bb = {0, -9.81, 0};
    cfg.precision = MPM_PREC_FP64_EXACT;
    cfg.relayout_every = 0;
    const BC fb = floor_bc();

    // ---- oracle
    Points o = make_points(0);
    const int n = o.num, nodec = (NE + 1) * (NE + 1) * (NE + 1);
    orc_state st;
    std::memset(&st, 0, sizeof st);
    for (int a = 0; a < 3; a++) {
        st.xmin[a] = 0;
        st.xmax[a] = 1;
        st.nex[a] = NE;
        st.idimc[a] = 1;
    }
    orc_grid_init(&st);
    st.dt = DT;
    st.mtol = cfg.mtol;
    st.bb[1] = -9.81;
    st.scheme = ORC_PIC;
    st.nlstep = 1;
    st.nmat = 1;
    st.iprop[0] = ORC_LINELAST;
    st.mat_prop[0][0] = 1e6;
    st.mat_prop[0][1] = 0.3;
    std::vector<int> obc(fb.nbc);
    std::vector<double> ofbc(fb.fbc);
    for (int c = 0; c < 3; c++) {
        st.ibc[c] = fb.ibc;
        st.nbc[c] = obc.data();
        st.fbc[c] = ofbc.data();
    }
    st.num = n;
    st.coord = &o.coord[0][0];
    st.vel = &o.vel[0][0];
    st.mass = o.mass.data();
    st.vol = o.vol.data();
    st.vol0 = o.vol0.data();
    st.stress = &o.stress[0][0];
    st.def_grad = &o.def_grad[0][0][0];
    st.det_def_grad = o.det_def_grad.data();
    st.trac_force = &o.trac_force[0][0];
    st.matid = o.matid.data();
    std::vector<int> idepf(st.nelem), idp2p(n), numep(st.nelem);
    std::vector<double> nmass(nodec), nvof(nodec), nmome(3 * nodec), nforce(3 * nodec), nvel(3 * nodec),
        naccel(3 * nodec), ddg(9 * (size_t)n);
    st.idepf = idepf.data();
    st.idp2p = idp2p.data();
    st.numep = numep.data();
    st.nmass = nmass.data();
    st.nvof = nvof.data();
    st.nmome = nmome.data();
    st.nforce = nforce.data();
    st.nvel = nvel.data();
    st.naccel = naccel.data();
    st.ddg = ddg.data();

    try {
        // ---- 2. linked lists on the initial state
        {
            Points p = make_points(0);
            mpm::ExplicitSolid eng(grid(), mats, cfg);
            eng.UploadPoints(p);
            eng.MeshPointLinklist();
            eng.DownloadLinklist(p);
            orc_bin(&st);
            EXPECT(p.idepf == idepf && p.idp2p == idp2p && p.numep == numep, "linked lists differ");
        }
        if (orc_run(&st, 1, NSTEP) != 0) {
            std::printf("FAIL: oracle run\n");
            return 2;
        }
        // ---- 1. FP64-exact, reference pass sequence
        {
            Points p = make_points(0);
            mpm::ExplicitSolid eng(grid(), mats, cfg);
            for (int c = 0; c < 3; c++) eng.SetBC(c, fb);
            eng.UploadPoints(p);
            for (int istep = 1; istep <= NSTEP; istep++) {
                const double facl = 1.0;  // nlstep = 1
                eng.MeshPointLinklist();
                eng.Particle2NodeS(facl);
                eng.SolveSolid();
                eng.Node2ParticleS(facl);
            }
            eng.DownloadPoints(p);
            eng.DownloadNodes(p);
            const double dx = maxdiff(p.coord, st.coord, 3), dv = maxdiff(p.vel, st.vel, 3),
                         ds = maxdiff(p.stress, st.stress, 6), df = maxdiff(p.def_grad, st.def_grad, 9),
                         dj = maxdiff(p.det_def_grad, st.det_def_grad, 1), dm = maxdiff(p.nmass, st.nmass, 1);
            std::printf("fp64-exact passes: max |diff| coord %.3g vel %.3g stress %.3g F %.3g J %.3g nmass %.3g\n",
                        dx, dv, ds, df, dj, dm);
            EXPECT(dx == 0 && dv == 0 && ds == 0 && df == 0 && dj == 0 && dm == 0, "FP64-exact not bit-identical");
        }
        // ---- 3. FP32 engine, whole steps
        {
            Points p = make_points(0);
            mpm::StepConfig c32 = cfg;
            c32.precision = MPM_PREC_FP32;
            c32.relayout_every = 8;
            mpm::ExplicitSolid eng(grid(), mats, c32);
            for (int c = 0; c < 3; c++) eng.SetBC(c, fb);
            eng.UploadPoints(p);
            eng.Step(1, NSTEP);
            eng.DownloadPoints(p);
            const double dx = maxdiff(p.coord, st.coord, 3);
            std::printf("fp32 steps: max |coord diff| %.3g\n", dx);
            EXPECT(dx < 1e-6, "FP32 positions off by %g", dx);
        }
        // ---- 4. error behaviour
        {
            Points p = make_points(1);
            mpm::ExplicitSolid eng(grid(), mats, cfg);
            int code = 0;
            try {
                eng.UploadPoints(p);
                eng.MeshPointLinklist();
            } catch (const mpm::Error &e) {
                code = e.code;
                std::printf("out-of-grid particle: %s\n", e.what());
            }
            EXPECT(code == MPM_ERR_OUT_OF_GRID, "expected MPM_ERR_OUT_OF_GRID, got %d", code);
        }
        // ---- 5. implicit transfers from the state the oracle reached
        {
            Points p = o;  // the oracle's state after the steps above
            p.accel.assign(n, {0, 0, 0});
            for (int i = 0; i < n; i++)
                for (int c = 0; c < 3; c++) p.accel[i][c] = 0.1 * std::sin(0.37 * i + c);
            std::vector<double> oacc(3 * (size_t)n);
            std::memcpy(oacc.data(), &p.accel[0][0], oacc.size() * sizeof(double));
            mpm::ExplicitSolid eng(grid(), mats, cfg);
            for (int c = 0; c < 3; c++) eng.SetBC(c, fb);
            eng.UploadPoints(p);
            eng.UploadAccel(p);
            eng.MeshPointLinklist();
            eng.Particle2Node(1.0);
            orc_bin(&st);
            orc_implicit_p2g(&st, oacc.data(), 1.0);
            eng.DownloadNodes(p);
            const double dnv = maxdiff(p.nvel, st.nvel, 1), dna = maxdiff(p.naccel, st.naccel, 1),
                         dnf = maxdiff(p.nforce, st.nforce, 1);
            std::vector<double> ndispl(3 * (size_t)nodec), nb(6);
            for (size_t q = 0; q < ndispl.size(); q++) ndispl[q] = 1e-6 * std::cos(0.11 * (double)q);
            orc_newmark_para(0.6, 0.3025, DT, nb.data());
            eng.Node2Particle(ndispl, nb, 0.6);
            if (orc_implicit_g2p(&st, oacc.data(), ndispl.data(), nb.data(), 0.6) != 0) {
                std::printf("FAIL: oracle implicit G2P\n");
                return 2;
            }
            eng.DownloadPoints(p);
            eng.DownloadAccel(p);
            const double dx = maxdiff(p.coord, st.coord, 3), dv = maxdiff(p.vel, st.vel, 3),
                         df = maxdiff(p.def_grad, st.def_grad, 9), ds = maxdiff(p.stress, st.stress, 6),
                         da = maxdiff(p.accel, oacc.data(), 3);
            std::printf("implicit transfers: max |diff| nvel %.3g naccel %.3g nforce %.3g coord %.3g vel %.3g F %.3g "
                        "stress %.3g accel %.3g\n",
                        dnv, dna, dnf, dx, dv, df, ds, da);
            EXPECT(dnv == 0 && dna == 0 && dnf == 0 && dx == 0 && dv == 0 && df == 0 && ds == 0 && da == 0,
                   "implicit transfers not bit-identical");
        }
    } catch (const mpm::Error &e) {
        std::printf("FAIL: %s\n", e.what());
        return 3;
    }
    std::printf(fails ? "FAILED %d\n" : "OK\n", fails);
    return fails ? 1 : 0;
}

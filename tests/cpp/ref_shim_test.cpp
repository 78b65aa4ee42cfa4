// Caller of the reference-signature host shim (include/mpm_ref/,
// libmpm_refshim.so), written the way the reference's own passes call that
// surface: globals set up as datain_para does (work/src_solid/explicit/
// datain_para.cpp:18-162), BuildMesh / BuildControlPoint, particle linked
// lists as MeshPointLinklist builds them (module/material_point.cpp:52-77),
// and the explicit step's passes over cells -> particles -> MakSf stencil
// (module/solid/explicit/var_trans_solid_explicit.cpp:20-311): P2G with
// StandardVarP2G-style sums and TPIC / APIC VarP2G, SolveSolid with
// BoundaryCondition, G2P with PIC / FLIP VarG2P, TPIC VarGradG2P, APIC
// VarBmatG2P / DmatG2P / InvDmatG2P, the MUSL remap, ComputeDeltaDefGrad and
// DetMat3, and the coordinate update.
//
// It includes only the shim's headers; the FP64 oracle (test
// infrastructure, a C library) is linked as the checker. Every compared
// field must be bit-identical to the oracle's for PIC, FLIP, TPIC and APIC
// with idimc (1,1,1), (2,2,2) and (2,1,2); MakSf and MakNodalVol likewise.
//
// With "--dump DIR" it also writes inputs and outputs of DetMat3, InvMat3,
// TraceMat3, PIC/FLIP VarG2P and BCSetVal/BCSetZero/BCSetDt as raw doubles,
// which the Python test checks against the reference's own code
// (oracle/_ref). Exit status 0 on success.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "bc.h"
#include "cal_mat.h"
#include "dataset.h"
#include "map_and_interpolate.h"
#include "mesh.h"
#include "shape_function.h"

#include "mpm_oracle.h"

using Mat3 = std::array<std::array<double, 3>, 3>;

static int g_fail = 0;
#define CHECK(cond, what)                                              \
    do {                                                               \
        if (!(cond)) {                                                 \
            std::printf("FAIL %s (%s:%d)\n", what, __FILE__, __LINE__); \
            g_fail++;                                                  \
        }                                                              \
    } while (0)

static uint64_t g_rng = 0x5EED5117ull;
static double urand() {
    uint64_t z = (g_rng += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static bool same(const double *a, const double *b, size_t n) { return std::memcmp(a, b, n * sizeof(double)) == 0; }

// ----- global set-up as datain_para does it, one rank -----
static void setup_grid(const double lo[3], const double hi[3], const int ne[3], const int order[3]) {
    for (int i = 0; i < 3; i++) {
        xyminw[i] = xymin[i] = lo[i];
        xymaxw[i] = xymax[i] = hi[i];
        xyelemw[i] = xyelem[i] = ne[i];
        idimc[i] = order[i];
        dxy[i] = (xymax[i] - xymin[i]) / xyelem[i];
        xynodew[i] = xynode[i] = xyelem[i] + 1;
        xynodecw[i] = xynodec[i] = xyelem[i] + idimc[i];
        aelemmin[i] = 0;
        aelemmax[i] = xyelem[i] - 1;
    }
    nelemw = nelem = xyelem[0] * xyelem[1] * xyelem[2];
    nodew = node = xynode[0] * xynode[1] * xynode[2];
    nodecw = nodec = xynodec[0] * xynodec[1] * xynodec[2];
    nuc = 0;
    nvc = nodec;
    nwc = 2 * nodec;
    BuildMesh();
    BuildControlPoint();
}

static void oracle_grid(orc_state &st) {
    std::memset(&st, 0, sizeof st);
    for (int i = 0; i < 3; i++) {
        st.xmin[i] = xymin[i];
        st.xmax[i] = xymax[i];
        st.nex[i] = xyelem[i];
        st.idimc[i] = idimc[i];
    }
    orc_grid_init(&st);
}

// ----- the caller's particle container (MaterialPoint's fields) -----
struct Points {
    int num = 0;
    std::vector<double> mass, vol, vol0, det_def_grad;
    std::vector<std::array<double, 3>> coord, vel, trac_force;
    std::vector<std::array<double, 6>> stress;
    std::vector<Mat3> def_grad;
    std::vector<int> idepf, idp2p, numep;
    std::vector<double> nmass, nvof, nmome, nforce, nvel, naccel;
    PIC pic;
    FLIP flip;
    TPIC tpic;
    APIC apic;
    MapScheme solswitch = MapScheme::PIC;
    BoundaryCondition usbc, vsbc, wsbc, rigid_bc;
};

// MeshPointLinklist: per-cell lists, particles appended in index order
static void link_list(Points &sp) {
    sp.idepf.assign(nelem, -1);
    sp.idp2p.assign(sp.num, -1);
    sp.numep.assign(nelem, 0);
    std::vector<int> tail(nelem, -1);
    for (int p = 0; p < sp.num; p++) {
        int ijk[3];
        for (int a = 0; a < 3; a++) ijk[a] = (int)std::floor((sp.coord[p][a] - xymin[a]) / dxy[a]);
        const int m = ijk[0] + ijk[1] * xyelem[0] + ijk[2] * xyelem[0] * xyelem[1];
        if (tail[m] < 0) sp.idepf[m] = p;
        else sp.idp2p[tail[m]] = p;
        tail[m] = p;
        sp.numep[m]++;
    }
}

static Points make_points(MapScheme scheme) {
    Points sp;
    sp.solswitch = scheme;
    // a block of 2x2x2 points per cell over the middle cells, jittered
    for (int k = 1; k < xyelem[2] - 1; k++)
        for (int j = 0; j < xyelem[1] - 1; j++)
            for (int i = 1; i < xyelem[0] - 1; i++)
                for (int q = 0; q < 8; q++) {
                    std::array<double, 3> x;
                    const int c[3] = {i, j, k};
                    for (int a = 0; a < 3; a++) {
                        const double f = 0.25 + 0.5 * ((q >> a) & 1) + 0.2 * (urand() - 0.5);
                        x[a] = xymin[a] + dxy[a] * (c[a] + f);
                    }
                    sp.coord.push_back(x);
                }
    sp.num = (int)sp.coord.size();
    const double volp = dxy[0] * dxy[1] * dxy[2] / 8.0;
    for (int p = 0; p < sp.num; p++) {
        sp.mass.push_back(1000.0 * volp * (0.9 + 0.2 * urand()));
        sp.vol0.push_back(volp);
        sp.vol.push_back(volp * (0.95 + 0.1 * urand()));
        sp.vel.push_back({urand() - 0.5, urand() - 0.5, urand() - 0.5});
        sp.trac_force.push_back({0.01 * (urand() - 0.5), 0.0, 0.02 * urand()});
        std::array<double, 6> s;
        for (double &v : s) v = 1e3 * (urand() - 0.5);
        sp.stress.push_back(s);
        Mat3 F;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) F[a][b] = (a == b ? 1.0 : 0.0) + 0.01 * (urand() - 0.5);
        sp.def_grad.push_back(F);
        sp.det_def_grad.push_back(DetMat3(F));
    }
    VectorAssign(sp.num, sp.tpic.vel_grad);
    VectorAssign(sp.num, sp.apic.vel_Bmat);
    VectorAssign(sp.num, sp.apic.inv_Dmat);
    for (int p = 0; p < sp.num; p++)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                sp.tpic.vel_grad[p][a][b] = urand() - 0.5;
                sp.apic.vel_Bmat[p][a][b] = 1e-3 * (urand() - 0.5);
                sp.apic.inv_Dmat[p][a][b] = (a == b ? 4.0 / (dxy[a] * dxy[a]) : 0.0) * (0.9 + 0.2 * urand());
            }
    // floor nodes fixed (all components), x = xmin roller (u), top w list
    // with a prescribed value
    const int nx = xynodec[0], ny = xynodec[1];
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            sp.rigid_bc.nbc.push_back(i + nx * j);
            sp.rigid_bc.fbc.push_back(0.0);
        }
    for (int k = 0; k < xynodec[2]; k++)
        for (int j = 0; j < ny; j++) {
            sp.usbc.nbc.push_back(nx * j + nx * ny * k);
            sp.usbc.fbc.push_back(0.0);
            sp.wsbc.nbc.push_back(nx - 1 + nx * j + nx * ny * k);
            sp.wsbc.fbc.push_back(0.125);
        }
    for (BoundaryCondition *b : {&sp.usbc, &sp.vsbc, &sp.wsbc, &sp.rigid_bc}) b->ibc = (int)b->nbc.size();
    return sp;
}

// ----- Particle2NodeS (var_trans_solid_explicit.cpp:20-99) -----
static void particle2node(Points &sp) {
    const double fx = bb[0] * facl, fy = bb[1] * facl, fz = bb[2] * facl;
    int nenode;
    std::vector<int> ncm;
    std::vector<double> sf;
    std::vector<std::array<double, 3>> dsf;
    VectorAssign(nodec, sp.nmass);
    VectorAssign(nodec, sp.nvof);
    VectorAssign(nodec * 3, sp.nmome);
    VectorAssign(nodec * 3, sp.nforce);
    for (int m = 0; m < nelem; m++) {
        for (int pid = sp.idepf[m]; pid != -1; pid = sp.idp2p[pid]) {
            const std::array<double, 3> xyp = sp.coord[pid];
            const std::array<double, 6> sts = sp.stress[pid];
            MakSf(m, xyp, idimc, xynodec, ncm, nenode, sf, dsf);
            for (int ni = 0; ni < nenode; ni++) {
                const int nid = ncm[ni];
                const double sfi = sf[ni];
                sp.nmass[nid] += sfi * sp.mass[pid];
                sp.nvof[nid] += sfi * sp.vol[pid];
                if (sp.solswitch == MapScheme::TPIC) {
                    sp.tpic.VarP2G(pid, nid, sfi, sp.tpic.vel_grad, sp.mass, sp.vel, sp.coord, sp.nmome);
                } else if (sp.solswitch == MapScheme::APIC) {
                    sp.apic.VarP2G(pid, nid, sfi, sp.apic.vel_Bmat, sp.mass, sp.vel, sp.coord, sp.nmome);
                } else {
                    sp.nmome[nid + nuc] += sfi * sp.mass[pid] * sp.vel[pid][0];
                    sp.nmome[nid + nvc] += sfi * sp.mass[pid] * sp.vel[pid][1];
                    sp.nmome[nid + nwc] += sfi * sp.mass[pid] * sp.vel[pid][2];
                }
                const double d1 = dsf[ni][0], d2 = dsf[ni][1], d3 = dsf[ni][2];
                const double fi[3] = {-sp.vol[pid] * (sts[0] * d1 + sts[5] * d2 + sts[4] * d3),
                                      -sp.vol[pid] * (sts[5] * d1 + sts[1] * d2 + sts[3] * d3),
                                      -sp.vol[pid] * (sts[4] * d1 + sts[3] * d2 + sts[2] * d3)};
                const double fe[3] = {sfi * (sp.mass[pid] * fx + sp.trac_force[pid][0]),
                                      sfi * (sp.mass[pid] * fy + sp.trac_force[pid][1]),
                                      sfi * (sp.mass[pid] * fz + sp.trac_force[pid][2])};
                sp.nforce[nid + nuc] += fi[0] + fe[0];
                sp.nforce[nid + nvc] += fi[1] + fe[1];
                sp.nforce[nid + nwc] += fi[2] + fe[2];
            }
        }
    }
    VectorAssign(nodec * 3, sp.nvel);
    for (int n = 0; n < nodec; n++) {
        if (sp.nmass[n] > mtol)
            for (int off : {nuc, nvc, nwc}) sp.nvel[n + off] = sp.nmome[n + off] / sp.nmass[n];
        if (sp.nvof[n] < mtol) sp.nvof[n] = 0.0e0;
    }
    sp.usbc.BCSetVal(nuc, sp.nvel);
    sp.vsbc.BCSetVal(nvc, sp.nvel);
    sp.wsbc.BCSetVal(nwc, sp.nvel);
    for (int off : {nuc, nvc, nwc}) sp.rigid_bc.BCSetVal(off, sp.nvel);
}

// ----- SolveSolid (solve_solid_explicit.cpp:8-31) -----
static void solve_solid(Points &sp) {
    VectorAssign(nodec * 3, sp.naccel);
    for (int n = 0; n < nodec; n++)
        if (sp.nmass[n] > mtol)
            for (int off : {nuc, nvc, nwc}) sp.naccel[n + off] = sp.nforce[n + off] / sp.nmass[n];
    sp.usbc.BCSetZero(nuc, sp.naccel);
    sp.vsbc.BCSetZero(nvc, sp.naccel);
    sp.wsbc.BCSetZero(nwc, sp.naccel);
    for (int off : {nuc, nvc, nwc}) sp.rigid_bc.BCSetZero(off, sp.naccel);
}

// ----- Node2ParticleS + MUSL + update_point_coord (var_trans_solid_explicit.cpp:101-311) -----
static void node2particle(Points &sp) {
    int nenode;
    std::vector<int> ncm;
    std::vector<double> sf;
    std::vector<std::array<double, 3>> dsf;
    if (sp.solswitch != MapScheme::FLIP) {
        for (int n = 0; n < 3 * nodec; n++) sp.nvel[n] += dt * sp.naccel[n];
        VectorAssign(sp.num, sp.vel);
    }
    if (sp.solswitch == MapScheme::TPIC) VectorAssign(sp.num, sp.tpic.vel_grad);
    if (sp.solswitch == MapScheme::APIC) {
        VectorAssign(sp.num, sp.apic.vel_Bmat);
        VectorAssign(sp.num, sp.apic.inv_Dmat);
    }
    for (int m = 0; m < nelem; m++)
        for (int pid = sp.idepf[m]; pid != -1; pid = sp.idp2p[pid]) {
            MakSf(m, sp.coord[pid], idimc, xynodec, ncm, nenode, sf, dsf);
            Mat3 ADp{};
            for (int ni = 0; ni < nenode; ni++) {
                const int nid = ncm[ni];
                const double sfi = sf[ni];
                if (sp.solswitch == MapScheme::FLIP) sp.flip.VarG2P(pid, nid, sfi, sp.vel, sp.naccel);
                else sp.pic.VarG2P(pid, nid, sfi, sp.vel, sp.nvel);
                if (sp.solswitch == MapScheme::TPIC) {
                    sp.tpic.VarGradG2P(pid, nid, ni, dsf, sp.tpic.vel_grad, sp.nvel);
                } else if (sp.solswitch == MapScheme::APIC) {
                    sp.apic.VarBmatG2P(pid, nid, sfi, sp.apic.vel_Bmat, sp.nvel, sp.coord);
                    sp.apic.DmatG2P(pid, nid, sfi, sp.coord, ADp);
                }
            }
            if (sp.solswitch == MapScheme::APIC) sp.apic.InvDmatG2P(pid, ADp);
        }
    // MUSL: momentum of the new particle velocities, node velocity, BCs
    VectorAssign(nodec * 3, sp.nmome);
    for (int m = 0; m < nelem; m++)
        for (int pid = sp.idepf[m]; pid != -1; pid = sp.idp2p[pid]) {
            MakSf(m, sp.coord[pid], idimc, xynodec, ncm, nenode, sf, dsf);
            for (int ni = 0; ni < nenode; ni++) {
                const int nid = ncm[ni];
                const double sfi = sf[ni];
                if (sp.solswitch == MapScheme::TPIC) {
                    sp.tpic.VarP2G(pid, nid, sfi, sp.tpic.vel_grad, sp.mass, sp.vel, sp.coord, sp.nmome);
                } else if (sp.solswitch == MapScheme::APIC) {
                    sp.apic.VarP2G(pid, nid, sfi, sp.apic.vel_Bmat, sp.mass, sp.vel, sp.coord, sp.nmome);
                } else {
                    sp.nmome[nid + nuc] += sfi * sp.mass[pid] * sp.vel[pid][0];
                    sp.nmome[nid + nvc] += sfi * sp.mass[pid] * sp.vel[pid][1];
                    sp.nmome[nid + nwc] += sfi * sp.mass[pid] * sp.vel[pid][2];
                }
            }
        }
    VectorAssign(nodec * 3, sp.nvel);
    for (int n = 0; n < nodec; n++)
        if (sp.nmass[n] > mtol)
            for (int off : {nuc, nvc, nwc}) sp.nvel[n + off] = sp.nmome[n + off] / sp.nmass[n];
    sp.usbc.BCSetVal(nuc, sp.nvel);
    sp.vsbc.BCSetVal(nvc, sp.nvel);
    sp.wsbc.BCSetVal(nwc, sp.nvel);
    for (int off : {nuc, nvc, nwc}) sp.rigid_bc.BCSetVal(off, sp.nvel);
    // ComputeDeltaDefGrad (af = 1), F = dF F, J = DetMat3(F)
    for (int m = 0; m < nelem; m++)
        for (int pid = sp.idepf[m]; pid != -1; pid = sp.idp2p[pid]) {
            MakSf(m, sp.coord[pid], idimc, xynodec, ncm, nenode, sf, dsf);
            Mat3 ddg{};
            for (int ni = 0; ni < nenode; ni++) {
                const int nid = ncm[ni];
                const double v[3] = {sp.nvel[nid + nuc], sp.nvel[nid + nvc], sp.nvel[nid + nwc]};
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) ddg[i][j] += dsf[ni][j] * v[i] * dt;
            }
            for (int i = 0; i < 3; i++) ddg[i][i] += 1.0e0;
            Mat3 F{};
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++)
                    for (int k = 0; k < 3; k++) F[i][j] += ddg[i][k] * sp.def_grad[pid][k][j];
            sp.def_grad[pid] = F;
            sp.det_def_grad[pid] = DetMat3(F);
        }
    // coordinates with the weights at the old position, V = J V0
    for (int m = 0; m < nelem; m++)
        for (int pid = sp.idepf[m]; pid != -1; pid = sp.idp2p[pid]) {
            const std::array<double, 3> xyp = sp.coord[pid];
            MakSf(m, xyp, idimc, xynodec, ncm, nenode, sf, dsf);
            for (int ni = 0; ni < nenode; ni++) {
                const int nid = ncm[ni];
                sp.coord[pid][0] += sf[ni] * dt * sp.nvel[nid + nuc];
                sp.coord[pid][1] += sf[ni] * dt * sp.nvel[nid + nvc];
                sp.coord[pid][2] += sf[ni] * dt * sp.nvel[nid + nwc];
            }
            sp.vol[pid] = sp.det_def_grad[pid] * sp.vol0[pid];
        }
}

template <typename A> static std::vector<double> flat(const std::vector<A> &v) {
    std::vector<double> out;
    for (const auto &a : v) {
        const double *p = reinterpret_cast<const double *>(&a);
        out.insert(out.end(), p, p + sizeof(A) / sizeof(double));
    }
    return out;
}

static void check_maksf(const char *tag) {
    orc_state st;
    oracle_grid(st);
    int nenode = 0;
    std::vector<int> ncm;
    std::vector<double> sf;
    std::vector<std::array<double, 3>> dsf;
    int32_t onc[27];
    double osf[27], odsf[81];
    int bad = 0;
    for (int m = 0; m < nelem; m++) {
        const int c[3] = {m % xyelem[0], (m / xyelem[0]) % xyelem[1], m / (xyelem[0] * xyelem[1])};
        for (int r = 0; r < 6; r++) {
            std::array<double, 3> x;
            for (int a = 0; a < 3; a++) x[a] = xymin[a] + dxy[a] * (c[a] + (r == 0 ? 0.0 : urand()));
            MakSf(m, x, idimc, xynodec, ncm, nenode, sf, dsf);
            const int on = orc_maksf(&st, m, x.data(), onc, osf, odsf);
            if (on != nenode || (int)ncm.size() != on) {
                bad++;
                continue;
            }
            for (int i = 0; i < on; i++)
                if (ncm[i] != onc[i] || std::memcmp(&sf[i], &osf[i], 8) || std::memcmp(dsf[i].data(), odsf + 3 * i, 24))
                    bad++;
        }
    }
    std::printf("%s MakSf: %d mismatches\n", tag, bad);
    CHECK(bad == 0, "MakSf bit-identical to the oracle");
}

static void check_nodal_volume(const char *tag) {
    orc_state st;
    oracle_grid(st);
    std::vector<double> onvol(nodec, 0.0);
    st.nvol = onvol.data();
    const int32_t ppc[3] = {2, 3, 2};
    for (int i = 0; i < 3; i++) npxye[i] = ppc[i];
    orc_nodal_vol(&st, ppc);
    MakNodalVol();
    std::printf("%s MakNodalVol: %s\n", tag, same(nvol.data(), onvol.data(), nodec) ? "equal" : "DIFFERENT");
    CHECK(same(nvol.data(), onvol.data(), nodec), "MakNodalVol bit-identical to the oracle");
}

static void check_step(MapScheme scheme, const char *tag) {
    Points sp = make_points(scheme);
    link_list(sp);
    orc_state st;
    oracle_grid(st);
    st.dt = dt;
    st.mtol = mtol;
    for (int i = 0; i < 3; i++) st.bb[i] = bb[i];
    st.scheme = (int)scheme;
    st.nlstep = 1;
    st.nmat = 1;
    st.iprop[0] = ORC_LINELAST;
    st.mat_prop[0][0] = 1e6;
    st.mat_prop[0][1] = 0.3;
    std::vector<double> coord = flat(sp.coord), vel = flat(sp.vel), stress = flat(sp.stress), F = flat(sp.def_grad),
                        trac = flat(sp.trac_force), mass = sp.mass, vol = sp.vol, vol0 = sp.vol0, J = sp.det_def_grad;
    std::vector<double> cmat = flat(scheme == MapScheme::APIC ? sp.apic.vel_Bmat : sp.tpic.vel_grad),
                        dinv = flat(sp.apic.inv_Dmat);
    std::vector<int32_t> idepf(nelem), idp2p(sp.num), numep(nelem);
    std::vector<double> nm(nodec), nvf(nodec), np(3 * nodec), nf(3 * nodec), nv(3 * nodec), na(3 * nodec),
        ddg(9 * sp.num), eps(sp.num, 0.0);
    st.num = sp.num;
    st.coord = coord.data();
    st.vel = vel.data();
    st.mass = mass.data();
    st.vol = vol.data();
    st.vol0 = vol0.data();
    st.stress = stress.data();
    st.def_grad = F.data();
    st.det_def_grad = J.data();
    st.trac_force = trac.data();
    st.eps_p = eps.data();
    st.idepf = idepf.data();
    st.idp2p = idp2p.data();
    st.numep = numep.data();
    st.nmass = nm.data();
    st.nvof = nvf.data();
    st.nmome = np.data();
    st.nforce = nf.data();
    st.nvel = nv.data();
    st.naccel = na.data();
    st.ddg = ddg.data();
    st.cmat = cmat.data();
    st.dinv = dinv.data();
    std::vector<int32_t> lists[4];
    std::vector<double> vals[4];
    BoundaryCondition *bcs[4] = {&sp.usbc, &sp.vsbc, &sp.wsbc, &sp.rigid_bc};
    for (int l = 0; l < 4; l++) {
        lists[l].assign(bcs[l]->nbc.begin(), bcs[l]->nbc.end());
        vals[l] = bcs[l]->fbc;
        st.ibc[l] = bcs[l]->ibc;
        st.nbc[l] = lists[l].data();
        st.fbc[l] = vals[l].data();
    }
    CHECK(orc_bin(&st) == 0, "oracle binning");
    CHECK(std::equal(sp.idepf.begin(), sp.idepf.end(), idepf.begin()) &&
              std::equal(sp.idp2p.begin(), sp.idp2p.end(), idp2p.begin()),
          "linked lists equal the oracle's");

    particle2node(sp);
    orc_p2g(&st, facl);
    const bool p2g_ok = same(sp.nmass.data(), nm.data(), nodec) && same(sp.nvof.data(), nvf.data(), nodec) &&
                        same(sp.nmome.data(), np.data(), 3 * nodec) && same(sp.nforce.data(), nf.data(), 3 * nodec) &&
                        same(sp.nvel.data(), nv.data(), 3 * nodec);
    CHECK(p2g_ok, "Particle2NodeS node fields");
    solve_solid(sp);
    orc_solve_solid(&st);
    CHECK(same(sp.naccel.data(), na.data(), 3 * nodec), "SolveSolid accelerations");
    node2particle(sp);
    orc_g2p(&st, facl);
    const bool vel_ok = same(flat(sp.vel).data(), vel.data(), 3 * sp.num);
    bool mat_ok = true;
    if (scheme == MapScheme::TPIC) mat_ok = same(flat(sp.tpic.vel_grad).data(), cmat.data(), 9 * sp.num);
    if (scheme == MapScheme::APIC)
        mat_ok = same(flat(sp.apic.vel_Bmat).data(), cmat.data(), 9 * sp.num) &&
                 same(flat(sp.apic.inv_Dmat).data(), dinv.data(), 9 * sp.num);
    const bool musl_ok = same(sp.nvel.data(), nv.data(), 3 * nodec);
    const bool f_ok = same(flat(sp.def_grad).data(), F.data(), 9 * sp.num) &&
                      same(sp.det_def_grad.data(), J.data(), sp.num);
    const bool x_ok = same(flat(sp.coord).data(), coord.data(), 3 * sp.num) && same(sp.vol.data(), vol.data(), sp.num);
    std::printf("%s %d points: P2G %s, SolveSolid %s, G2P vel %s, matrices %s, MUSL %s, F/J %s, x/V %s\n", tag,
                sp.num, p2g_ok ? "=" : "DIFF", same(sp.naccel.data(), na.data(), 3 * nodec) ? "=" : "DIFF",
                vel_ok ? "=" : "DIFF", mat_ok ? "=" : "DIFF", musl_ok ? "=" : "DIFF", f_ok ? "=" : "DIFF",
                x_ok ? "=" : "DIFF");
    CHECK(vel_ok, "G2P particle velocities");
    CHECK(mat_ok, "TPIC / APIC particle matrices");
    CHECK(musl_ok, "MUSL node velocities");
    CHECK(f_ok, "deformation gradient and det");
    CHECK(x_ok, "coordinates and volume");
}

static void write_raw(const std::string &path, const std::vector<double> &v) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) {
        g_fail++;
        return;
    }
    std::fwrite(v.data(), sizeof(double), v.size(), f);
    std::fclose(f);
}

// Inputs and outputs of the helpers that the reference's own code also
// provides (checked against oracle/_ref by tests/test_ref_shim.py).
static void dump(const std::string &dir) {
    const int nm = 200;
    std::vector<double> mats, det, inv, tr;
    for (int k = 0; k < nm; k++) {
        Mat3 a;
        for (auto &r : a)
            for (double &v : r) v = 2.0 * urand() - 1.0;
        for (int i = 0; i < 3; i++) a[i][i] += 1.5;
        for (auto &r : a) mats.insert(mats.end(), r.begin(), r.end());
        det.push_back(DetMat3(a));
        const Mat3 b = InvMat3(a);
        for (auto &r : b) inv.insert(inv.end(), r.begin(), r.end());
        tr.push_back(TraceMat3(a));
    }
    write_raw(dir + "/mats.bin", mats);
    write_raw(dir + "/det.bin", det);
    write_raw(dir + "/inv.bin", inv);
    write_raw(dir + "/trace.bin", tr);
    // VarG2P over random (particle, node, weight) triples on 50 particles and
    // a [u|v|w] node vector of 40 nodes
    const int np = 50, nn = 40, npair = 600;
    nuc = 0;
    nvc = nn;
    nwc = 2 * nn;
    dt = 3.7e-4;
    std::vector<double> nodev(3 * nn), pairs;
    for (double &v : nodev) v = urand() - 0.5;
    PIC pic;
    FLIP flip;
    std::vector<std::array<double, 3>> ppic(np), pflip(np);
    for (int k = 0; k < np; k++)
        for (int c = 0; c < 3; c++) ppic[k][c] = pflip[k][c] = urand();
    const std::vector<double> p0 = flat(ppic);
    for (int k = 0; k < npair; k++) {
        const int pid = (int)(urand() * np), nid = (int)(urand() * nn);
        const double w = urand();
        pairs.insert(pairs.end(), {(double)pid, (double)nid, w});
        pic.VarG2P(pid, nid, w, ppic, nodev);
        flip.VarG2P(pid, nid, w, pflip, nodev);
    }
    write_raw(dir + "/g2p_nodes.bin", nodev);
    write_raw(dir + "/g2p_pairs.bin", pairs);
    write_raw(dir + "/g2p_p0.bin", p0);
    write_raw(dir + "/g2p_pic.bin", flat(ppic));
    write_raw(dir + "/g2p_flip.bin", flat(pflip));
    // BCSetVal / BCSetZero / BCSetDt with repeated node ids (the later entry
    // wins)
    BoundaryCondition bc;
    std::vector<double> bcin;
    for (int i = 0; i < 30; i++) {
        bc.nbc.push_back((int)(urand() * 20));
        bc.fbc.push_back(urand() - 0.5);
    }
    bc.ibc = 30;
    facl = 0.625;
    std::vector<double> var(3 * 20), out;
    for (double &v : var) v = urand();
    for (int i = 0; i < 30; i++) bcin.insert(bcin.end(), {(double)bc.nbc[i], bc.fbc[i]});
    for (int op = 0; op < 3; op++)
        for (int nn2 : {0, 20, 40}) {
            std::vector<double> v = var;
            if (op == 0) bc.BCSetVal(nn2, v);
            else if (op == 1) bc.BCSetZero(nn2, v);
            else bc.BCSetDt(nn2, v);
            out.insert(out.end(), v.begin(), v.end());
        }
    write_raw(dir + "/bc_lists.bin", bcin);
    write_raw(dir + "/bc_var.bin", var);
    write_raw(dir + "/bc_out.bin", out);
    // BCOutput then BCInput: the same lists back
    {
        std::ofstream o(dir + "/bc.txt");
        bc.BCOutput(o, "usbc");
    }
    std::ifstream in(dir + "/bc.txt");
    BoundaryCondition back;
    back.BCInput(in);
    CHECK(back.ibc == bc.ibc && back.nbc == bc.nbc, "BCOutput / BCInput round trip");
}

int main(int argc, char **argv) {
    dt = 1.0e-4;
    mtol = 1.0e-12;
    facl = 1.0;
    bb[0] = 0.0;
    bb[1] = -9.81;
    bb[2] = 0.5;
    const double lo[3] = {0.0, -0.5, 1.0}, hi[3] = {1.2, 0.5, 2.4};
    const int ne[3] = {6, 5, 7};
    const int orders[3][3] = {{1, 1, 1}, {2, 2, 2}, {2, 1, 2}};
    const MapScheme schemes[4] = {MapScheme::PIC, MapScheme::FLIP, MapScheme::TPIC, MapScheme::APIC};
    const char *names[4] = {"PIC", "FLIP", "TPIC", "APIC"};
    for (const auto &o : orders) {
        setup_grid(lo, hi, ne, o);
        char tag[64];
        std::snprintf(tag, sizeof tag, "idimc %d%d%d", o[0], o[1], o[2]);
        check_maksf(tag);
        check_nodal_volume(tag);
        for (int s = 0; s < 4; s++) {
            std::snprintf(tag, sizeof tag, "idimc %d%d%d %s", o[0], o[1], o[2], names[s]);
            check_step(schemes[s], tag);
        }
    }
    CHECK(ParseMapScheme("APIC") == MapScheme::APIC && ParseMapScheme("FLIP") == MapScheme::FLIP,
          "ParseMapScheme");
    if (argc > 2 && std::string(argv[1]) == "--dump") dump(argv[2]);
    std::printf(g_fail ? "FAILED (%d)\n" : "OK\n", g_fail);
    return g_fail ? 1 : 0;
}

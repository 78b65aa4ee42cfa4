// C++17 host interface of the explicit solid MPM step, mirroring the
// reference's pass functions over its own containers.
//
// The reference drives one explicit step as (work/src_solid/explicit/main.cpp:55-61)
//     sp.MeshPointLinklist();   // module/material_point.cpp:52-77
//     Particle2NodeS();         // module/solid/explicit/var_trans_solid_explicit.cpp:20-99
//     solve_linear_system();    // SolveSolid, module/solid/explicit/solve_solid_explicit.cpp:8-31
//     Node2ParticleS();         // var_trans_solid_explicit.cpp:187-311 (with MUSL :101-148)
// over the global point container `sp` (module/material_point.h:113-153) and
// the grid globals (module/mesh.h, module/dataset.h). ExplicitSolid keeps the
// same pass names and argument meaning; the state lives on the GPU between
// passes and is copied from/to the caller's containers only by
// UploadPoints / DownloadPoints / DownloadNodes / DownloadLinklist.
//
// Errors: the reference has no checks on this path (an out-of-grid particle
// is undefined behaviour there). Here every call checks the engine's status
// and throws mpm::Error (code = mpm_status, message = mpm_last_error()).
#pragma once

#include <algorithm>
#include <array>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "mpm_capi.h"
#include "mpm_surrogate.h"

namespace mpm {

class Error : public std::runtime_error {
  public:
    Error(int code, const std::string &msg) : std::runtime_error(msg), code(code) {}
    int code;
};

inline void check(int rc) {
    if (rc != MPM_OK) {
        const char *m = mpm_last_error();
        throw Error(rc, "mpm error " + std::to_string(rc) + ": " + (m ? m : ""));
    }
}

struct Grid {
    std::array<double, 3> xmin{};  // xyminw
    std::array<double, 3> xmax{};  // xymaxw
    std::array<int, 3> nelem{};    // xyelemw
    std::array<int, 3> idimc{{1, 1, 1}};
};

struct Material {
    int model = MPM_MODEL_LINELAST;  // iprop
    std::vector<double> prop;        // mat_prop row
};

struct StepConfig {
    double dt = 0.0;
    double mtol = 1e-12;
    std::array<double, 3> bb{};  // body force per unit mass
    int scheme = MPM_SCHEME_PIC; // sp.solswitch
    int nlstep = 1;
    int precision = MPM_PREC_FP32;
    int relayout_every = 32;
};

namespace detail {
template <class V>
double *dptr(V &v) {
    using E = typename V::value_type;
    static_assert(sizeof(E) % sizeof(double) == 0, "AoS rows of doubles expected");
    return v.empty() ? nullptr : reinterpret_cast<double *>(v.data());
}
}  // namespace detail

class ExplicitSolid {
  public:
    ExplicitSolid(const Grid &g, const std::vector<Material> &mats, const StepConfig &c) {
        mpm_grid_desc gd{};
        for (int a = 0; a < 3; a++) {
            gd.xmin[a] = g.xmin[a];
            gd.xmax[a] = g.xmax[a];
            gd.nelem[a] = g.nelem[a];
            gd.idimc[a] = g.idimc[a];
        }
        std::vector<mpm_material_desc> md(mats.size());
        for (size_t i = 0; i < mats.size(); i++) {
            md[i].model = mats[i].model;
            md[i].nprop = (int32_t)std::min<size_t>(mats[i].prop.size(), 8);
            for (int k = 0; k < md[i].nprop; k++) md[i].prop[k] = mats[i].prop[k];
        }
        mpm_step_params sp{};
        sp.dt = c.dt;
        sp.mtol = c.mtol;
        for (int a = 0; a < 3; a++) sp.bb[a] = c.bb[a];
        sp.scheme = c.scheme;
        sp.nlstep = c.nlstep;
        sp.precision = c.precision;
        sp.relayout_every = c.relayout_every;
        nodec_ = 1;
        for (int a = 0; a < 3; a++) nodec_ *= (int64_t)(g.nelem[a] + g.idimc[a]);
        nelem_ = (int64_t)g.nelem[0] * g.nelem[1] * g.nelem[2];
        check(mpm_engine_create(&gd, md.data(), (int32_t)md.size(), &sp, &h_));
    }
    ~ExplicitSolid() {
        if (h_) mpm_engine_destroy(h_);
    }
    ExplicitSolid(const ExplicitSolid &) = delete;
    ExplicitSolid &operator=(const ExplicitSolid &) = delete;
    ExplicitSolid(ExplicitSolid &&o) noexcept : h_(std::exchange(o.h_, nullptr)), n_(o.n_), nodec_(o.nodec_), nelem_(o.nelem_) {}

    // comp: 0 = ubc, 1 = vbc, 2 = wbc, 3 = rigid_bc (module/material_point.h:28-32);
    // BC is any type with ibc, nbc, fbc (BoundaryCondition, module/bc.h:7-22)
    template <class BC>
    void SetBC(int comp, const BC &bc) {
        std::vector<double> f(bc.fbc.begin(), bc.fbc.end());
        f.resize(bc.ibc, 0.0);
        check(mpm_set_bc(h_, comp, bc.nbc.data(), f.data(), bc.ibc));
    }

    // SP: the reference's point container (coord, vel, mass, vol, vol0,
    // stress, def_grad, det_def_grad, trac_force, matid; num = count)
    template <class SP>
    void UploadPoints(SP &sp) {
        mpm_particles_view v = view(sp);
        check(mpm_upload_particles(h_, &v));
        n_ = v.n;
    }
    template <class SP>
    void DownloadPoints(SP &sp) {
        mpm_particles_view v = view(sp);
        check(mpm_download_particles(h_, &v));
    }
    // nmass, nvof [nodec]; nmome, nforce, nvel, naccel [3*nodec] ([u|v|w] blocks)
    template <class SP>
    void DownloadNodes(SP &sp) {
        sp.nmass.assign(nodec_, 0.0);
        sp.nvof.assign(nodec_, 0.0);
        sp.nmome.assign(3 * nodec_, 0.0);
        sp.nforce.assign(3 * nodec_, 0.0);
        sp.nvel.assign(3 * nodec_, 0.0);
        sp.naccel.assign(3 * nodec_, 0.0);
        mpm_nodes_view nv{nodec_, sp.nmass.data(), sp.nvof.data(), sp.nmome.data(),
                          sp.nforce.data(), sp.nvel.data(), sp.naccel.data()};
        check(mpm_download_nodes(h_, &nv));
    }
    // idepf / numep per cell, idp2p per particle (linked lists of
    // module/material_point.cpp:66-74)
    template <class SP>
    void DownloadLinklist(SP &sp) {
        sp.idepf.assign(nelem_, -1);
        sp.numep.assign(nelem_, 0);
        sp.idp2p.assign(n_, -1);
        check(mpm_download_bins(h_, sp.idepf.data(), sp.idp2p.data(), sp.numep.data(), nullptr));
    }

    void MeshPointLinklist() { check(mpm_bin(h_)); }
    void Particle2NodeS(double facl) { check(mpm_p2g(h_, facl)); }
    void SolveSolid() { check(mpm_solve_grid(h_)); }
    void Node2ParticleS(double facl) { check(mpm_g2p(h_, facl)); }
    // istep0 .. istep0 + nsteps - 1, facl from nlstep as main.cpp:43-48
    void Step(int istep0, int nsteps) { check(mpm_step(h_, istep0, nsteps)); }

    // Implicit path (FP64-exact engine, PIC/FLIP): the transfers of
    // SolidMaterialPoint::Particle2Node / Node2Particle
    // (module/solid/implicit/var_trans_solid_implicit.cpp:20-80, :106-167)
    // around the reference's own implicit solve, which supplies the converged
    // nodal displacement increment ndispl ([u|v|w] blocks of nodec). nb_para
    // and gamma_nb are the reference's Newmark values (NewmarkBetaParaSet,
    // module/nonlinear_time_integration.cpp:43-57). SP::accel is the point
    // container's vector<array<double,3>> of accelerations.
    template <class SP>
    void UploadAccel(SP &sp) {
        check(mpm_upload_accel(h_, detail::dptr(sp.accel)));
    }
    template <class SP>
    void DownloadAccel(SP &sp) {
        sp.accel.resize((size_t)n_);
        check(mpm_download_accel(h_, detail::dptr(sp.accel)));
    }
    void Particle2Node(double facl) { check(mpm_implicit_p2g(h_, facl)); }
    void Node2Particle(const std::vector<double> &ndispl, const std::vector<double> &nb_para, double gamma_nb) {
        if ((int64_t)ndispl.size() != 3 * nodec_ || nb_para.size() < 6)
            throw Error(MPM_ERR_INVALID_ARGUMENT, "ndispl needs 3 * nodec values and nb_para 6");
        check(mpm_implicit_g2p(h_, ndispl.data(), nb_para.data(), gamma_nb));
    }
    void Sync() { check(mpm_sync(h_)); }

    mpm_engine *handle() { return h_; }
    int64_t nodec() const { return nodec_; }

  private:
    template <class SP>
    mpm_particles_view view(SP &sp) {
        mpm_particles_view v{};
        v.n = (int64_t)sp.coord.size();
        v.coord = detail::dptr(sp.coord);
        v.vel = detail::dptr(sp.vel);
        v.mass = sp.mass.empty() ? nullptr : sp.mass.data();
        v.vol = sp.vol.empty() ? nullptr : sp.vol.data();
        v.vol0 = sp.vol0.empty() ? nullptr : sp.vol0.data();
        v.stress = detail::dptr(sp.stress);
        v.def_grad = detail::dptr(sp.def_grad);
        v.det_def_grad = sp.det_def_grad.empty() ? nullptr : sp.det_def_grad.data();
        v.trac_force = detail::dptr(sp.trac_force);
        v.matid = sp.matid.empty() ? nullptr : reinterpret_cast<int32_t *>(sp.matid.data());
        static_assert(sizeof(typename decltype(sp.matid)::value_type) == sizeof(int32_t), "int matid expected");
        v.eps_p = nullptr;
        return v;
    }

    mpm_engine *h_ = nullptr;
    int64_t n_ = 0, nodec_ = 0, nelem_ = 0;
};

}  // namespace mpm

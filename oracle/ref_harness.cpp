// Test harness around the reference's own numerics, compiled from the sources
// where they lie under the reference tree (see oracle/Makefile). TEST
// INFRASTRUCTURE: only tests load the resulting oracle/_ref/libmpm_ref.so, to
// check the FP64 oracle bit for bit. It includes the reference headers
// module/cal_mat.h, module/bc.h, module/map_and_interpolate.h and
// module/solid/constitutive_model.h and links module/bc.cpp and
// module/solid/constitutive_model.cpp; nothing is stubbed.
#include "bc.h"
#include "cal_mat.h"
#include "dataset.h"
#include "map_and_interpolate.h"
#include "mesh.h"
#include "solid/constitutive_model.h"

#include <array>
#include <vector>

extern "C" {

int ref_update_stress(int model, double *stress6, const double *F9, const double *dF9, const double *prop,
                      int nprop, double jac, double jac_bar) {
    ConstitutiveModel cm;
    cm.implicit_flag = false;
    std::array<double, 6> s;
    std::array<std::array<double, 3>, 3> F, dF;
    for (int i = 0; i < 6; i++) s[i] = stress6[i];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            F[i][j] = F9[3 * i + j];
            dF[i][j] = dF9[3 * i + j];
        }
    std::vector<double> p(prop, prop + nprop);
    cm.UpdateStress(model, s, F, dF, p, jac, jac_bar);
    for (int i = 0; i < 6; i++) stress6[i] = s[i];
    return 0;
}

static BoundaryCondition make_bc(int ibc, const int *nbc, const double *fbc) {
    BoundaryCondition bc;
    bc.ibc = ibc;
    bc.nbc.assign(nbc, nbc + ibc);
    if (fbc) bc.fbc.assign(fbc, fbc + ibc);
    return bc;
}

int ref_bc_apply(int op, int ibc, const int *nbc, const double *fbc, int nn, double *var, int nvar, double facl_in,
                 double dt_in) {
    BoundaryCondition bc = make_bc(ibc, nbc, fbc);
    std::vector<double> v(var, var + nvar);
    facl = facl_in;
    dt = dt_in;
    if (op == 0) bc.BCSetVal(nn, v);
    else if (op == 1) bc.BCSetZero(nn, v);
    else bc.BCSetDt(nn, v);
    for (int i = 0; i < nvar; i++) var[i] = v[i];
    return 0;
}

double ref_det3(const double *a9) {
    std::array<std::array<double, 3>, 3> a;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) a[i][j] = a9[3 * i + j];
    return DetMat3(a);
}

void ref_inv3(const double *a9, double *o9) {
    std::array<std::array<double, 3>, 3> a;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) a[i][j] = a9[3 * i + j];
    auto o = InvMat3(a);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) o9[3 * i + j] = o[i][j];
}

// PIC::VarG2P / FLIP::VarG2P (map_and_interpolate.h:12-55) on a particle
// vector of n points and a [u|v|w] node vector of 3*nodec entries.
int ref_var_g2p(int scheme, int n, double *pvar3, int nodec, const double *node3, int npairs, const int *pid,
                const int *nid, const double *sfi, double dt_in) {
    std::vector<std::array<double, 3>> pv(n);
    for (int i = 0; i < n; i++)
        for (int c = 0; c < 3; c++) pv[i][c] = pvar3[3 * i + c];
    std::vector<double> nv(node3, node3 + 3 * nodec);
    nuc = 0;
    nvc = nodec;
    nwc = 2 * nodec;
    dt = dt_in;
    FLIP flip;
    PIC pic;
    for (int k = 0; k < npairs; k++) {
        if (scheme == 1) flip.VarG2P(pid[k], nid[k], sfi[k], pv, nv);
        else pic.VarG2P(pid[k], nid[k], sfi[k], pv, nv);
    }
    for (int i = 0; i < n; i++)
        for (int c = 0; c < 3; c++) pvar3[3 * i + c] = pv[i][c];
    return 0;
}

int ref_trace3(const double *a9, double *out) {
    std::array<std::array<double, 3>, 3> a;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) a[i][j] = a9[3 * i + j];
    *out = TraceMat3(a);
    return 0;
}
}

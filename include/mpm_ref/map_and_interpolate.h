// Reference-signature host shim: particle <-> node transfer classes.
//
// FLIP, PIC, TPIC and APIC with the member templates and data members of the
// reference's module/map_and_interpolate.h:12-176. Node vectors are
// [u | v | w] component blocks at the offsets nuc, nvc, nwc (mesh.h); dt is
// the global of dataset.h. Each call adds one stencil node's contribution in
// the reference's operation order.
//
// TPIC/APIC P2G read the control-point coordinate as xyc[nid][j] (xyc is a
// vector of 3-arrays indexed by node); the reference writes xyc[j][nid]
// (map_and_interpolate.h:69, :123), which indexes the node vector by the
// axis and reads out of bounds for j beyond the third node.
#pragma once

#include <array>
#include <vector>

#include "cal_mat.h"
#include "dataset.h"
#include "mesh.h"

namespace mpm_ref_detail {
// Velocity v_p + M (x_a - x_p) that particle pid carries to node nid, for a
// 3x3 matrix M (TPIC velocity gradient, APIC B D^-1); the sum over j runs
// from zero as in the reference.
template <typename T>
inline std::array<double, 3> affine_velocity(const std::array<std::array<double, 3>, 3> &M,
                                             const std::array<T, 3> &xp, int nid) {
    std::array<double, 3> out{};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out[i] += M[i][j] * (xyc[nid][j] - xp[j]);
    return out;
}

template <typename T>
inline void add_momentum(std::vector<T> &node_var, int nid, double w, const std::array<T, 3> &v,
                         const std::array<double, 3> &dv) {
    node_var[nid + nuc] += w * (v[0] + dv[0]);
    node_var[nid + nvc] += w * (v[1] + dv[1]);
    node_var[nid + nwc] += w * (v[2] + dv[2]);
}
}  // namespace mpm_ref_detail

class FLIP {
  public:
    // explicit FLIP: the particle value moves by dt times the node rate
    template <typename T>
    void VarG2P(int pid, int nid, double sfi, std::vector<T> &point_var, const std::vector<T> &node_var) {
        point_var[pid] += sfi * dt * node_var[nid];
    }

    template <typename T>
    void VarG2P(int pid, int nid, double sfi, std::vector<std::array<T, 3>> &point_var,
                const std::vector<T> &node_var) {
        const int off[3] = {nuc, nvc, nwc};
        for (int c = 0; c < 3; c++) point_var[pid][c] += sfi * dt * node_var[nid + off[c]];
    }
};

class PIC {
  public:
    template <typename T>
    void VarG2P(int pid, int nid, double sfi, std::vector<T> &point_var, const std::vector<T> &node_var) {
        point_var[pid] += sfi * node_var[nid];
    }

    template <typename T>
    void VarG2P(int pid, int nid, double sfi, std::vector<std::array<T, 3>> &point_var,
                const std::vector<T> &node_var) {
        const int off[3] = {nuc, nvc, nwc};
        for (int c = 0; c < 3; c++) point_var[pid][c] += sfi * node_var[nid + off[c]];
    }
};

class TPIC {
  public:
    std::vector<std::array<std::array<double, 3>, 3>> vel_grad, accel_grad;

    // node momentum of the Taylor-expanded particle velocity v_p + L (x_a - x_p)
    template <typename T>
    void VarP2G(int pid, int nid, double sfi, std::vector<std::array<std::array<T, 3>, 3>> &point_var_grad,
                const std::vector<T> &point_mass, const std::vector<std::array<T, 3>> &point_var,
                const std::vector<std::array<T, 3>> &point_coord, std::vector<T> &node_var) {
        const std::array<double, 3> dv =
            mpm_ref_detail::affine_velocity(point_var_grad[pid], point_coord[pid], nid);
        mpm_ref_detail::add_momentum(node_var, nid, sfi * point_mass[pid], point_var[pid], dv);
    }

    // particle gradient L_ij += dN_j v_i
    template <typename T>
    void VarGradG2P(int pid, int nid, int ni, const std::vector<std::array<T, 3>> &dsf,
                    std::vector<std::array<std::array<T, 3>, 3>> &point_var_grad, const std::vector<T> &node_var) {
        const int off[3] = {nuc, nvc, nwc};
        const double d[3] = {dsf[ni][0], dsf[ni][1], dsf[ni][2]};
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) point_var_grad[pid][i][j] += d[j] * node_var[nid + off[i]];
    }
};

class APIC {
  public:
    std::vector<std::array<std::array<double, 3>, 3>> vel_Bmat, accel_Bmat, inv_Dmat;

    // node momentum of v_p + B D^-1 (x_a - x_p)
    template <typename T>
    void VarP2G(int pid, int nid, double sfi, std::vector<std::array<std::array<T, 3>, 3>> &point_var_Bmat,
                const std::vector<T> &point_mass, const std::vector<std::array<T, 3>> &point_var,
                const std::vector<std::array<T, 3>> &point_coord, std::vector<T> &node_var) {
        std::array<std::array<double, 3>, 3> BD{};
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++) BD[i][j] += point_var_Bmat[pid][i][k] * this->inv_Dmat[pid][k][j];
        const std::array<double, 3> dv = mpm_ref_detail::affine_velocity(BD, point_coord[pid], nid);
        mpm_ref_detail::add_momentum(node_var, nid, sfi * point_mass[pid], point_var[pid], dv);
    }

    // B_ij += N v_i (x_a - x_p)_j
    template <typename T>
    void VarBmatG2P(int pid, int nid, double sfi, std::vector<std::array<std::array<T, 3>, 3>> &point_var_Bmat,
                    const std::vector<T> &node_var, const std::vector<std::array<T, 3>> &point_coord) {
        const int off[3] = {nuc, nvc, nwc};
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                point_var_Bmat[pid][i][j] += sfi * node_var[nid + off[i]] * (xyc[nid][j] - point_coord[pid][j]);
    }

    // D_ij += N (x_a - x_p)_i (x_a - x_p)_j
    template <typename T>
    void DmatG2P(int pid, int nid, double sfi, const std::vector<std::array<T, 3>> &point_coord,
                 std::array<std::array<T, 3>, 3> &ADp) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                ADp[i][j] += sfi * (xyc[nid][i] - point_coord[pid][i]) * (xyc[nid][j] - point_coord[pid][j]);
    }

    template <typename T> void InvDmatG2P(int pid, const std::array<std::array<T, 3>, 3> &ADp) {
        this->inv_Dmat[pid] = InvMat3(ADp);
    }
};

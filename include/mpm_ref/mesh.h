// Reference-signature host shim: mesh and control-point globals.
//
// Same names, types and meaning as the reference's module/mesh.h:7-21, so that
// code written against the reference (callers of MakSf, the transfer
// classes and BoundaryCondition) compiles unchanged against this directory.
// C++17 inline variables: one instance per process, shared by the shim
// library (libmpm_refshim.so) and the caller. Single rank: the local grid is
// the whole grid (xymin = xyminw, aelemmin = 0, xynodec = xynodecw).
#pragma once

#include <array>
#include <vector>

// ----- mesh (module/mesh.h:7-15) -----
inline int node, nodew, nelem, nelemw, nu, nv, nw, np;
inline std::vector<int> xynode(3), xyelem(3), aelemmin(3), aelemmax(3);
inline std::vector<int> xynodew(3), xyelemw(3);
inline std::vector<double> dxy(3), xymin(3), xymax(3), xyminw(3), xymaxw(3);
inline std::vector<double> rhoe, rmue;
inline std::vector<std::array<int, 8>> nc;
inline std::vector<std::array<double, 3>> xyn;

// ----- control points (module/mesh.h:17-21) -----
inline int nodec, nodecw, nuc, nvc, nwc, npc;
inline std::vector<int> xynodec(3), xynodecw(3), idimc(3), idiml(3);
inline std::vector<double> nvol;
inline std::vector<std::vector<int>> ncc;
inline std::vector<std::array<double, 3>> xyc;

// Point offsets of npxye[i] points per cell along axis i from the cell
// centre (module/mesh.cpp:12-54).
void GaussianDistribution(std::array<std::array<double, 3>, 6> &dec2p);

// nvol of the completely filled grid (module/mesh.cpp:160-202), one rank.
void MakNodalVol();

// xyn and nc of the linear mesh (module/mesh.cpp:56-99).
void BuildMesh();

// ncc and the control-point coordinates xyc (module/mesh.cpp:102-158).
void BuildControlPoint();

// Coordinate of control point nid along axis dir (module/mesh.h:32-47):
// the node itself for the tent functions; for the open quadratic B-spline the
// Greville abscissae, i.e. the domain ends for the first and last control
// point and the cell centres in between.
template <int dir> void DefineCpPos(int nid, const std::vector<int> &ijk_l, const std::vector<int> &ijk_g) {
    const int order = idimc[dir];
    double x = 0.0;
    if (order == 1) {
        x = xymin[dir] + dxy[dir] * double(ijk_l[dir]);
    } else if (order == 2) {
        const int last = xynodecw[dir] - 1;
        if (ijk_g[dir] == 0)
            x = xyminw[dir];
        else if (ijk_g[dir] == last)
            x = xymaxw[dir];
        else
            x = xymin[dir] + dxy[dir] * (double(ijk_l[dir]) - 0.5e0);
    } else {
        return;
    }
    xyc[nid][dir] = x;
}

// Reference-signature host shim (CPU, FP64): MakSf, the mesh and
// control-point builders, the filled-grid nodal volume and
// BoundaryCondition, with the signatures of the reference's headers
// (include/mpm_ref/). Code written against the reference's call surface
// compiles and links against this library unchanged; whole explicit steps
// go to the GPU through the C ABI (include/mpm/mpm_capi.h) instead.
//
// Built with -ffp-contract=off: every a * b + c rounds twice, as in the
// reference's x86-64 build, so results equal the reference's arithmetic.
#include <cmath>
#include <iomanip>

#include "bc.h"
#include "cal_mat.h"
#include "dataset.h"
#include "map_and_interpolate.h"
#include "mesh.h"
#include "shape_function.h"

namespace {

// Weight and derivative along one axis for control point ixyn (local index)
// and coordinate x (module/shape_function.cpp:38-104).
void axis_weight(int ii, int order, int ixyn, double x, double &s, double &d) {
    const double h = dxy[ii];
    if (order == 1) {
        // tent function; at x == x_a the derivative takes the right-hand side
        const double dxa = x - (xymin[ii] + h * double(ixyn));
        s = 1.0e0 - std::fabs(dxa) / h;
        d = (dxa < 0.0e0) ? 1.0e0 / h : -1.0e0 / h;
        return;
    }
    // open quadratic B-spline: u = (x - x_{ixyn-2}) / h in [0, 3) over the
    // support, one quadratic piece per knot span, with the clamped pieces of
    // the first two and last two control points
    const int g = aelemmin[ii] + ixyn;
    const int last = xynodecw[ii] - 1;
    double u = (x - (xymin[ii] + h * double(ixyn - 2))) / h;
    // piece coefficients: s = (c2 u + c1) u + c0, ds/du = e1 u + e0
    double c2, c1, c0, e1, e0;
    if (g == 0) {
        u -= 2.0e0;
        c2 = 1.0e0, c1 = -2.0e0, c0 = 1.0e0, e1 = 2.0e0, e0 = -2.0e0;
    } else if (g == 1) {
        u -= 1.0e0;
        if (u <= 1.0e0)
            c2 = -1.5e0, c1 = 2.0e0, c0 = 0.0e0, e1 = -3.0e0, e0 = 2.0e0;
        else
            c2 = 0.5e0, c1 = -2.0e0, c0 = 2.0e0, e1 = 1.0e0, e0 = -2.0e0;
    } else if (g == last) {
        c2 = 1.0e0, c1 = 0.0e0, c0 = 0.0e0, e1 = 2.0e0, e0 = 0.0e0;
    } else if (g == last - 1) {
        if (u <= 1.0e0)
            c2 = 0.5e0, c1 = 0.0e0, c0 = 0.0e0, e1 = 1.0e0, e0 = 0.0e0;
        else
            c2 = -1.5e0, c1 = 4.0e0, c0 = -2.0e0, e1 = -3.0e0, e0 = 4.0e0;
    } else if (u <= 1.0e0) {
        c2 = 0.5e0, c1 = 0.0e0, c0 = 0.0e0, e1 = 1.0e0, e0 = 0.0e0;
    } else if (u <= 2.0e0) {
        c2 = -1.0e0, c1 = 3.0e0, c0 = -1.5e0, e1 = -2.0e0, e0 = 3.0e0;
    } else {
        c2 = 0.5e0, c1 = -3.0e0, c0 = 4.5e0, e1 = 1.0e0, e0 = -3.0e0;
    }
    s = (c2 * u + c1) * u + c0;
    d = (e1 * u + e0) / h;
}

}  // namespace

void MakSf(const int m, const std::array<double, 3> &xyg, const std::vector<int> &idimc,
           const std::vector<int> &xynodec, std::vector<int> &nc, int &nenode, std::vector<double> &sf,
           std::vector<std::array<double, 3>> &dsf) {
    const int nen = (idimc[0] + 1) * (idimc[1] + 1) * (idimc[2] + 1);
    const int nex = xyelem[0], ney = xyelem[1];
    const int cell[3] = {m % nex, (m / nex) % ney, m / (nex * ney)};
    nc.resize(nen);
    sf.resize(nen);
    dsf.resize(nen);
    // per-axis weights of the idimc + 1 control points of the cell
    double s[3][3], d[3][3];
    for (int a = 0; a < 3; a++)
        for (int o = 0; o <= idimc[a]; o++) axis_weight(a, idimc[a], cell[a] + o, xyg[a], s[a][o], d[a][o]);
    int id = 0;
    for (int k = 0; k <= idimc[2]; k++)
        for (int j = 0; j <= idimc[1]; j++)
            for (int i = 0; i <= idimc[0]; i++, id++) {
                nc[id] = (cell[0] + i) + xynodec[0] * (cell[1] + j) + xynodec[0] * xynodec[1] * (cell[2] + k);
                sf[id] = s[0][i] * s[1][j] * s[2][k];
                dsf[id][0] = d[0][i] * s[1][j] * s[2][k];
                dsf[id][1] = s[0][i] * d[1][j] * s[2][k];
                dsf[id][2] = s[0][i] * s[1][j] * d[2][k];
            }
    nenode = id;
}

void GaussianDistribution(std::array<std::array<double, 3>, 6> &dec2p) {
    // npx points per axis at equal spacing h / npx around the cell centre
    // (offsets (2 q + 1 - npx) / (2 npx) h, q = 0 .. npx - 1), written with
    // the reference's constants (module/mesh.cpp:12-54)
    dec2p = {};
    for (int i = 0; i < 3; i++) {
        const double h = dxy[i];
        switch (npxye[i]) {
        case 2: {
            const double o[2] = {-0.25e0, 0.25e0};
            for (int q = 0; q < 2; q++) dec2p[q][i] = o[q] * h;
            break;
        }
        case 3:
            dec2p[0][i] = -1.0e0 / 3.0e0 * h;
            dec2p[2][i] = 1.0e0 / 3.0e0 * h;
            break;
        case 4: {
            const double o[4] = {-0.375e0, -0.125e0, 0.125e0, 0.375e0};
            for (int q = 0; q < 4; q++) dec2p[q][i] = o[q] * h;
            break;
        }
        case 5: {
            const double o[5] = {-0.4e0, -0.2e0, 0.0e0, 0.2e0, 0.4e0};
            for (int q = 0; q < 5; q++) dec2p[q][i] = o[q] * h;
            break;
        }
        case 6:
            for (int q = 0; q < 6; q++) dec2p[q][i] = double(2 * q - 5) / 12.0 * h;
            break;
        default: break;  // one point: the cell centre
        }
    }
}

void BuildMesh() {
    const int nx = xynode[0], ny = xynode[1], nz = xynode[2];
    xyn.assign(node, std::array<double, 3>{});
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const int id = i + nx * j + nx * ny * k;
                xyn[id] = {dxy[0] * i + xymin[0], dxy[1] * j + xymin[1], dxy[2] * k + xymin[2]};
            }
    // hexahedron connectivity, counter-clockwise bottom face then top face
    nc.assign(nelem, std::array<int, 8>{});
    for (int k = 0; k < xyelem[2]; k++)
        for (int j = 0; j < xyelem[1]; j++)
            for (int i = 0; i < xyelem[0]; i++) {
                const int mid = i + xyelem[0] * j + xyelem[0] * xyelem[1] * k;
                const int b = nx * ny * k + nx * j + i, t = b + nx * ny;
                nc[mid] = {b, b + 1, b + nx + 1, b + nx, t, t + 1, t + nx + 1, t + nx};
            }
}

void BuildControlPoint() {
    const int nen = (idimc[0] + 1) * (idimc[1] + 1) * (idimc[2] + 1);
    const int nex = xyelem[0], ney = xyelem[1];
    ncc.assign(nelem, std::vector<int>(nen));
    for (int m = 0; m < nelem; m++) {
        const int c[3] = {m % nex, (m / nex) % ney, m / (nex * ney)};
        int id = 0;
        for (int k = 0; k <= idimc[2]; k++)
            for (int j = 0; j <= idimc[1]; j++)
                for (int i = 0; i <= idimc[0]; i++)
                    ncc[m][id++] = (c[0] + i) + xynodec[0] * (c[1] + j) + xynodec[0] * xynodec[1] * (c[2] + k);
    }
    xyc.assign(nodec, std::array<double, 3>{});
    std::vector<int> l(3), gl(3);
    for (int k = aelemmin[2]; k <= aelemmax[2] + idimc[2]; k++)
        for (int j = aelemmin[1]; j <= aelemmax[1] + idimc[1]; j++)
            for (int i = aelemmin[0]; i <= aelemmax[0] + idimc[0]; i++) {
                gl = {i, j, k};
                l = {i - aelemmin[0], j - aelemmin[1], k - aelemmin[2]};
                const int nid = l[0] + xynodec[0] * l[1] + xynodec[0] * xynodec[1] * l[2];
                DefineCpPos<0>(nid, l, gl);
                DefineCpPos<1>(nid, l, gl);
                DefineCpPos<2>(nid, l, gl);
            }
}

void MakNodalVol() {
    // every cell filled with npxye points of volume dx dy dz / (npx npy npz);
    // single rank (no node exchange)
    std::array<std::array<double, 3>, 6> off;
    GaussianDistribution(off);
    const double volp = dxy[0] * dxy[1] * dxy[2] / (npxye[0] * npxye[1] * npxye[2]);
    const int nex = xyelem[0], ney = xyelem[1];
    int nenode = 0;
    std::vector<int> ncm;
    std::vector<double> sf;
    std::vector<std::array<double, 3>> dsf;
    VectorAssign(nodec, nvol);
    for (int m = 0; m < nelem; m++) {
        const int c[3] = {m % nex, (m / nex) % ney, m / (nex * ney)};
        std::array<double, 3> centre, xp;
        for (int a = 0; a < 3; a++) centre[a] = xymin[a] + dxy[a] * (double(c[a]) + 0.5e0);
        for (int iz = 0; iz < npxye[2]; iz++) {
            xp[2] = centre[2] + off[iz][2];
            for (int iy = 0; iy < npxye[1]; iy++) {
                xp[1] = centre[1] + off[iy][1];
                for (int ix = 0; ix < npxye[0]; ix++) {
                    xp[0] = centre[0] + off[ix][0];
                    MakSf(m, xp, idimc, xynodec, ncm, nenode, sf, dsf);
                    for (int ni = 0; ni < nenode; ni++) nvol[ncm[ni]] += sf[ni] * volp;
                }
            }
        }
    }
}

void BoundaryCondition::BCInput(std::ifstream &infile, bool if_fbc) {
    infile >> ibc;
    infile.ignore(1000, '\n');
    nbc.resize(ibc);
    if (if_fbc) fbc.resize(ibc);
    for (int i = 0; i < ibc; i++) {
        infile >> nbc[i];
        if (if_fbc) infile >> fbc[i];
        infile.ignore(1000, '\n');
    }
}

void BoundaryCondition::BCOutput(std::ofstream &outfile, std::string bc_name, bool if_fbc) {
    outfile << std::setw(15) << ibc << std::setw(15) << "! ---- " << bc_name << ".ibc ----" << "\n";
    for (int i = 0; i < ibc; i++) {
        outfile << std::setw(15) << nbc[i];
        if (if_fbc) outfile << std::setw(15) << fbc[i];
        outfile << "\n";
    }
}

void BoundaryCondition::BCSetVal(int nn, std::vector<double> &variable) {
    for (int i = 0; i < ibc; i++) variable[nbc[i] + nn] = fbc[i] * facl;
}

void BoundaryCondition::BCSetZero(int nn, std::vector<double> &variable) {
    for (int i = 0; i < ibc; i++) variable[nbc[i] + nn] = 0.0e0;
}

void BoundaryCondition::BCSetDt(int nn, std::vector<double> &variable) {
    for (int i = 0; i < ibc; i++) variable[nbc[i] + nn] = fbc[i] * dt * facl;
}

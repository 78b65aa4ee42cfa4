/*
 * FP64 CPU oracle for the explicit solid MPM step -- TEST INFRASTRUCTURE.
 *
 * Scalar restatement of the reference explicit step; see mpm_oracle.h for the
 * layouts. Must be compiled with -ffp-contract=off so that every a*b+c is two
 * rounded operations, as in the reference build on x86-64.
 *
 * How each function is pinned (checked by tests/test_oracle*.py):
 *   orc_maksf (idimc=1)   closed form 1-|x-x_a|/h, partition of unity,
 *                         sum of gradients = 0, linear reproduction, and the
 *                         dN sign rule at x == x_a (shape_function.cpp:45-50)
 *   orc_maksf (idimc=2)   partition of unity, gradient sum, linear
 *                         reproduction in every cell, boundary cells included
 *                         (with orc_node_coord, below)
 *   orc_cell_of/orc_bin   brute force on tiny grids, faces map to the higher
 *                         cell, linked-list order = stable sort by cell
 *   orc_p2g/orc_g2p       conservation of mass and momentum, sum of internal
 *                         forces = 0, free fall v = g t exactly, rigid
 *                         translation exact under PIC, against brute force
 *   orc_update_stress     bit-exact against the reference's own
 *                         constitutive_model.cpp compiled under oracle/_ref
 *   orc_bc_set_*          bit-exact against the reference's bc.cpp (oracle/_ref)
 *   orc_det3/orc_inv3     bit-exact against the reference's cal_mat.h templates
 *   orc_node_coord        linear reproduction sum_a N_a(x) x_a = x and
 *                         sum_a x_a (x) grad N_a = I in every cell, which
 *                         also pins the quadratic B-spline boundary branches
 *                         of orc_maksf (the end control points are the
 *                         Greville abscissae of the open B-spline)
 *   TPIC/APIC in orc_p2g/orc_g2p  G2P of a linear nodal field v = A x + b
 *                         returns C = A (TPIC) and B D^-1 = A (APIC) for a
 *                         non-symmetric A; P2G of particles carrying
 *                         v = A x_p + b, C = A returns A x_a + b at every
 *                         node with mass; the affine term adds no net
 *                         momentum
 *   orc_fbar              conservation of the volume-weighted J by the
 *                         nodal projection, reproduction of a constant J,
 *                         Fbar = F for a homogeneous deformation, and
 *                         det Fbar = Jbar along a run
 *   orc_j2_return         no reference counterpart: lands on the yield surface,
 *                         dgamma >= 0, elastic steps unchanged, closed-form
 *                         uniaxial hardening response
 */
#include "mpm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

void orc_grid_init(orc_state *st) {
    /* datain_para.cpp:86-90: dxy = (xymaxw - xyminw) / xyelemw,
     * xynodecw = xyelemw + idimc; data_io.cpp:53-56 node offsets. */
    for (int i = 0; i < 3; i++) {
        st->dx[i] = (st->xmax[i] - st->xmin[i]) / (double)st->nex[i];
        st->nxc[i] = st->nex[i] + st->idimc[i];
    }
    st->nelem = st->nex[0] * st->nex[1] * st->nex[2];
    st->nodec = st->nxc[0] * st->nxc[1] * st->nxc[2];
}

/* MakSf, module/shape_function.cpp:8-119. Linear branch :40-51, quadratic
 * B-spline :53-104, tensor product :107-110. fabs replaces the reference's
 * unqualified abs(double) (:43), which is the intended floating absolute value. */
int orc_maksf(const orc_state *st, int m, const double xyg[3], int32_t *nc, double *sf, double *dsf) {
    const int nex = st->nex[0], ney = st->nex[1];
    const int nxc = st->nxc[0], nyc = st->nxc[1];
    int ize = m / (nex * ney);
    int iye = (m - ize * (nex * ney)) / nex;
    int ixe = m - ize * (nex * ney) - iye * nex;
    int ixyn[3];
    int id = 0;
    for (int k = 0; k <= st->idimc[2]; k++) {
        ixyn[2] = ize + k;
        for (int j = 0; j <= st->idimc[1]; j++) {
            ixyn[1] = iye + j;
            for (int i = 0; i <= st->idimc[0]; i++) {
                ixyn[0] = ixe + i;
                nc[id] = ixyn[0] + nxc * ixyn[1] + nxc * nyc * ixyn[2];
                double sfxy[3], dsfxy[3];
                for (int ii = 0; ii < 3; ++ii) {
                    double dxi = st->dx[ii];
                    if (st->idimc[ii] == 1) {
                        double xy = st->xmin[ii] + dxi * (double)ixyn[ii];
                        double dxnpi = xyg[ii] - xy;
                        double xi = fabs(dxnpi) / dxi;
                        sfxy[ii] = 1.0 - xi;
                        if (dxnpi < 0.0)
                            dsfxy[ii] = 1.0 / dxi;
                        else
                            dsfxy[ii] = -1.0 / dxi;
                    } else {
                        /* single rank: aelemmin = 0, xynodecw = xynodec */
                        int ncxy = ixyn[ii];
                        int nw = st->nxc[ii];
                        double xy = st->xmin[ii] + dxi * (double)(ixyn[ii] - 2);
                        double dxnpi = xyg[ii] - xy;
                        double dxh = dxnpi / dxi;
                        double s1, s2, s3, d1, d2;
                        if (ncxy == 0) {
                            dxh -= 2.0;
                            s1 = 1.0; s2 = -2.0; s3 = 1.0; d1 = 2.0; d2 = -2.0;
                        } else if (ncxy == 1) {
                            dxh -= 1.0;
                            if (dxh <= 1.0) {
                                s1 = -1.5; s2 = 2.0; s3 = 0.0; d1 = -3.0; d2 = 2.0;
                            } else {
                                s1 = 0.5; s2 = -2.0; s3 = 2.0; d1 = 1.0; d2 = -2.0;
                            }
                        } else if (ncxy == nw - 1) {
                            s1 = 1.0; s2 = 0.0; s3 = 0.0; d1 = 2.0; d2 = 0.0;
                        } else if (ncxy == nw - 2) {
                            if (dxh <= 1.0) {
                                s1 = 0.5; s2 = 0.0; s3 = 0.0; d1 = 1.0; d2 = 0.0;
                            } else {
                                s1 = -1.5; s2 = 4.0; s3 = -2.0; d1 = -3.0; d2 = 4.0;
                            }
                        } else {
                            if (dxh <= 1.0) {
                                s1 = 0.5; s2 = 0.0; s3 = 0.0; d1 = 1.0; d2 = 0.0;
                            } else if (dxh <= 2.0) {
                                s1 = -1.0; s2 = 3.0; s3 = -1.5; d1 = -2.0; d2 = 3.0;
                            } else {
                                s1 = 0.5; s2 = -3.0; s3 = 4.5; d1 = 1.0; d2 = -3.0;
                            }
                        }
                        sfxy[ii] = (s1 * dxh + s2) * dxh + s3;
                        dsfxy[ii] = (d1 * dxh + d2) / dxi;
                    }
                }
                sf[id] = sfxy[0] * sfxy[1] * sfxy[2];
                dsf[3 * id + 0] = dsfxy[0] * sfxy[1] * sfxy[2];
                dsf[3 * id + 1] = sfxy[0] * dsfxy[1] * sfxy[2];
                dsf[3 * id + 2] = sfxy[0] * sfxy[1] * dsfxy[2];
                id++;
            }
        }
    }
    return id;
}

int orc_maksf_batch(const orc_state *st, int n, const int32_t *cell, const double *xyg, int32_t *nc, double *sf,
                    double *dsf) {
    int nen = (st->idimc[0] + 1) * (st->idimc[1] + 1) * (st->idimc[2] + 1);
    for (int p = 0; p < n; p++)
        orc_maksf(st, cell[p], xyg + 3 * p, nc + (size_t)nen * p, sf + (size_t)nen * p, dsf + (size_t)3 * nen * p);
    return nen;
}

/* material_point.cpp:59-65 */
int orc_cell_of(const orc_state *st, const double x[3]) {
    double xx = x[0] - st->xmin[0];
    double yy = x[1] - st->xmin[1];
    double zz = x[2] - st->xmin[2];
    int iex = (int)floor(xx / st->dx[0]);
    int iey = (int)floor(yy / st->dx[1]);
    int iez = (int)floor(zz / st->dx[2]);
    if (iex < 0 || iey < 0 || iez < 0 || iex >= st->nex[0] || iey >= st->nex[1] || iez >= st->nex[2]) return -1;
    return iex + iey * st->nex[0] + iez * st->nex[0] * st->nex[1];
}

/* MeshPointLinklist, material_point.cpp:52-77. Returns the number of
 * particles outside the grid (the reference has no check; this is an error). */
int orc_bin(orc_state *st) {
    int bad = 0;
    int32_t *idepl = (int32_t *)malloc(sizeof(int32_t) * (size_t)st->nelem);
    for (int m = 0; m < st->nelem; m++) {
        idepl[m] = -1;
        st->numep[m] = 0;
        st->idepf[m] = -1;
    }
    for (int ip = 0; ip < st->num; ip++) st->idp2p[ip] = -1;
    for (int ip = 0; ip < st->num; ip++) {
        int ne = orc_cell_of(st, st->coord + 3 * ip);
        if (ne < 0) {
            bad++;
            continue;
        }
        st->numep[ne]++;
        if (st->numep[ne] == 1) {
            st->idepf[ne] = ip;
            idepl[ne] = ip;
        } else {
            st->idp2p[idepl[ne]] = ip;
            idepl[ne] = ip;
        }
    }
    free(idepl);
    return bad;
}

/* BoundaryCondition::BCSetVal / BCSetZero, module/bc.cpp:45-67 */
void orc_bc_set_val(const orc_state *st, int l, int nn, double *var, double facl) {
    for (int i = 0; i < st->ibc[l]; i++) var[st->nbc[l][i] + nn] = st->fbc[l][i] * facl;
}
void orc_bc_set_zero(const orc_state *st, int l, int nn, double *var) {
    for (int i = 0; i < st->ibc[l]; i++) var[st->nbc[l][i] + nn] = 0.0;
}

static void apply_vel_bc(const orc_state *st, double *nvel, double facl) {
    /* var_trans_solid_explicit.cpp:91-96 */
    const int nn = st->nodec;
    orc_bc_set_val(st, 0, 0, nvel, facl);
    orc_bc_set_val(st, 1, nn, nvel, facl);
    orc_bc_set_val(st, 2, 2 * nn, nvel, facl);
    orc_bc_set_val(st, 3, 0, nvel, facl);
    orc_bc_set_val(st, 3, nn, nvel, facl);
    orc_bc_set_val(st, 3, 2 * nn, nvel, facl);
}

#define MAXNEN 27

/* DefineCpPos (module/mesh.h:32-47) on one rank (xymin == xyminw):
 * linear axes put control points on the cell corners; quadratic B-spline
 * axes put the two end points on the grid ends and the others at the cell
 * centres. Pinned in tests/test_oracle_affine.py by linear reproduction,
 * sum_a N_a(x) x_a == x, for both shape functions. */
double orc_node_coord(const orc_state *st, int axis, int idx) {
    if (st->idimc[axis] == 1) return st->xmin[axis] + st->dx[axis] * (double)idx;
    if (idx == 0) return st->xmin[axis];
    if (idx == st->nxc[axis] - 1) return st->xmax[axis];
    return st->xmin[axis] + st->dx[axis] * ((double)idx - 0.5);
}

static void node_offset(const orc_state *st, int nid, const double *xp, double d[3]) {
    int i = nid % st->nxc[0];
    int j = (nid / st->nxc[0]) % st->nxc[1];
    int k = nid / (st->nxc[0] * st->nxc[1]);
    d[0] = orc_node_coord(st, 0, i) - xp[0];
    d[1] = orc_node_coord(st, 1, j) - xp[1];
    d[2] = orc_node_coord(st, 2, k) - xp[2];
}

/* Velocity carried to node nid by particle pid (module/map_and_interpolate.h
 * TPIC::VarP2G :62-78, APIC::VarP2G :107-131): v_p + C (x_a - x_p) with C
 * the TPIC velocity gradient or the APIC product B D^-1 (summed over k, then
 * over j, as there). The reference indexes the node coordinate as
 * xyc[j][nid] (:69, :123), out of bounds for vector<array<double,3>>; this
 * restatement uses xyc[nid][j], the coordinate the APIC G2P side uses. */
static void affine_vel(const orc_state *st, int pid, int nid, const double *xp, const double *v, double out[3]) {
    const double *C = st->cmat + 9 * (size_t)pid;
    double d[3], M[9];
    node_offset(st, nid, xp, d);
    if (st->scheme == ORC_APIC) {
        const double *Di = st->dinv + 9 * (size_t)pid;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double acc = 0.0;
                for (int k = 0; k < 3; k++) acc += C[3 * i + k] * Di[3 * k + j];
                M[3 * i + j] = acc;
            }
        C = M;
    }
    for (int i = 0; i < 3; i++) {
        double t = 0.0;
        for (int j = 0; j < 3; j++) t += C[3 * i + j] * d[j];
        out[i] = v[i] + t;
    }
}

/* Particle2NodeS, var_trans_solid_explicit.cpp:20-99 (all four transfer schemes). */
void orc_p2g(orc_state *st, double facl) {
    const int nn = st->nodec;
    double fx = st->bb[0] * facl, fy = st->bb[1] * facl, fz = st->bb[2] * facl;
    int32_t ncm[MAXNEN];
    double sf[MAXNEN], dsf[3 * MAXNEN];
    memset(st->nmass, 0, sizeof(double) * nn);
    memset(st->nvof, 0, sizeof(double) * nn);
    memset(st->nmome, 0, sizeof(double) * 3 * nn);
    memset(st->nforce, 0, sizeof(double) * 3 * nn);
    for (int m = 0; m < st->nelem; m++) {
        int pid = st->idepf[m];
        while (pid != -1) {
            const double *xyp = st->coord + 3 * pid;
            const double *sts = st->stress + 6 * pid;
            const double mp = st->mass[pid], vp = st->vol[pid];
            const double *vel = st->vel + 3 * pid;
            double t0 = 0.0, t1 = 0.0, t2 = 0.0;
            if (st->trac_force) {
                t0 = st->trac_force[3 * pid];
                t1 = st->trac_force[3 * pid + 1];
                t2 = st->trac_force[3 * pid + 2];
            }
            int nenode = orc_maksf(st, m, xyp, ncm, sf, dsf);
            for (int ni = 0; ni < nenode; ni++) {
                int nid = ncm[ni];
                double sfi = sf[ni];
                double d1 = dsf[3 * ni], d2 = dsf[3 * ni + 1], d3 = dsf[3 * ni + 2];
                st->nmass[nid] += sfi * mp;
                st->nvof[nid] += sfi * vp;
                double va[3] = {vel[0], vel[1], vel[2]};
                if (st->scheme >= ORC_TPIC) affine_vel(st, pid, nid, xyp, vel, va);
                st->nmome[nid] += sfi * mp * va[0];
                st->nmome[nid + nn] += sfi * mp * va[1];
                st->nmome[nid + 2 * nn] += sfi * mp * va[2];
                double fi0 = -vp * (sts[0] * d1 + sts[5] * d2 + sts[4] * d3);
                double fi1 = -vp * (sts[5] * d1 + sts[1] * d2 + sts[3] * d3);
                double fi2 = -vp * (sts[4] * d1 + sts[3] * d2 + sts[2] * d3);
                double fe0 = sfi * (mp * fx + t0);
                double fe1 = sfi * (mp * fy + t1);
                double fe2 = sfi * (mp * fz + t2);
                st->nforce[nid] += fi0 + fe0;
                st->nforce[nid + nn] += fi1 + fe1;
                st->nforce[nid + 2 * nn] += fi2 + fe2;
            }
            pid = st->idp2p[pid];
        }
    }
    for (int n = 0; n < nn; n++) {
        if (st->nmass[n] > st->mtol) {
            st->nvel[n] = st->nmome[n] / st->nmass[n];
            st->nvel[n + nn] = st->nmome[n + nn] / st->nmass[n];
            st->nvel[n + 2 * nn] = st->nmome[n + 2 * nn] / st->nmass[n];
        } else {
            st->nvel[n] = 0.0;
            st->nvel[n + nn] = 0.0;
            st->nvel[n + 2 * nn] = 0.0;
        }
        if (st->nvof[n] < st->mtol) st->nvof[n] = 0.0;
    }
    apply_vel_bc(st, st->nvel, facl);
}

/* SolveSolid, solve_solid_explicit.cpp:8-31 */
void orc_solve_solid(orc_state *st) {
    const int nn = st->nodec;
    for (int n = 0; n < nn; n++) {
        if (st->nmass[n] > st->mtol) {
            st->naccel[n] = st->nforce[n] / st->nmass[n];
            st->naccel[n + nn] = st->nforce[n + nn] / st->nmass[n];
            st->naccel[n + 2 * nn] = st->nforce[n + 2 * nn] / st->nmass[n];
        } else {
            st->naccel[n] = 0.0;
            st->naccel[n + nn] = 0.0;
            st->naccel[n + 2 * nn] = 0.0;
        }
    }
    orc_bc_set_zero(st, 0, 0, st->naccel);
    orc_bc_set_zero(st, 1, nn, st->naccel);
    orc_bc_set_zero(st, 2, 2 * nn, st->naccel);
    orc_bc_set_zero(st, 3, 0, st->naccel);
    orc_bc_set_zero(st, 3, nn, st->naccel);
    orc_bc_set_zero(st, 3, 2 * nn, st->naccel);
}

/* cal_mat.h:38-46 (cofactor order) */
double orc_det3(const double a[9]) {
    return a[0] * (a[4] * a[8] - a[5] * a[7]) + a[1] * (a[5] * a[6] - a[3] * a[8]) +
           a[2] * (a[3] * a[7] - a[4] * a[6]);
}

/* cal_mat.h:48-66 (no singularity check, as in the reference) */
void orc_inv3(const double a[9], double o[9]) {
    double det = orc_det3(a);
    det = 1.0 / det;
    o[0] = (a[4] * a[8] - a[5] * a[7]) * det;
    o[1] = (a[2] * a[7] - a[1] * a[8]) * det;
    o[2] = (a[1] * a[5] - a[2] * a[4]) * det;
    o[3] = (a[5] * a[6] - a[3] * a[8]) * det;
    o[4] = (a[0] * a[8] - a[2] * a[6]) * det;
    o[5] = (a[2] * a[3] - a[0] * a[5]) * det;
    o[6] = (a[3] * a[7] - a[4] * a[6]) * det;
    o[7] = (a[1] * a[6] - a[0] * a[7]) * det;
    o[8] = (a[0] * a[4] - a[1] * a[3]) * det;
}

static const double EYE[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};

/* MatLinElast, constitutive_model.cpp:16-60 (hypoelastic increment) */
static void mat_linelast(double s[6], const double dF[9], const double *p) {
    double EE = p[0], nu = p[1];
    double C0 = (EE * (1.0 - nu)) / ((1.0 - 2.0 * nu) * (1.0 + nu));
    double C1 = (EE * nu) / ((1.0 - 2.0 * nu) * (1.0 + nu));
    double C2 = EE / (2.0 * (1.0 + nu));
    double e0 = dF[0] - 1.0, e1 = dF[4] - 1.0, e2 = dF[8] - 1.0;
    double e3 = dF[5] + dF[7];
    double e4 = dF[6] + dF[2];
    double e5 = dF[1] + dF[3];
    double d0 = C0 * e0 + C1 * e1 + C1 * e2;
    double d1 = C1 * e0 + C0 * e1 + C1 * e2;
    double d2 = C1 * e0 + C1 * e1 + C0 * e2;
    s[0] += d0;
    s[1] += d1;
    s[2] += d2;
    s[3] += C2 * e3;
    s[4] += C2 * e4;
    s[5] += C2 * e5;
}

static void bmat(const double F[9], double be[9]) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            be[3 * i + j] = F[3 * i] * F[3 * j] + F[3 * i + 1] * F[3 * j + 1] + F[3 * i + 2] * F[3 * j + 2];
}

static void store_voigt(double s[6], const double sig[9]) {
    s[0] = sig[0];
    s[1] = sig[4];
    s[2] = sig[8];
    s[3] = sig[5];
    s[4] = sig[2];
    s[5] = sig[1];
}

/* MatNeoHookean, constitutive_model.cpp:62-120 (explicit branch) */
static void mat_neohookean(double s[6], const double F[9], const double *p, double jac, double jac_bar) {
    double yng = p[0], nu = p[1];
    double lmd = nu * yng / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double mu = yng / (2.0 * (1.0 + nu));
    double lnjac = log(jac_bar);
    double be[9], sig[9];
    bmat(F, be);
    for (int i = 0; i < 9; i++) sig[i] = (mu * (be[i] - EYE[i]) + lmd * lnjac * EYE[i]) / jac;
    store_voigt(s, sig);
}

/* MatStVk, constitutive_model.cpp:122-185 (explicit branch) */
static void mat_stvk(double s[6], const double F[9], const double *p, double jac) {
    double yng = p[0], nu = p[1];
    double lmd = nu * yng / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double mu = yng / (2.0 * (1.0 + nu));
    double be[9], be2[9], sig[9];
    bmat(F, be);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            be2[3 * i + j] = be[3 * i] * be[j] + be[3 * i + 1] * be[3 + j] + be[3 * i + 2] * be[6 + j];
    double tr_be = 0.0;
    for (int i = 0; i < 3; i++) tr_be += be[4 * i];
    double tr_gls = 0.5 * (tr_be - 3.0);
    for (int i = 0; i < 9; i++) sig[i] = (mu * be2[i] + (lmd * tr_gls - mu) * be[i]) / jac;
    store_voigt(s, sig);
}

/* MatMooneyRivlin, constitutive_model.cpp:187-285 (explicit branch) */
static void mat_mooneyrivlin(double s[6], const double F[9], const double *p, double jac) {
    double c1 = p[0], c2 = p[1], bulk = p[2];
    double Fi[9], be[9], be2[9], sig[9];
    for (int i = 0; i < 9; i++) Fi[i] = 1.0 / cbrt(jac) * F[i];
    bmat(Fi, be);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            be2[3 * i + j] = be[3 * i] * be[j] + be[3 * i + 1] * be[3 + j] + be[3 * i + 2] * be[6 + j];
    double I1 = 0.0, trb2 = 0.0;
    for (int i = 0; i < 3; i++) I1 += be[4 * i];
    for (int i = 0; i < 3; i++) trb2 += be2[4 * i];
    double I2 = 0.5 * (pow(I1, 2) - trb2);
    double hp = bulk * log(jac) / jac;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            sig[3 * i + j] = 2.0 * ((c1 + c2 * I1) * be[3 * i + j] - c2 * be2[3 * i + j]) / jac;
        sig[4 * i] += hp - 2.0 / (3.0 * jac) * (c1 * I1 + 2.0 * c2 * I2);
    }
    store_voigt(s, sig);
}

/* J2 von Mises plasticity with linear isotropic hardening, hypoelastic
 * small-increment form; prop = {E, nu, sigma_y, H}. Radial return after
 * Simo & Hughes, Computational Inelasticity (1998), Box 3.2, written with the
 * equivalent plastic strain increment dgamma = f_trial / (3 mu + H).
 * No reference counterpart. Returns 1 when plastic. */
int orc_j2_return(const double prop[8], double s[6], double *eps_p, const double de[6], double *dgamma,
                  double *resid) {
    double E = prop[0], nu = prop[1], sy = prop[2], H = prop[3];
    double lmd = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double mu = E / (2.0 * (1.0 + nu));
    double tr = de[0] + de[1] + de[2];
    double st[6];
    st[0] = s[0] + lmd * tr + 2.0 * mu * de[0];
    st[1] = s[1] + lmd * tr + 2.0 * mu * de[1];
    st[2] = s[2] + lmd * tr + 2.0 * mu * de[2];
    st[3] = s[3] + mu * de[3];
    st[4] = s[4] + mu * de[4];
    st[5] = s[5] + mu * de[5];
    double pm = (st[0] + st[1] + st[2]) / 3.0;
    double d0 = st[0] - pm, d1 = st[1] - pm, d2 = st[2] - pm;
    double j2 = 0.5 * (d0 * d0 + d1 * d1 + d2 * d2) + st[3] * st[3] + st[4] * st[4] + st[5] * st[5];
    double q = sqrt(3.0 * j2);
    double ep = *eps_p;
    double f = q - (sy + H * ep);
    if (f <= 0.0) {
        memcpy(s, st, sizeof(st));
        *dgamma = 0.0;
        if (resid) *resid = f > 0.0 ? f : 0.0;
        return 0;
    }
    double dg = f / (3.0 * mu + H);
    double scale = 1.0 - 3.0 * mu * dg / q;
    s[0] = pm + d0 * scale;
    s[1] = pm + d1 * scale;
    s[2] = pm + d2 * scale;
    s[3] = st[3] * scale;
    s[4] = st[4] * scale;
    s[5] = st[5] * scale;
    *eps_p = ep + dg;
    *dgamma = dg;
    if (resid) {
        double qn = q * scale;
        *resid = fabs(qn - (sy + H * (*eps_p))) / sy;
    }
    return 1;
}

/* ConstitutiveModel::UpdateStress, constitutive_model.cpp:287-305 */
void orc_update_stress(int model, double s[6], const double F[9], const double dF[9], const double *prop, double jac,
                       double jac_bar) {
    switch (model) {
    case ORC_RIGID:
        for (int i = 0; i < 6; i++) s[i] = 0.0;
        break;
    case ORC_LINELAST: mat_linelast(s, dF, prop); break;
    case ORC_NEOHOOKEAN: mat_neohookean(s, F, prop, jac, jac_bar); break;
    case ORC_STVK: mat_stvk(s, F, prop, jac); break;
    case ORC_MOONEYRIVLIN: mat_mooneyrivlin(s, F, prop, jac); break;
    default: break;
    }
}

static void j2_update(const orc_state *st, int pid, const double *prop, const double dF[9]) {
    double de[6];
    de[0] = dF[0] - 1.0;
    de[1] = dF[4] - 1.0;
    de[2] = dF[8] - 1.0;
    de[3] = dF[5] + dF[7];
    de[4] = dF[6] + dF[2];
    de[5] = dF[1] + dF[3];
    double dg, r;
    double ep = st->eps_p ? st->eps_p[pid] : 0.0;
    orc_j2_return(prop, st->stress + 6 * pid, &ep, de, &dg, &r);
    if (st->eps_p) st->eps_p[pid] = ep;
}

/* Node2ParticleS + MUSL + update_point_coord_and_stress,
 * var_trans_solid_explicit.cpp:101-311 (F-bar off, PIC/FLIP). */
/* MakNodalVol, module/mesh.cpp:160-202; offsets of GaussianDistribution :12-54 */
void orc_nodal_vol(orc_state *st, const int32_t npxye[3]) {
    double dec[6][3];
    memset(dec, 0, sizeof(dec));
    for (int a = 0; a < 3; a++) {
        const double h = st->dx[a];
        switch (npxye[a]) {
        case 1: dec[0][a] = 0.0; break;
        case 2: dec[0][a] = -0.25 * h; dec[1][a] = 0.25 * h; break;
        case 3: dec[0][a] = -1.0 / 3.0 * h; dec[1][a] = 0.0; dec[2][a] = 1.0 / 3.0 * h; break;
        case 4:
            dec[0][a] = -0.375 * h; dec[1][a] = -0.125 * h; dec[2][a] = 0.125 * h; dec[3][a] = 0.375 * h;
            break;
        case 5:
            dec[0][a] = -0.4 * h; dec[1][a] = -0.2 * h; dec[2][a] = 0.0; dec[3][a] = 0.2 * h; dec[4][a] = 0.4 * h;
            break;
        case 6:
            dec[0][a] = -5.0 / 12.0 * h; dec[1][a] = -3.0 / 12.0 * h; dec[2][a] = -1.0 / 12.0 * h;
            dec[3][a] = 1.0 / 12.0 * h; dec[4][a] = 3.0 / 12.0 * h; dec[5][a] = 5.0 / 12.0 * h;
            break;
        default: break;
        }
    }
    const double volp = st->dx[0] * st->dx[1] * st->dx[2] / (double)(npxye[0] * npxye[1] * npxye[2]);
    const int nex = st->nex[0], ney = st->nex[1];
    int32_t ncm[27];
    double sf[27], dsf[81];
    memset(st->nvol, 0, sizeof(double) * st->nodec);
    for (int m = 0; m < st->nelem; m++) {
        const int ize = m / (nex * ney);
        const int iye = (m - ize * (nex * ney)) / nex;
        const int ixe = m - ize * (nex * ney) - iye * nex;
        double xye[3], xyp[3];
        xye[0] = st->xmin[0] + st->dx[0] * ((double)ixe + 0.5);
        xye[1] = st->xmin[1] + st->dx[1] * ((double)iye + 0.5);
        xye[2] = st->xmin[2] + st->dx[2] * ((double)ize + 0.5);
        for (int iz = 0; iz < npxye[2]; iz++) {
            xyp[2] = xye[2] + dec[iz][2];
            for (int iy = 0; iy < npxye[1]; iy++) {
                xyp[1] = xye[1] + dec[iy][1];
                for (int ix = 0; ix < npxye[0]; ix++) {
                    xyp[0] = xye[0] + dec[ix][0];
                    int nenode = orc_maksf(st, m, xyp, ncm, sf, dsf);
                    for (int ni = 0; ni < nenode; ni++) st->nvol[ncm[ni]] += sf[ni] * volp;
                }
            }
        }
    }
}

/* Planted mistakes for the tests that pin orc_shift_delta (finite-difference
 * gradient of the excess-volume norm and the Gauss-Newton step length): the
 * parity build leaves ORC_SHIFT_MUTANT undefined; _build/libmpm_oracle_mut1.so
 * drops the factor 2 of the gradient, _build/libmpm_oracle_mut2.so takes the
 * initial volume vol0 where the reference takes vol. Those tests must fail on
 * both. */
#if defined(ORC_SHIFT_MUTANT) && ORC_SHIFT_MUTANT == 1
#define ORC_SHIFT_GRAD_FACTOR 1.0
#else
#define ORC_SHIFT_GRAD_FACTOR 2.0
#endif
#if defined(ORC_SHIFT_MUTANT) && ORC_SHIFT_MUTANT == 2
#define ORC_SHIFT_VOL vol0
#else
#define ORC_SHIFT_VOL vol
#endif

/* DeltaCorrectionParticleShifting, module/material_point.cpp:174-227 */
void orc_shift_delta(const orc_state *st, double *delta) {
    const int nn = st->nodec, num = st->num;
    double *nei = malloc(sizeof(double) * nn);
    double *geup = calloc(3 * (size_t)num, sizeof(double));
    /* excess nodal volume and its squared norm */
    double eu_norm = 0.0;
    for (int b0 = 0; b0 < nn; b0 += ORC_SUM_BLOCK) {
        double s = 0.0;
        for (int i = b0; i < nn && i < b0 + ORC_SUM_BLOCK; i++) {
            const double e = -st->nvol[i] + st->nvof[i];
            nei[i] = e > 0.0 ? e : 0.0;
            s += nei[i] * nei[i];
        }
        eu_norm += s;
    }
    /* gradient of the excess volume at the particles */
    int32_t ncm[27];
    double sf[27], dsf[81];
    for (int m = 0; m < st->nelem; m++) {
        for (int pid = st->idepf[m]; pid != -1; pid = st->idp2p[pid]) {
            int nenode = orc_maksf(st, m, st->coord + 3 * pid, ncm, sf, dsf);
            double *g = geup + 3 * pid;
            for (int ni = 0; ni < nenode; ni++) {
                const int nid = ncm[ni];
                g[0] += dsf[3 * ni] * nei[nid];
                g[1] += dsf[3 * ni + 1] * nei[nid];
                g[2] += dsf[3 * ni + 2] * nei[nid];
            }
        }
    }
    double geup_dot = 0.0;
    for (int b0 = 0; b0 < num; b0 += ORC_SUM_BLOCK) {
        double s = 0.0;
        for (int n = b0; n < num && n < b0 + ORC_SUM_BLOCK; n++) {
            double *g = geup + 3 * n;
            for (int i = 0; i < 3; i++) g[i] *= ORC_SHIFT_GRAD_FACTOR * st->ORC_SHIFT_VOL[n];
            for (int i = 0; i < 3; i++) s += g[i] * g[i];
        }
        geup_dot += s;
    }
    const double b_0 = (geup_dot < st->mtol) ? 0.0 : (eu_norm / geup_dot);
    for (int n = 0; n < 3 * num; n++) delta[n] = -b_0 * geup[n];
    free(nei);
    free(geup);
}

void orc_g2p(orc_state *st, double facl) {
    const int nn = st->nodec;
    const double dt = st->dt;
    int32_t ncm[MAXNEN];
    double sf[MAXNEN], dsf[3 * MAXNEN];

    /* :194-202 */
    if (st->scheme != ORC_FLIP) {
        for (int n = 0; n < 3 * nn; n++) st->nvel[n] += dt * st->naccel[n];
        memset(st->vel, 0, sizeof(double) * 3 * (size_t)st->num);
    }
    /* :203-209 */
    if (st->scheme == ORC_TPIC || st->scheme == ORC_APIC) memset(st->cmat, 0, sizeof(double) * 9 * (size_t)st->num);
    if (st->scheme == ORC_APIC) memset(st->dinv, 0, sizeof(double) * 9 * (size_t)st->num);
    /* :211-240 */
    for (int m = 0; m < st->nelem; m++) {
        for (int pid = st->idepf[m]; pid != -1; pid = st->idp2p[pid]) {
            const double *xp = st->coord + 3 * pid;
            int nenode = orc_maksf(st, m, xp, ncm, sf, dsf);
            double *v = st->vel + 3 * pid;
            double *Cm = st->cmat ? st->cmat + 9 * (size_t)pid : NULL;
            double ADp[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int ni = 0; ni < nenode; ni++) {
                int nid = ncm[ni];
                double sfi = sf[ni];
                const double nv[3] = {st->nvel[nid], st->nvel[nid + nn], st->nvel[nid + 2 * nn]};
                if (st->scheme != ORC_FLIP) {
                    v[0] += sfi * nv[0];
                    v[1] += sfi * nv[1];
                    v[2] += sfi * nv[2];
                } else {
                    v[0] += sfi * dt * st->naccel[nid];
                    v[1] += sfi * dt * st->naccel[nid + nn];
                    v[2] += sfi * dt * st->naccel[nid + 2 * nn];
                }
                if (st->scheme == ORC_TPIC) {
                    /* TPIC::VarGradG2P (map_and_interpolate.h:81-99): C_ij += dN_j v_i */
                    for (int i = 0; i < 3; i++)
                        for (int j = 0; j < 3; j++) Cm[3 * i + j] += dsf[3 * ni + j] * nv[i];
                } else if (st->scheme == ORC_APIC) {
                    /* APIC::VarBmatG2P :134-148 and DmatG2P :151-164 */
                    double d[3];
                    node_offset(st, nid, xp, d);
                    for (int i = 0; i < 3; i++)
                        for (int j = 0; j < 3; j++) Cm[3 * i + j] += sfi * nv[i] * d[j];
                    for (int i = 0; i < 3; i++)
                        for (int j = 0; j < 3; j++) ADp[3 * i + j] += sfi * d[i] * d[j];
                }
            }
            /* APIC::InvDmatG2P :167-175 (InvMat3, cal_mat.h:48-66) */
            if (st->scheme == ORC_APIC) orc_inv3(ADp, st->dinv + 9 * (size_t)pid);
        }
    }
    /* MUSL :101-148 */
    memset(st->nmome, 0, sizeof(double) * 3 * nn);
    for (int m = 0; m < st->nelem; m++) {
        for (int pid = st->idepf[m]; pid != -1; pid = st->idp2p[pid]) {
            const double *xp = st->coord + 3 * pid;
            int nenode = orc_maksf(st, m, xp, ncm, sf, dsf);
            const double mp = st->mass[pid];
            const double *v = st->vel + 3 * pid;
            for (int ni = 0; ni < nenode; ni++) {
                int nid = ncm[ni];
                double sfi = sf[ni];
                double va[3] = {v[0], v[1], v[2]};
                if (st->scheme >= ORC_TPIC) affine_vel(st, pid, nid, xp, v, va);
                st->nmome[nid] += sfi * mp * va[0];
                st->nmome[nid + nn] += sfi * mp * va[1];
                st->nmome[nid + 2 * nn] += sfi * mp * va[2];
            }
        }
    }
    memset(st->nvel, 0, sizeof(double) * 3 * nn);
    for (int n = 0; n < nn; n++) {
        if (st->nmass[n] > st->mtol) {
            st->nvel[n] = st->nmome[n] / st->nmass[n];
            st->nvel[n + nn] = st->nmome[n + nn] / st->nmass[n];
            st->nvel[n + 2 * nn] = st->nmome[n + 2 * nn] / st->nmass[n];
        }
    }
    apply_vel_bc(st, st->nvel, facl);

    /* deformation gradient increment :244-296 */
    for (int m = 0; m < st->nelem; m++) {
        for (int pid = st->idepf[m]; pid != -1; pid = st->idp2p[pid]) {
            int nenode = orc_maksf(st, m, st->coord + 3 * pid, ncm, sf, dsf);
            double *dd = st->ddg + 9 * pid;
            for (int i = 0; i < 9; i++) dd[i] = 0.0;
            for (int ni = 0; ni < nenode; ni++) {
                int nid = ncm[ni];
                double d1 = dsf[3 * ni], d2 = dsf[3 * ni + 1], d3 = dsf[3 * ni + 2];
                double vu = st->nvel[nid], vv = st->nvel[nid + nn], vw = st->nvel[nid + 2 * nn];
                dd[0] += d1 * vu * dt;
                dd[1] += d2 * vu * dt;
                dd[2] += d3 * vu * dt;
                dd[3] += d1 * vv * dt;
                dd[4] += d2 * vv * dt;
                dd[5] += d3 * vv * dt;
                dd[6] += d1 * vw * dt;
                dd[7] += d2 * vw * dt;
                dd[8] += d3 * vw * dt;
            }
            dd[0] += 1.0;
            dd[4] += 1.0;
            dd[8] += 1.0;
            double *Fo = st->def_grad + 9 * pid;
            double F[9];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    double acc = 0.0;
                    for (int k = 0; k < 3; k++) acc += dd[3 * i + k] * Fo[3 * k + j];
                    F[3 * i + j] = acc;
                }
            memcpy(Fo, F, sizeof(F));
            st->det_def_grad[pid] = orc_det3(F);
            if (st->fbar) st->det_ddg[pid] = orc_det3(dd);
        }
    }
    /* :298-301 */
    if (st->fbar) orc_fbar(st);

    /* update_point_coord_and_stress :150-185; the shifting correction is
     * computed first, from the positions and volumes before the update */
    double *delta = NULL;
    if (st->shifting && st->nvol) {
        delta = malloc(sizeof(double) * 3 * (size_t)st->num);
        orc_shift_delta(st, delta);
    }
    for (int m = 0; m < st->nelem; m++) {
        for (int pid = st->idepf[m]; pid != -1; pid = st->idp2p[pid]) {
            double xyp[3] = {st->coord[3 * pid], st->coord[3 * pid + 1], st->coord[3 * pid + 2]};
            int nenode = orc_maksf(st, m, xyp, ncm, sf, dsf);
            double *x = st->coord + 3 * pid;
            for (int ni = 0; ni < nenode; ni++) {
                int nid = ncm[ni];
                double sfi = sf[ni];
                x[0] += sfi * dt * st->nvel[nid];
                x[1] += sfi * dt * st->nvel[nid + nn];
                x[2] += sfi * dt * st->nvel[nid + 2 * nn];
            }
            if (delta)
                for (int i = 0; i < 3; i++) x[i] += delta[3 * pid + i];
            st->vol[pid] = st->det_def_grad[pid] * st->vol0[pid];
            int mat = st->matid ? st->matid[pid] : 0;
            int model = st->iprop[mat];
            const double *prop = st->mat_prop[mat];
            if (model == ORC_J2)
                j2_update(st, pid, prop, st->ddg + 9 * pid);
            else if (model == ORC_TAIT)
                orc_tait_stress(prop, st->det_def_grad[pid], st->ddg + 9 * pid, dt, st->stress + 6 * pid);
            else if (st->fbar)  /* F-bar: F and J-bar of the projection, true J (solid_material_point.cpp:143-149) */
                orc_update_stress(model, st->stress + 6 * pid, st->def_grad_bar + 9 * pid, st->ddg + 9 * pid, prop,
                                  st->det_def_grad[pid], st->det_def_grad_bar[pid]);
            else
                orc_update_stress(model, st->stress + 6 * pid, st->def_grad + 9 * pid, st->ddg + 9 * pid, prop,
                                  st->det_def_grad[pid], st->det_def_grad[pid]);
        }
    }
    free(delta);
}

/* cal_def_grad_bar (module/solid/explicit/fbar_projection.cpp:18-94). The
 * reference declares delta_def_grad as vector<array<array<double,3>,3>>
 * ([pid][i][j]) while its caller passes the [i][j][pid] nested vectors built
 * at var_trans_solid_explicit.cpp:242-244, which does not compile; this is
 * the projection with delta_def_grad[pid][i][j] = dF_ij of particle pid.
 * Nodal J-bar: sum of N (Jbar_p det dF_p) V_p over cells ascending and
 * particles in list order, divided by nvof where nvof > mtol (else 0);
 * particle J-bar: sum of N nJbar; dF *= cbrt(Jbar / (Jbar_old det dF));
 * Fbar = dF Fbar_old. Pinned in tests/test_oracle_fbar.py: a homogeneous
 * deformation gives Fbar = F; det Fbar = Jbar after many steps; the nodal
 * projection conserves the volume-weighted J. */
void orc_fbar(orc_state *st) {
    const int nn = st->nodec;
    int32_t ncm[MAXNEN];
    double sf[MAXNEN], dsf[3 * MAXNEN];
    double *nj = st->njbar;
    memset(nj, 0, sizeof(double) * nn);
    for (int m = 0; m < st->nelem; m++) {
        for (int pid = st->idepf[m]; pid != -1; pid = st->idp2p[pid]) {
            int nenode = orc_maksf(st, m, st->coord + 3 * pid, ncm, sf, dsf);
            const double jp = st->det_def_grad_bar[pid] * st->det_ddg[pid];
            for (int ni = 0; ni < nenode; ni++) nj[ncm[ni]] += sf[ni] * jp * st->vol[pid];
        }
    }
    for (int n = 0; n < nn; n++) {
        if (st->nvof[n] > st->mtol)
            nj[n] /= st->nvof[n];
        else
            nj[n] = 0.0;
    }
    for (int m = 0; m < st->nelem; m++) {
        for (int pid = st->idepf[m]; pid != -1; pid = st->idp2p[pid]) {
            int nenode = orc_maksf(st, m, st->coord + 3 * pid, ncm, sf, dsf);
            double jbar = 0.0;
            for (int ni = 0; ni < nenode; ni++) jbar += sf[ni] * nj[ncm[ni]];
            const double J = st->det_def_grad_bar[pid] * st->det_ddg[pid];
            double *dd = st->ddg + 9 * pid;
            for (int q = 0; q < 9; q++) dd[q] *= cbrt((jbar / J));
            double *Fb = st->def_grad_bar + 9 * pid;
            double Fn[9];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    double acc = 0.0;
                    for (int k = 0; k < 3; k++) acc += dd[3 * i + k] * Fb[3 * k + j];
                    Fn[3 * i + j] = acc;
                }
            memcpy(Fb, Fn, sizeof(Fn));
            st->det_def_grad_bar[pid] = jbar;
        }
    }
}

/* Tait equation of state + Newtonian viscosity (textbook weakly
 * compressible fluid; pinned by closed forms in tests/test_oracle_invariants.py) */
void orc_tait_stress(const double *prop, double jac, const double dF[9], double dt, double s[6]) {
    const double rho0 = prop[0], c = prop[1], gam = prop[2], mu = prop[3];
    const double B = rho0 * c * c / gam;
    const double p = B * expm1(-gam * log1p(jac - 1.0));
    const double idt = 1.0 / dt;
    double l[9];
    for (int q = 0; q < 9; q++) l[q] = (dF[q] - ((q == 0 || q == 4 || q == 8) ? 1.0 : 0.0)) * idt;
    const double tr = (l[0] + l[4] + l[8]) / 3.0;
    s[0] = -p + 2.0 * mu * (l[0] - tr);
    s[1] = -p + 2.0 * mu * (l[4] - tr);
    s[2] = -p + 2.0 * mu * (l[8] - tr);
    s[3] = mu * (l[5] + l[7]);
    s[4] = mu * (l[6] + l[2]);
    s[5] = mu * (l[1] + l[3]);
}

/* work/src_solid/explicit/main.cpp:43-48 */
double orc_facl(const orc_state *st, int istep) {
    if (istep <= st->nlstep) return (1.0 / (double)st->nlstep) * (double)istep;
    return 1.0;
}

/* one pass of the time loop body, main.cpp:51-57 */
int orc_step(orc_state *st, int istep) {
    double facl = orc_facl(st, istep);
    int bad = orc_bin(st);
    if (bad) return -bad;
    orc_p2g(st, facl);
    orc_solve_solid(st);
    orc_g2p(st, facl);
    return 0;
}

int orc_run(orc_state *st, int ista, int iend) {
    for (int istep = ista; istep <= iend; istep++) {
        int rc = orc_step(st, istep);
        if (rc) return rc;
    }
    return 0;
}

/* ------------------------------------------------------------------------
 * Implicit-path transfers: the P2G and G2P halves of the reference's
 * implicit solid step, without its PETSc solve (the nodal displacement
 * increment comes from the caller). PIC and FLIP; particle shifting and
 * F-bar off.
 * ------------------------------------------------------------------------ */

/* NewmarkBetaParaSet, module/nonlinear_time_integration.cpp:43-57 */
void orc_newmark_para(double gamma_nb, double beta_nb, double dt, double nb[6]) {
    nb[0] = gamma_nb / (beta_nb * dt);
    nb[1] = gamma_nb / beta_nb - 1.0;
    nb[2] = dt / 2.0 * (gamma_nb / beta_nb - 2.0);
    nb[3] = 1.0 / (beta_nb * dt * dt);
    nb[4] = 1.0 / (beta_nb * dt);
    nb[5] = 1.0 / (2.0 * beta_nb) - 1.0;
}

/* SolidMaterialPoint::Particle2Node, module/solid/implicit/var_trans_solid_implicit.cpp:20-80:
 * nmass, nvof, nmome = sum N m v (VelP2G) and nforce = sum N m a (AccelP2G,
 * module/map_and_interpolate.cpp:12-49) in the cell / particle-list order;
 * nvel = nmome / nmass and naccel = nforce / nmass where nmass > mtol (else
 * 0); velocity BCs set (u, v, w lists, then the rigid list on all three
 * components), acceleration BCs zeroed. accel: [num][3]. */
void orc_implicit_p2g(orc_state *st, const double *accel, double facl) {
    const int nn = st->nodec;
    int32_t ncm[MAXNEN];
    double sf[MAXNEN], dsf[3 * MAXNEN];
    memset(st->nmass, 0, sizeof(double) * nn);
    memset(st->nvof, 0, sizeof(double) * nn);
    memset(st->nmome, 0, sizeof(double) * 3 * nn);
    memset(st->nforce, 0, sizeof(double) * 3 * nn);
    for (int m = 0; m < st->nelem; m++) {
        for (int pid = st->idepf[m]; pid != -1; pid = st->idp2p[pid]) {
            const double *xyp = st->coord + 3 * pid;
            const double mp = st->mass[pid], vp = st->vol[pid];
            const double *v = st->vel + 3 * pid, *a = accel + 3 * pid;
            int nenode = orc_maksf(st, m, xyp, ncm, sf, dsf);
            for (int ni = 0; ni < nenode; ni++) {
                int nid = ncm[ni];
                double sfi = sf[ni];
                st->nmass[nid] += sfi * mp;
                st->nvof[nid] += sfi * vp;
                for (int c = 0; c < 3; c++) st->nmome[nid + c * nn] += sfi * mp * v[c];
                for (int c = 0; c < 3; c++) st->nforce[nid + c * nn] += sfi * mp * a[c];
            }
        }
    }
    memset(st->nvel, 0, sizeof(double) * 3 * nn);
    memset(st->naccel, 0, sizeof(double) * 3 * nn);
    for (int n = 0; n < nn; n++) {
        if (st->nmass[n] > st->mtol) {
            for (int c = 0; c < 3; c++) {
                st->nvel[n + c * nn] = st->nmome[n + c * nn] / st->nmass[n];
                st->naccel[n + c * nn] = st->nforce[n + c * nn] / st->nmass[n];
            }
        }
    }
    apply_vel_bc(st, st->nvel, facl);
    orc_bc_set_zero(st, 0, 0, st->naccel);
    orc_bc_set_zero(st, 1, nn, st->naccel);
    orc_bc_set_zero(st, 2, 2 * nn, st->naccel);
    orc_bc_set_zero(st, 3, 0, st->naccel);
    orc_bc_set_zero(st, 3, nn, st->naccel);
    orc_bc_set_zero(st, 3, 2 * nn, st->naccel);
}

/* SolidMaterialPoint::Node2Particle, var_trans_solid_implicit.cpp:106-167,
 * for a converged nodal displacement increment ndispl ([u|v|w] blocks):
 *   PredictNewmarkBetaVelAndAccel + CommitNodalKinematics
 *     (nonlinear_time_integration.cpp:59-77): nvel, naccel from ndispl;
 *   per particle: dF = I + sum dN (x) ndispl (ComputeDeltaDefGrad,
 *     module/material_point.cpp:127-151 with af = 1, the identity added
 *     last), F = dF F, J, V = J V0, the stress update; for non-rigid
 *     materials displ = sum N ndispl, accel = sum N naccel and (not FLIP)
 *     vel = sum N nvel;
 *   CommitParticleKinematics (nonlinear_time_integration.cpp:79-96): FLIP
 *     vel += dt ((1 - gamma) accel_old + gamma accel); x += displ.
 * accel: [num][3], updated. Returns -1 for TPIC/APIC, -2 with shifting or
 * F-bar on. */
int orc_implicit_g2p(orc_state *st, double *accel, const double *ndispl, const double nb[6], double gamma_nb) {
    if (st->scheme != ORC_PIC && st->scheme != ORC_FLIP) return -1;
    if (st->shifting || st->fbar) return -2;
    const int nn = st->nodec, np = st->num;
    const double dt = st->dt, af = 1.0;
    const int flip = st->scheme == ORC_FLIP;
    int32_t ncm[MAXNEN];
    double sf[MAXNEN], dsf[3 * MAXNEN];
    for (int n = 0; n < 3 * nn; n++) {
        const double u = ndispl[n], v = st->nvel[n], a = st->naccel[n];
        const double vk = nb[0] * u - nb[1] * v - nb[2] * a;
        const double ak = nb[3] * u - nb[4] * v - nb[5] * a;
        st->nvel[n] = vk;
        st->naccel[n] = ak;
    }
    double *accel_old = NULL;
    if (flip) {
        accel_old = malloc(sizeof(double) * 3 * (size_t)np);
        memcpy(accel_old, accel, sizeof(double) * 3 * (size_t)np);
    }
    memset(accel, 0, sizeof(double) * 3 * (size_t)np);
    if (!flip) memset(st->vel, 0, sizeof(double) * 3 * (size_t)np);
    double *displ = calloc(3 * (size_t)np, sizeof(double));
    for (int m = 0; m < st->nelem; m++) {
        for (int pid = st->idepf[m]; pid != -1; pid = st->idp2p[pid]) {
            const double *xp = st->coord + 3 * pid;
            int nenode = orc_maksf(st, m, xp, ncm, sf, dsf);
            double *dd = st->ddg + 9 * pid;
            for (int i = 0; i < 9; i++) dd[i] = 0.0;
            for (int ni = 0; ni < nenode; ni++) {
                int nid = ncm[ni];
                double d1 = dsf[3 * ni], d2 = dsf[3 * ni + 1], d3 = dsf[3 * ni + 2];
                double uu = ndispl[nid], uv = ndispl[nid + nn], uw = ndispl[nid + 2 * nn];
                dd[0] += d1 * af * uu;
                dd[1] += d2 * af * uu;
                dd[2] += d3 * af * uu;
                dd[3] += d1 * af * uv;
                dd[4] += d2 * af * uv;
                dd[5] += d3 * af * uv;
                dd[6] += d1 * af * uw;
                dd[7] += d2 * af * uw;
                dd[8] += d3 * af * uw;
            }
            dd[0] += 1.0;
            dd[4] += 1.0;
            dd[8] += 1.0;
            double *Fo = st->def_grad + 9 * pid;
            double F[9];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    double acc = 0.0;
                    for (int k = 0; k < 3; k++) acc += dd[3 * i + k] * Fo[3 * k + j];
                    F[3 * i + j] = acc;
                }
            memcpy(Fo, F, sizeof(F));
            st->det_def_grad[pid] = orc_det3(F);
            st->vol[pid] = st->det_def_grad[pid] * st->vol0[pid];
            int mat = st->matid ? st->matid[pid] : 0;
            int model = st->iprop[mat];
            const double *prop = st->mat_prop[mat];
            if (model == ORC_J2)
                j2_update(st, pid, prop, dd);
            else if (model == ORC_TAIT)
                orc_tait_stress(prop, st->det_def_grad[pid], dd, dt, st->stress + 6 * pid);
            else
                orc_update_stress(model, st->stress + 6 * pid, F, dd, prop, st->det_def_grad[pid],
                                  st->det_def_grad[pid]);
            if (model != ORC_RIGID) {
                for (int ni = 0; ni < nenode; ni++) {
                    int nid = ncm[ni];
                    double sfi = sf[ni];
                    for (int c = 0; c < 3; c++) displ[3 * pid + c] += sfi * ndispl[nid + c * nn];
                    for (int c = 0; c < 3; c++) accel[3 * pid + c] += sfi * st->naccel[nid + c * nn];
                    if (!flip)
                        for (int c = 0; c < 3; c++) st->vel[3 * pid + c] += sfi * st->nvel[nid + c * nn];
                }
            }
        }
    }
    for (int p = 0; p < np; p++)
        for (int i = 0; i < 3; i++) {
            if (flip)
                st->vel[3 * p + i] += dt * ((1.0 - gamma_nb) * accel_old[3 * p + i] + gamma_nb * accel[3 * p + i]);
            st->coord[3 * p + i] += displ[3 * p + i];
        }
    free(displ);
    free(accel_old);
    return 0;
}

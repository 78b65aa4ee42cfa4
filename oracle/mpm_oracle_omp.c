/*
 * Multithreaded CPU port of the explicit step -- CPU BASELINE ONLY.
 *
 * The same step as orc_step (mpm_oracle.c; var_trans_solid_explicit.cpp
 * order: binning, P2G, SolveSolid, G2P, MUSL, dF/F/J, x/V/stress) spread over
 * OpenMP threads, for the CPU-baseline figure of the benchmark. Only the
 * benchmark's baseline leg and the tests load it; it checks nothing and the
 * engine never calls it.
 *
 * Node sums use cell colouring: cells whose indices agree modulo
 * (idimc + 1) on every axis touch disjoint stencils, so each colour scatters
 * in parallel without atomics. The summation order per node is colour order,
 * then the reference's particle order, so results equal the oracle's to
 * rounding, not bitwise (tests/test_oracle_omp.py). Covered: PIC and FLIP,
 * linear and quadratic shape functions, the models of orc_update_stress and
 * Tait; not covered (returns -1000): J2, TPIC/APIC, shifting, F-bar.
 */
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "mpm_oracle.h"

#define MAXNEN 27

static void bin_parallel(orc_state *st, int32_t *cell, int *bad) {
    int nbad = 0;
#pragma omp parallel for reduction(+ : nbad) schedule(static)
    for (int ip = 0; ip < st->num; ip++) {
        cell[ip] = orc_cell_of(st, st->coord + 3 * ip);
        if (cell[ip] < 0) nbad++;
    }
    *bad = nbad;
    if (nbad) return;
    /* material_point.cpp:52-77: per-cell lists in ascending particle order */
#pragma omp parallel for schedule(static)
    for (int m = 0; m < st->nelem; m++) {
        st->idepf[m] = -1;
        st->numep[m] = 0;
    }
    for (int ip = st->num - 1; ip >= 0; ip--) {
        int m = cell[ip];
        st->idp2p[ip] = st->idepf[m];
        st->idepf[m] = ip;
        st->numep[m]++;
    }
}

static void apply_bc_vel(const orc_state *st, double *nvel, double facl) {
    const int nn = st->nodec;
    orc_bc_set_val(st, 0, 0, nvel, facl);
    orc_bc_set_val(st, 1, nn, nvel, facl);
    orc_bc_set_val(st, 2, 2 * nn, nvel, facl);
    orc_bc_set_val(st, 3, 0, nvel, facl);
    orc_bc_set_val(st, 3, nn, nvel, facl);
    orc_bc_set_val(st, 3, 2 * nn, nvel, facl);
}

/* Scatter over the cells of one colour; full = 1: mass, volume, momentum
 * and force (Particle2NodeS), full = 0: momentum only (MUSL). */
static void scatter_colour(orc_state *st, int c0, int c1, int c2, int full, double facl) {
    const int nn = st->nodec;
    const int s0 = st->idimc[0] + 1, s1 = st->idimc[1] + 1, s2 = st->idimc[2] + 1;
    const int n0 = (st->nex[0] - c0 + s0 - 1) / s0, n1 = (st->nex[1] - c1 + s1 - 1) / s1,
              n2 = (st->nex[2] - c2 + s2 - 1) / s2;
    const double fx = st->bb[0] * facl, fy = st->bb[1] * facl, fz = st->bb[2] * facl;
    const long total = (long)n0 * n1 * n2;
#pragma omp parallel for schedule(dynamic, 16)
    for (long t = 0; t < total; t++) {
        int i = c0 + s0 * (int)(t % n0);
        int j = c1 + s1 * (int)((t / n0) % n1);
        int k = c2 + s2 * (int)(t / ((long)n0 * n1));
        int m = i + st->nex[0] * (j + st->nex[1] * k);
        int32_t ncm[MAXNEN];
        double sf[MAXNEN], dsf[3 * MAXNEN];
        for (int pid = st->idepf[m]; pid != -1; pid = st->idp2p[pid]) {
            const double *xyp = st->coord + 3 * pid;
            const double *v = st->vel + 3 * pid;
            const double mp = st->mass[pid];
            int nen = orc_maksf(st, m, xyp, ncm, sf, dsf);
            if (!full) {
                for (int a = 0; a < nen; a++) {
                    int nid = ncm[a];
                    st->nmome[nid] += sf[a] * mp * v[0];
                    st->nmome[nid + nn] += sf[a] * mp * v[1];
                    st->nmome[nid + 2 * nn] += sf[a] * mp * v[2];
                }
                continue;
            }
            const double *sts = st->stress + 6 * pid;
            const double vp = st->vol[pid];
            double t0 = 0.0, t1 = 0.0, t2 = 0.0;
            if (st->trac_force) {
                t0 = st->trac_force[3 * pid];
                t1 = st->trac_force[3 * pid + 1];
                t2 = st->trac_force[3 * pid + 2];
            }
            for (int a = 0; a < nen; a++) {
                int nid = ncm[a];
                double sfi = sf[a], d1 = dsf[3 * a], d2 = dsf[3 * a + 1], d3 = dsf[3 * a + 2];
                st->nmass[nid] += sfi * mp;
                st->nvof[nid] += sfi * vp;
                st->nmome[nid] += sfi * mp * v[0];
                st->nmome[nid + nn] += sfi * mp * v[1];
                st->nmome[nid + 2 * nn] += sfi * mp * v[2];
                st->nforce[nid] += -vp * (sts[0] * d1 + sts[5] * d2 + sts[4] * d3) + sfi * (mp * fx + t0);
                st->nforce[nid + nn] += -vp * (sts[5] * d1 + sts[1] * d2 + sts[3] * d3) + sfi * (mp * fy + t1);
                st->nforce[nid + 2 * nn] += -vp * (sts[4] * d1 + sts[3] * d2 + sts[2] * d3) + sfi * (mp * fz + t2);
            }
        }
    }
}

static void scatter(orc_state *st, int full, double facl) {
    for (int c2 = 0; c2 <= st->idimc[2]; c2++)
        for (int c1 = 0; c1 <= st->idimc[1]; c1++)
            for (int c0 = 0; c0 <= st->idimc[0]; c0++) scatter_colour(st, c0, c1, c2, full, facl);
}

static void zero(double *a, long n) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) a[i] = 0.0;
}

static int step_omp(orc_state *st, int istep, int32_t *cell) {
    const int nn = st->nodec;
    const double dt = st->dt;
    const double facl = orc_facl(st, istep);
    int bad = 0;
    bin_parallel(st, cell, &bad);
    if (bad) return -bad;

    /* Particle2NodeS */
    zero(st->nmass, nn);
    zero(st->nvof, nn);
    zero(st->nmome, 3L * nn);
    zero(st->nforce, 3L * nn);
    scatter(st, 1, facl);
#pragma omp parallel for schedule(static)
    for (int n = 0; n < nn; n++) {
        double m = st->nmass[n];
        for (int c = 0; c < 3; c++) st->nvel[n + c * nn] = (m > st->mtol) ? st->nmome[n + c * nn] / m : 0.0;
        if (st->nvof[n] < st->mtol) st->nvof[n] = 0.0;
    }
    apply_bc_vel(st, st->nvel, facl);

    /* SolveSolid */
#pragma omp parallel for schedule(static)
    for (int n = 0; n < nn; n++) {
        double m = st->nmass[n];
        for (int c = 0; c < 3; c++) st->naccel[n + c * nn] = (m > st->mtol) ? st->nforce[n + c * nn] / m : 0.0;
    }
    for (int l = 0; l < 4; l++)
        for (int c = 0; c < 3; c++)
            if (l == c || l == 3) orc_bc_set_zero(st, l, c * nn, st->naccel);

    /* Node2ParticleS: velocity */
    if (st->scheme != ORC_FLIP) {
#pragma omp parallel for schedule(static)
        for (long n = 0; n < 3L * nn; n++) st->nvel[n] += dt * st->naccel[n];
    }
#pragma omp parallel for schedule(static)
    for (int pid = 0; pid < st->num; pid++) {
        int32_t ncm[MAXNEN];
        double sf[MAXNEN], dsf[3 * MAXNEN];
        int nen = orc_maksf(st, cell[pid], st->coord + 3 * pid, ncm, sf, dsf);
        double *v = st->vel + 3 * pid;
        if (st->scheme != ORC_FLIP) v[0] = v[1] = v[2] = 0.0;
        for (int a = 0; a < nen; a++) {
            int nid = ncm[a];
            if (st->scheme != ORC_FLIP) {
                v[0] += sf[a] * st->nvel[nid];
                v[1] += sf[a] * st->nvel[nid + nn];
                v[2] += sf[a] * st->nvel[nid + 2 * nn];
            } else {
                v[0] += sf[a] * dt * st->naccel[nid];
                v[1] += sf[a] * dt * st->naccel[nid + nn];
                v[2] += sf[a] * dt * st->naccel[nid + 2 * nn];
            }
        }
    }

    /* MUSL */
    zero(st->nmome, 3L * nn);
    scatter(st, 0, facl);
#pragma omp parallel for schedule(static)
    for (int n = 0; n < nn; n++) {
        double m = st->nmass[n];
        for (int c = 0; c < 3; c++) st->nvel[n + c * nn] = (m > st->mtol) ? st->nmome[n + c * nn] / m : 0.0;
    }
    apply_bc_vel(st, st->nvel, facl);

    /* dF, F, J, then x, V, stress (:244-296, :150-185) */
    int nonpos = 0;
#pragma omp parallel for schedule(static) reduction(+ : nonpos)
    for (int pid = 0; pid < st->num; pid++) {
        int32_t ncm[MAXNEN];
        double sf[MAXNEN], dsf[3 * MAXNEN];
        double *x = st->coord + 3 * pid;
        int nen = orc_maksf(st, cell[pid], x, ncm, sf, dsf);
        double dd[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, dx[3] = {0, 0, 0};
        for (int a = 0; a < nen; a++) {
            int nid = ncm[a];
            const double vu = st->nvel[nid], vv = st->nvel[nid + nn], vw = st->nvel[nid + 2 * nn];
            for (int q = 0; q < 3; q++) {
                dd[q] += dsf[3 * a + q] * vu * dt;
                dd[3 + q] += dsf[3 * a + q] * vv * dt;
                dd[6 + q] += dsf[3 * a + q] * vw * dt;
            }
            dx[0] += sf[a] * dt * vu;
            dx[1] += sf[a] * dt * vv;
            dx[2] += sf[a] * dt * vw;
        }
        dd[0] += 1.0;
        dd[4] += 1.0;
        dd[8] += 1.0;
        double *Fo = st->def_grad + 9 * pid, F[9];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double acc = 0.0;
                for (int k = 0; k < 3; k++) acc += dd[3 * i + k] * Fo[3 * k + j];
                F[3 * i + j] = acc;
            }
        memcpy(Fo, F, sizeof(F));
        const double J = orc_det3(F);
        st->det_def_grad[pid] = J;
        if (!(J > 0.0)) nonpos++;
        for (int c = 0; c < 3; c++) x[c] += dx[c];
        st->vol[pid] = J * st->vol0[pid];
        int mat = st->matid ? st->matid[pid] : 0;
        if (st->iprop[mat] == ORC_TAIT)
            orc_tait_stress(st->mat_prop[mat], J, dd, dt, st->stress + 6 * pid);
        else
            orc_update_stress(st->iprop[mat], st->stress + 6 * pid, F, dd, st->mat_prop[mat], J, J);
    }
    return nonpos ? -2000 : 0;
}

int orc_run_omp(orc_state *st, int ista, int iend, int nthreads) {
    if (st->scheme > ORC_FLIP || st->shifting || st->fbar) return -1000;
    for (int m = 0; m < st->nmat; m++)
        if (st->iprop[m] == ORC_J2) return -1000;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    int32_t *cell = (int32_t *)malloc(sizeof(int32_t) * (size_t)(st->num > 0 ? st->num : 1));
    int rc = 0;
    for (int istep = ista; istep <= iend && !rc; istep++) rc = step_omp(st, istep, cell);
    free(cell);
    return rc;
}

int orc_omp_max_threads(void) { return omp_get_max_threads(); }

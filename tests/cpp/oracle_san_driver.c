/* Runs the FP64 oracle's step on small scenes, every constitutive model,
 * transfer scheme and shape-function order, with particle shifting, F-bar,
 * two materials, traction and boundary lists, plus the network oracle. Built
 * with AddressSanitizer and UndefinedBehaviorSanitizer (tests/cpp/Makefile
 * `sanitize`): a run that touches memory out of bounds or hits undefined
 * behaviour aborts. It checks only that the results stay finite; numerics
 * are checked by the parity tests. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mpm_oracle.h"

static int g_fail = 0;

static void *zalloc(size_t n) {
    void *p = calloc(n ? n : 1, 1);
    if (!p) {
        fprintf(stderr, "out of memory\n");
        exit(2);
    }
    return p;
}

/* xorshift for reproducible jitter */
static unsigned long long g_rng = 0x9E3779B97F4A7C15ull;
static double urand(void) {
    g_rng ^= g_rng << 13;
    g_rng ^= g_rng >> 7;
    g_rng ^= g_rng << 17;
    return (double)(g_rng >> 11) * (1.0 / 9007199254740992.0);
}

typedef struct {
    int model;
    double prop[ORC_MAXPROP];
} material;

static void run_case(const char *tag, const int idimc[3], int scheme, const material *mats, int nmat, int shifting,
                     int fbar, int nsteps) {
    orc_state st;
    memset(&st, 0, sizeof st);
    const int ne[3] = {6, 5, 4};
    for (int a = 0; a < 3; a++) {
        st.xmin[a] = 0.0;
        st.xmax[a] = 0.6 + 0.1 * a;
        st.nex[a] = ne[a];
        st.idimc[a] = idimc[a];
    }
    orc_grid_init(&st);
    st.dt = 1e-5;
    st.mtol = 1e-12;
    st.bb[1] = -9.81;
    st.scheme = scheme;
    st.nlstep = 3;
    st.nmat = nmat;
    for (int m = 0; m < nmat; m++) {
        st.iprop[m] = mats[m].model;
        memcpy(st.mat_prop[m], mats[m].prop, sizeof mats[m].prop);
    }
    /* a block of 2x2x2 points per cell over the inner cells, jittered */
    const int lo[3] = {1, 1, 1}, hi[3] = {ne[0] - 1, ne[1] - 2, ne[2] - 1};
    int n = 0;
    for (int a = 0; a < 3; a++) n = (a ? n : 1) * (hi[a] - lo[a]) * 2;
    st.num = n;
    st.coord = zalloc(sizeof(double) * 3 * n);
    st.vel = zalloc(sizeof(double) * 3 * n);
    st.mass = zalloc(sizeof(double) * n);
    st.vol = zalloc(sizeof(double) * n);
    st.vol0 = zalloc(sizeof(double) * n);
    st.stress = zalloc(sizeof(double) * 6 * n);
    st.def_grad = zalloc(sizeof(double) * 9 * n);
    st.det_def_grad = zalloc(sizeof(double) * n);
    st.trac_force = zalloc(sizeof(double) * 3 * n);
    st.matid = zalloc(sizeof(int32_t) * n);
    st.eps_p = zalloc(sizeof(double) * n);
    int p = 0;
    const double vp = st.dx[0] * st.dx[1] * st.dx[2] / 8.0;
    for (int k = 2 * lo[2]; k < 2 * hi[2]; k++)
        for (int j = 2 * lo[1]; j < 2 * hi[1]; j++)
            for (int i = 2 * lo[0]; i < 2 * hi[0]; i++) {
                const int ijk[3] = {i, j, k};
                for (int a = 0; a < 3; a++) {
                    st.coord[3 * p + a] = st.xmin[a] + st.dx[a] * (0.5 * ijk[a] + 0.25 + 0.05 * (urand() - 0.5));
                    st.vel[3 * p + a] = 0.01 * (urand() - 0.5);
                    st.trac_force[3 * p + a] = (a == 1 && j == 2 * hi[1] - 1) ? -1e-3 : 0.0;
                }
                st.mass[p] = 1000.0 * vp;
                st.vol[p] = st.vol0[p] = vp;
                st.def_grad[9 * p] = st.def_grad[9 * p + 4] = st.def_grad[9 * p + 8] = 1.0;
                st.det_def_grad[p] = 1.0;
                st.matid[p] = nmat > 1 ? (i + j + k) % nmat : 0;
                p++;
            }
    const int nelem = st.nelem, nodec = st.nodec;
    st.idepf = zalloc(sizeof(int32_t) * nelem);
    st.idp2p = zalloc(sizeof(int32_t) * n);
    st.numep = zalloc(sizeof(int32_t) * nelem);
    st.nmass = zalloc(sizeof(double) * nodec);
    st.nvof = zalloc(sizeof(double) * nodec);
    st.nmome = zalloc(sizeof(double) * 3 * nodec);
    st.nforce = zalloc(sizeof(double) * 3 * nodec);
    st.nvel = zalloc(sizeof(double) * 3 * nodec);
    st.naccel = zalloc(sizeof(double) * 3 * nodec);
    st.nvol = zalloc(sizeof(double) * nodec);
    st.ddg = zalloc(sizeof(double) * 9 * n);
    st.cmat = zalloc(sizeof(double) * 9 * n);
    st.dinv = zalloc(sizeof(double) * 9 * n);
    st.def_grad_bar = zalloc(sizeof(double) * 9 * n);
    st.det_def_grad_bar = zalloc(sizeof(double) * n);
    st.njbar = zalloc(sizeof(double) * nodec);
    st.det_ddg = zalloc(sizeof(double) * n);
    for (int q = 0; q < n; q++) {
        st.def_grad_bar[9 * q] = st.def_grad_bar[9 * q + 4] = st.def_grad_bar[9 * q + 8] = 1.0;
        st.det_def_grad_bar[q] = 1.0;
    }
    st.fbar = fbar;
    /* floor nodes fixed in every component, one side rolling in x, a rigid list */
    int32_t *nbc[4];
    double *fbc[4];
    for (int c = 0; c < 4; c++) {
        nbc[c] = zalloc(sizeof(int32_t) * nodec);
        fbc[c] = zalloc(sizeof(double) * nodec);
        int cnt = 0;
        for (int kk = 0; kk < st.nxc[2]; kk++)
            for (int jj = 0; jj < st.nxc[1]; jj++)
                for (int ii = 0; ii < st.nxc[0]; ii++) {
                    const int nid = ii + st.nxc[0] * jj + st.nxc[0] * st.nxc[1] * kk;
                    const int on = (c < 3 && jj == 0) || (c == 0 && ii == 0) || (c == 3 && ii == st.nxc[0] - 1 && kk == 0);
                    if (on) {
                        nbc[c][cnt] = nid;
                        fbc[c][cnt] = 0.0;
                        cnt++;
                    }
                }
        st.ibc[c] = cnt;
        st.nbc[c] = nbc[c];
        st.fbc[c] = fbc[c];
    }
    if (shifting) {
        const int32_t npxye[3] = {2, 2, 2};
        orc_nodal_vol(&st, npxye);
        st.shifting = 1;
    }
    const int rc = orc_run(&st, 1, nsteps);
    int finite = 1;
    for (int q = 0; q < 3 * n; q++) finite = finite && isfinite(st.coord[q]);
    for (int q = 0; q < 6 * n; q++) finite = finite && isfinite(st.stress[q]);
    printf("%-44s %4d points, %d steps: rc %d, %s\n", tag, n, nsteps, rc, finite ? "finite" : "NON-FINITE");
    if (rc != 0 || !finite) g_fail++;
    void *bufs[] = {st.coord, st.vel, st.mass, st.vol, st.vol0, st.stress, st.def_grad, st.det_def_grad,
                    st.trac_force, st.matid, st.eps_p, st.idepf, st.idp2p, st.numep, st.nmass, st.nvof, st.nmome,
                    st.nforce, st.nvel, st.naccel, st.nvol, st.ddg, st.cmat, st.dinv, st.def_grad_bar,
                    st.det_def_grad_bar, st.njbar, st.det_ddg};
    for (size_t b = 0; b < sizeof bufs / sizeof bufs[0]; b++) free(bufs[b]);
    for (int c = 0; c < 4; c++) {
        free(nbc[c]);
        free(fbc[c]);
    }
}

static void run_mlp(void) {
    /* 16 -> 64 -> 64 -> 16 parameters and normalisation (sizes from the oracle header) */
    const long n = 37;
    double *params = zalloc(sizeof(double) * ORC_MLP_NPARAMS);
    double *grad = zalloc(sizeof(double) * ORC_MLP_NPARAMS);
    double scaling[64];
    float *in = zalloc(sizeof(float) * 16 * n), *tgt = zalloc(sizeof(float) * 16 * n);
    double *out = zalloc(sizeof(double) * 16 * n);
    for (int q = 0; q < ORC_MLP_NPARAMS; q++) params[q] = 0.1 * (urand() - 0.5);
    for (int q = 0; q < 16; q++) {
        scaling[q] = 0.1 * urand();
        scaling[16 + q] = 1.0 + urand();
        scaling[32 + q] = 0.1 * urand();
        scaling[48 + q] = 1.0 + urand();
    }
    for (long q = 0; q < 16 * n; q++) {
        in[q] = (float)(urand() - 0.5);
        tgt[q] = (float)(urand() - 0.5);
    }
    for (int bf = 0; bf < 2; bf++) {
        orc_mlp_forward(params, scaling, in, n, bf, out);
        const double loss = orc_mlp_grad(params, scaling, in, tgt, n, bf, grad);
        printf("network oracle, bf16 %d: loss %.3e\n", bf, loss);
        if (!isfinite(loss)) g_fail++;
    }
    free(params);
    free(grad);
    free(in);
    free(tgt);
    free(out);
}

int main(void) {
    const material le = {ORC_LINELAST, {1e6, 0.3}}, nh = {ORC_NEOHOOKEAN, {5e5, 0.35}},
                   svk = {ORC_STVK, {5e5, 0.3}}, mr = {ORC_MOONEYRIVLIN, {1e5, 2e4, 1e6}},
                   j2 = {ORC_J2, {2e6, 0.3, 2e3, 1e4}}, tait = {ORC_TAIT, {1000.0, 30.0, 7.0, 1e-3}},
                   rigid = {ORC_RIGID, {0.0}};
    const int lin[3] = {1, 1, 1}, quad[3] = {2, 2, 2}, mixed[3] = {2, 1, 2};
    const material singles[] = {le, nh, svk, mr, j2, tait};
    const char *names[] = {"linear elastic", "Neo-Hookean", "St.Venant-Kirchhoff", "Mooney-Rivlin", "J2", "Tait"};
    char tag[96];
    for (int m = 0; m < 6; m++) {
        snprintf(tag, sizeof tag, "%s, tent, PIC", names[m]);
        run_case(tag, lin, ORC_PIC, &singles[m], 1, 0, 0, 10);
    }
    const int *orders[3] = {lin, quad, mixed};
    const char *onames[3] = {"tent", "quadratic", "2 1 2"};
    const char *snames[4] = {"PIC", "FLIP", "TPIC", "APIC"};
    const material two[2] = {nh, rigid};
    for (int o = 0; o < 3; o++)
        for (int s = 0; s < 4; s++) {
            snprintf(tag, sizeof tag, "NH + rigid, %s, %s", onames[o], snames[s]);
            run_case(tag, orders[o], s, two, 2, 0, 0, 8);
        }
    for (int o = 0; o < 3; o++) {
        snprintf(tag, sizeof tag, "NH, %s, FLIP, shifting + F-bar", onames[o]);
        run_case(tag, orders[o], ORC_FLIP, &nh, 1, 1, 1, 8);
        snprintf(tag, sizeof tag, "LE, %s, PIC, F-bar", onames[o]);
        run_case(tag, orders[o], ORC_PIC, &le, 1, 0, 1, 8);
    }
    run_mlp();
    printf(g_fail ? "FAILED (%d)\n" : "OK\n", g_fail);
    return g_fail ? 1 : 0;
}

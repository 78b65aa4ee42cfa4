/*
 * FP64 CPU oracle for the explicit solid MPM time step.
 *
 * TEST INFRASTRUCTURE ONLY. This is a scalar restatement of the reference
 * explicit update-stress-last / MUSL step, written from the reference's
 * semantics. It exists to check the HIP engine; only tests, the smoke check
 * and the CPU-baseline leg of the benchmark may load it. The product path
 * never calls it.
 *
 * Layouts follow the reference (module/dataset.h, module/mesh.h,
 * module/data_io.cpp:53-56):
 *   particles  AoS FP64: coord[n][3], vel[n][3], stress[n][6] (Voigt
 *              xx,yy,zz,yz,xz,xy per module/dataset.h:15), def_grad[n][3][3]
 *   nodes      component blocks [u | v | w] of length nodec (nuc=0,
 *              nvc=nodec, nwc=2*nodec)
 *   node id    i + nxc*j + nxc*nyc*k   (module/shape_function.cpp:34)
 *   cell id    ix + iy*nex + iz*nex*ney (module/material_point.cpp:65)
 *
 * Parity status per function (how each one is pinned) is listed in the
 * comment above its definition in mpm_oracle.c.
 */
#ifndef MPM_ORACLE_H
#define MPM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXMAT 8
#define ORC_MAXPROP 8

/* Transfer schemes, same meaning as MapScheme in module/dataset.h:19. */
enum { ORC_PIC = 0, ORC_FLIP = 1, ORC_TPIC = 2, ORC_APIC = 3 };

/* Constitutive model codes, same as iprop in module/solid/constitutive_model.cpp:287-305,
 * plus J2 plasticity (no reference counterpart). */
enum {
    ORC_RIGID = -1,
    ORC_LINELAST = 0,
    ORC_NEOHOOKEAN = 1,
    ORC_STVK = 2,
    ORC_MOONEYRIVLIN = 3,
    ORC_J2 = 10,
    ORC_TAIT = 20
};

typedef struct orc_state {
    /* grid: single rank, so xymin == xyminw */
    double xmin[3];
    double xmax[3];
    double dx[3];
    int32_t nex[3];   /* cells per axis (xyelem) */
    int32_t idimc[3]; /* 1 = linear tent, 2 = quadratic B-spline */
    int32_t nxc[3];   /* control points per axis (xynodec = nex + idimc) */
    int32_t nelem;
    int32_t nodec;

    /* step parameters */
    double dt;
    double mtol;
    double bb[3];
    int32_t scheme;   /* ORC_PIC / ORC_FLIP */
    int32_t nlstep;
    int32_t shifting; /* 0: particle shifting off (declared default) */
    int32_t pad0;

    /* materials */
    int32_t nmat;
    int32_t iprop[ORC_MAXMAT];
    double mat_prop[ORC_MAXMAT][ORC_MAXPROP];

    /* boundary-condition lists: 0=u, 1=v, 2=w, 3=rigid (all comps) */
    int32_t ibc[4];
    int32_t *nbc[4];
    double *fbc[4];

    /* particles */
    int32_t num;
    double *coord;
    double *vel;
    double *mass;
    double *vol;
    double *vol0;
    double *stress;
    double *def_grad;
    double *det_def_grad;
    double *trac_force; /* may be NULL (zero traction) */
    int32_t *matid;     /* may be NULL (all material 0) */
    double *eps_p;      /* J2 equivalent plastic strain, may be NULL */

    /* binning (module/material_point.cpp:52-77) */
    int32_t *idepf;
    int32_t *idp2p;
    int32_t *numep;

    /* nodes, [u|v|w] blocks */
    double *nmass;
    double *nvof;
    double *nmome;
    double *nforce;
    double *nvel;
    double *naccel;
    double *nvol; /* nodal volume for shifting, may be NULL when shifting=0 */

    /* scratch for the deformation-gradient increment, [num][9] */
    double *ddg;

    /* affine transfer state, [num][3][3] row-major, zero at start
     * (solid_material_point.cpp:36-43): TPIC vel_grad, APIC vel_Bmat in
     * cmat; APIC inv_Dmat in dinv. May be NULL for PIC/FLIP. */
    double *cmat;
    double *dinv;

    /* F-bar projection (fbar_projection.cpp:18-94), off unless fbar != 0:
     * def_grad_bar [num][9] (identity at start), det_def_grad_bar [num] (one
     * at start), nodal scratch njbar [nodec] and per-particle det dF [num]. */
    int32_t fbar;
    int32_t pad1;
    double *def_grad_bar;
    double *det_def_grad_bar;
    double *njbar;
    double *det_ddg;
} orc_state;

/* grid set-up helper: dx = (xmax - xmin) / nex as in datain_para.cpp:87 */
void orc_grid_init(orc_state *st);

int orc_maksf(const orc_state *st, int m, const double xyg[3], int32_t *nc, double *sf, double *dsf);
int orc_maksf_batch(const orc_state *st, int n, const int32_t *cell, const double *xyg, int32_t *nc, double *sf,
                    double *dsf);

/* cell of a position, FP64 floor formula of material_point.cpp:59-65 */
int orc_cell_of(const orc_state *st, const double x[3]);

int orc_bin(orc_state *st);
void orc_p2g(orc_state *st, double facl);
void orc_solve_solid(orc_state *st);
void orc_g2p(orc_state *st, double facl);
double orc_facl(const orc_state *st, int istep);
int orc_step(orc_state *st, int istep);
int orc_run(orc_state *st, int ista, int iend);

void orc_bc_set_val(const orc_state *st, int comp_list, int nn, double *var, double facl);
void orc_bc_set_zero(const orc_state *st, int comp_list, int nn, double *var);

/* Nodal volume of a fully filled grid, MakNodalVol (module/mesh.cpp:160-202):
 * npxye points per cell at the GaussianDistribution offsets (:12-54), each of
 * volume dx dy dz / (npx npy npz). Fills st->nvol (nodec). */
void orc_nodal_vol(orc_state *st, const int32_t npxye[3]);

/* Particle shifting correction, DeltaCorrectionParticleShifting
 * (module/material_point.cpp:174-227), single rank (dbc = 1). The two global
 * sums (an MPI_Allreduce there, so their order follows the decomposition)
 * are taken in blocks of ORC_SUM_BLOCK consecutive entries, each block in
 * index order, then over the blocks in order. delta: [num][3]. */
#define ORC_SUM_BLOCK 256
void orc_shift_delta(const orc_state *st, double *delta);

/* F-bar projection of one step (after the deformation-gradient update of
 * orc_g2p, before the stress update): scales st->ddg and updates
 * def_grad_bar / det_def_grad_bar. Called by orc_g2p when st->fbar is set. */
void orc_fbar(orc_state *st);

/* OpenMP port of orc_run for the CPU baseline (mpm_oracle_omp.c, built into
 * _build/libmpm_oracle_omp.so only): PIC/FLIP, no J2, shifting or F-bar
 * (returns -1000 for those). nthreads <= 0 keeps the OpenMP default. */
int orc_run_omp(orc_state *st, int ista, int iend, int nthreads);

/* Implicit-path transfers (the P2G / G2P halves of the reference's implicit
 * solid step; the nodal displacement increment ndispl, [u|v|w] blocks, is
 * the caller's). Newmark parameters of NewmarkBetaParaSet
 * (module/nonlinear_time_integration.cpp:43-57); Particle2Node
 * (module/solid/implicit/var_trans_solid_implicit.cpp:20-80) with particle
 * accelerations accel [num][3]; Node2Particle (:106-167) for PIC/FLIP
 * (-1 otherwise; -2 with shifting or F-bar), updating accel. */
void orc_newmark_para(double gamma_nb, double beta_nb, double dt, double nb_para[6]);
void orc_implicit_p2g(orc_state *st, const double *accel, double facl);
int orc_implicit_g2p(orc_state *st, double *accel, const double *ndispl, const double nb_para[6], double gamma_nb);

/* Control-point coordinate on one axis (DefineCpPos, module/mesh.h:32-47, one rank). */
double orc_node_coord(const orc_state *st, int axis, int idx);

/* constitutive update, semantics of ConstitutiveModel::UpdateStress */
void orc_update_stress(int model, double stress[6], const double F[9], const double dF[9], const double *prop,
                       double jac, double jac_bar);
/* J2 radial return (Simo & Hughes 1998, Box 3.2), small-strain increment form */
int orc_j2_return(const double prop[8], double stress[6], double *eps_p, const double deps[6], double *dgamma,
                  double *resid);

/* Constitutive-surrogate MLP 16 -> 64 -> 64 -> 16 (ReLU), no reference
 * counterpart. params: 6288 values [w1 64x16 | b1 64 | w2 64x64 | b2 64 |
 * w3 16x64 | b3 16], weights [out][in]; scaling: 64 values [in_mu 16 |
 * in_isd 16 | out_mu 16 | out_sd 16]; in/out/tgt rows of 16 floats.
 * bf16 = 1 rounds weights, inputs, hidden activations and back-propagated
 * activation gradients to bfloat16 where the GPU kernels do; bf16 = 0 is the
 * exact FP64 network. */
#define ORC_MLP_NPARAMS 6288
void orc_mlp_forward(const double *params, const double *scaling, const float *in, long n, int bf16, double *out);
/* gradient of L = 1/(2n) sum_i sum_{o<7} (y_io - t_io)^2 in normalised
 * output units; returns L */
double orc_mlp_grad(const double *params, const double *scaling, const float *in, const float *tgt, long n, int bf16,
                    double *grad);

/* Weakly compressible Newtonian fluid, no reference counterpart:
 * prop = {rho0, c, gamma, mu}; p = (rho0 c^2 / gamma) ((rho/rho0)^gamma - 1)
 * with rho = rho0 / J (Tait), evaluated as B expm1(-gamma log1p(J - 1));
 * sigma = -p I + 2 mu dev(sym(L)), L = (dF - I) / dt. */
void orc_tait_stress(const double *prop, double jac, const double dF[9], double dt, double s[6]);

double orc_det3(const double m[9]);
void orc_inv3(const double m[9], double out[9]);

#ifdef __cplusplus
}
#endif

#endif

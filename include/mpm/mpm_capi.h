/*
 * mpm_capi.h -- C ABI of the MI355X explicit-MPM step engine.
 *
 * Plain C types only: pointers, sizes and status codes. Host buffers are
 * owned by the caller and only read or written during a call; the engine
 * owns its device memory (allocated by mpm_engine_create and
 * mpm_upload_particles, freed by mpm_engine_destroy).
 *
 * Reference layouts honoured at the boundary (reference = HPC-MPM-Code):
 *   particle arrays  AoS FP64 as MaterialPoint keeps them
 *                    (module/material_point.h:113-121): coord[n][3],
 *                    vel[n][3], stress[n][6] in Voigt order xx,yy,zz,yz,xz,xy
 *                    (module/dataset.h:15), def_grad[n][3][3] row-major
 *   node arrays      component blocks [u | v | w] of length nodec
 *                    (nuc = 0, nvc = nodec, nwc = 2*nodec; module/data_io.cpp:53-56)
 *   node id          i + nxc*j + nxc*nyc*k (module/shape_function.cpp:34)
 *   cell id          ix + iy*nex + iz*nex*ney (module/material_point.cpp:65)
 *
 * Every entry point returns MPM_OK (0) or a negative mpm_status. No C++
 * exception crosses this boundary. Calls on one engine must be serialised
 * by the caller; distinct engines are independent.
 */
#ifndef MPM_CAPI_H
#define MPM_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum mpm_status {
    MPM_OK = 0,
    MPM_ERR_INVALID_ARGUMENT = -1,
    MPM_ERR_OUT_OF_GRID = -2,   /* a particle left the background grid */
    MPM_ERR_NONFINITE = -3,     /* NaN/Inf in particle state */
    MPM_ERR_JACOBIAN = -4,      /* det F <= 0 */
    MPM_ERR_HIP = -5,           /* a HIP runtime call failed */
    MPM_ERR_UNSUPPORTED = -6,   /* option not implemented for this precision/shape */
    MPM_ERR_NO_DEVICE = -7,
    MPM_ERR_INTERNAL = -8,      /* an internal consistency check failed (the step placed nothing) */
    MPM_ERR_HOST_MEMORY = -9    /* a host allocation failed */
} mpm_status;

/* Transfer scheme; same meaning and codes as MapScheme (module/dataset.h:19).
 * TPIC and APIC (module/map_and_interpolate.h:57-176) run in the FP64-exact
 * engine only; their P2G reads the node coordinate as xyc[nid][j] where the
 * reference writes xyc[j][nid] (:69, :123), an out-of-bounds read of its
 * vector<array<double,3>>. */
enum { MPM_SCHEME_PIC = 0, MPM_SCHEME_FLIP = 1, MPM_SCHEME_TPIC = 2, MPM_SCHEME_APIC = 3 };

/* Constitutive models; codes -1..3 are the reference's iprop values
 * (module/solid/constitutive_model.cpp:287-305). MPM_MODEL_J2 and
 * MPM_MODEL_TAIT have no reference counterpart: J2 plasticity with linear
 * hardening, prop = {E, nu, sigma_y, H}; weakly compressible Newtonian fluid
 * with the Tait equation of state, prop = {rho0, c, gamma, mu},
 * p = (rho0 c^2 / gamma) ((rho / rho0)^gamma - 1) with rho = rho0 / J,
 * sigma = -p I + 2 mu dev(sym(grad v)). */
enum {
    MPM_MODEL_RIGID = -1,
    MPM_MODEL_LINELAST = 0,
    MPM_MODEL_NEOHOOKEAN = 1,
    MPM_MODEL_STVK = 2,
    MPM_MODEL_MOONEYRIVLIN = 3,
    MPM_MODEL_J2 = 10,
    MPM_MODEL_TAIT = 20
};

/* Arithmetic of the step kernels.
 * MPM_PREC_FP64_EXACT: FP64 throughout, in the reference's summation order
 *   (cells ascending, particles in index order, stencil nodes k,j,i); with
 *   libm-free models (linear elastic) the result equals the reference's
 *   arithmetic bit for bit.
 * MPM_PREC_FP32: FP32 particle/grid arithmetic with compensated (Kahan) node
 *   sums; det F and the particle volume in FP64, FP64 division at
 *   near-empty nodes. Positions are stored as cell index + FP32 offset in
 *   the cell. */
enum { MPM_PREC_FP64_EXACT = 0, MPM_PREC_FP32 = 1 };

typedef struct mpm_grid_desc {
    double xmin[3];    /* xyminw */
    double xmax[3];    /* xymaxw */
    int32_t nelem[3];  /* xyelemw: cells per axis */
    int32_t idimc[3];  /* shape-function order per axis (1 = linear tent, 2 = quadratic B-spline) */
} mpm_grid_desc;

typedef struct mpm_material_desc {
    int32_t model;     /* MPM_MODEL_* (reference iprop) */
    int32_t nprop;
    double prop[8];    /* mat_prop: E, nu (LinElast/NeoHookean/StVK); c1, c2, bulk (Mooney-Rivlin) */
} mpm_material_desc;

typedef struct mpm_step_params {
    double dt;
    double mtol;       /* node mass threshold (input.txt) */
    double bb[3];      /* body force per unit mass (gravity) */
    int32_t scheme;    /* MPM_SCHEME_* (sp.solswitch) */
    int32_t nlstep;    /* load ramp: facl = istep/nlstep for istep <= nlstep, else 1 */
    int32_t precision; /* MPM_PREC_* */
    int32_t relayout_every; /* FP32: physically reorder particles every k steps (0 = never) */
} mpm_step_params;

/* Host view of particle arrays in the reference layout. Optional pointers
 * may be NULL: vel (zero), stress (zero), def_grad (identity), det_def_grad
 * (one), vol (vol0), trac_force (zero), matid (all 0), eps_p (zero). */
typedef struct mpm_particles_view {
    int64_t n;
    double *coord;        /* [n][3] */
    double *vel;          /* [n][3] */
    double *mass;         /* [n]    */
    double *vol;          /* [n]    */
    double *vol0;         /* [n]    */
    double *stress;       /* [n][6] */
    double *def_grad;     /* [n][9] */
    double *det_def_grad; /* [n]    */
    double *trac_force;   /* [n][3] */
    int32_t *matid;       /* [n]    */
    double *eps_p;        /* [n] equivalent plastic strain (J2) */
} mpm_particles_view;

/* Host view of node arrays, [u|v|w] blocks of length nodec. Any pointer may
 * be NULL (not downloaded). */
typedef struct mpm_nodes_view {
    int64_t nodec;
    double *nmass;   /* [nodec]   */
    double *nvof;    /* [nodec]   */
    double *nmome;   /* [3*nodec] */
    double *nforce;  /* [3*nodec] */
    double *nvel;    /* [3*nodec] */
    double *naccel;  /* [3*nodec] */
} mpm_nodes_view;

/* Per-kernel timing of the last mpm_step call (HIP events recorded on the
 * engine's stream between the kernels), milliseconds summed over the steps
 * of that call. In FP64-exact mode only bin_ms, p2g_ms and g2p_upd_ms (the
 * whole G2P pass) are filled. */
typedef struct mpm_kernel_times {
    double bin_ms;        /* cell keys + radix sort + cell ranges (+ relayout) */
    double p2g_ms;        /* p2g_cells_ms + p2g_nodes_ms */
    double p2g_cells_ms;  /* per-cell corner partials */
    double p2g_nodes_ms;  /* node sums + grid update */
    double g2p_vel_ms;    /* musl_cells_ms + musl_nodes_ms */
    double musl_cells_ms; /* G2P velocity + MUSL momentum partials */
    double musl_nodes_ms; /* MUSL node sums + velocity + BCs */
    double g2p_upd_ms;    /* dF, F, J, x, V, stress update */
    double total_ms;
    int32_t steps;
    int32_t active_cells; /* cells holding particles at the last bin */
    int32_t active_nodes; /* nodes of those cells */
    int32_t pad;
} mpm_kernel_times;

typedef struct mpm_engine mpm_engine;

/* Version string of the library ("mpm-mi355x <semver> gfx950"). */
const char *mpm_version(void);
/* Human-readable text of the last error on this thread. */
const char *mpm_last_error(void);

/* Creates an engine on the current HIP device. Replaces the global set-up
 * of datain_para + BuildMesh + BuildControlPoint
 * (work/src_solid/explicit/datain_para.cpp:18-162, module/mesh.cpp:56-158).
 * nmat: 1..1024 materials, selected per particle by matid. */
int mpm_engine_create(const mpm_grid_desc *grid, const mpm_material_desc *mats, int32_t nmat,
                      const mpm_step_params *params, mpm_engine **out);
int mpm_engine_destroy(mpm_engine *eng);

/* What an engine holds: particle count (0 before an upload), grid sizes and
 * the run options. */
typedef struct mpm_engine_info {
    int64_t n;
    int32_t nelem, nodec;
    int32_t precision, scheme;
    int32_t shifting, fbar;
} mpm_engine_info;
int mpm_get_engine_info(mpm_engine *eng, mpm_engine_info *out);

/* Node-list Dirichlet BC; comp 0/1/2 = usbc/vsbc/wsbc, 3 = rigid_bc (all
 * three components). Semantics of BoundaryCondition::BCSetVal/BCSetZero
 * (module/bc.cpp:45-67): velocity set to fbc*facl, acceleration zeroed.
 * Lists are applied in the order u, v, w, rigid; a later entry wins. */
int mpm_set_bc(mpm_engine *eng, int32_t comp, const int32_t *nbc, const double *fbc, int32_t ibc);

/* Copies particles host -> device (replaces InputPointData + the initial
 * state set-up of datain_para.cpp:104-126). */
int mpm_upload_particles(mpm_engine *eng, const mpm_particles_view *p);
/* Copies every non-NULL field of the view device -> host, in the original
 * particle order (mass, vol0, matid and trac_force included, so that the
 * view re-uploads as a restart). The FP32 engine returns its FP32 copies of
 * mass, vol0 and trac_force. Once steps have run, its det_def_grad and vol
 * are det(I + H) and det(I + H) vol0 in FP64 from the stored displacement
 * gradient H, exactly as its update kernel computes them. Particles outside
 * the grid keep the values they had when they left, or that were uploaded. */
int mpm_download_particles(mpm_engine *eng, mpm_particles_view *p);
int mpm_download_nodes(mpm_engine *eng, mpm_nodes_view *nodes);

/* Per-pass entry points (one call = one reference pass on the device):
 *   mpm_bin          MaterialPoint::MeshPointLinklist (material_point.cpp:52-77)
 *   mpm_p2g          Particle2NodeS (var_trans_solid_explicit.cpp:20-99)
 *   mpm_solve_grid   SolveSolid (solve_solid_explicit.cpp:8-31)
 *   mpm_g2p          Node2ParticleS incl. MUSL and the coordinate/stress
 *                    update (var_trans_solid_explicit.cpp:101-311)
 * facl is the load factor of the current step (main.cpp:43-48). */
int mpm_bin(mpm_engine *eng);
int mpm_p2g(mpm_engine *eng, double facl);
int mpm_solve_grid(mpm_engine *eng);
int mpm_g2p(mpm_engine *eng, double facl);

/* Downloads the binning in the reference's linked-list form:
 * idepf[nelem] (first particle per cell or -1), idp2p[n] (next particle or
 * -1), numep[nelem] (count); cell_of[n] may be NULL. */
int mpm_download_bins(mpm_engine *eng, int32_t *idepf, int32_t *idp2p, int32_t *numep, int32_t *cell_of);

/* Runs steps istep0 .. istep0+nsteps-1 of the time loop
 * (work/src_solid/explicit/main.cpp:39-73: bin, P2G, grid, G2P) with
 * everything resident on the device. Returns MPM_ERR_OUT_OF_GRID /
 * MPM_ERR_NONFINITE when the device-side checks fired. */
int mpm_step(mpm_engine *eng, int32_t istep0, int32_t nsteps);

/* Waits for the engine's stream. */
int mpm_sync(mpm_engine *eng);

int mpm_get_kernel_times(mpm_engine *eng, mpm_kernel_times *out);

/* Engine options. MPM_OPT_FULL_NODE_OUTPUT (default 1): the last step of
 * each mpm_step call also writes node momentum, force and volume, so that
 * mpm_download_nodes returns every field; 0 skips that work in every step
 * (FP32 precision only; node arrays are then valid at nodes next to
 * particles). MPM_OPT_LANES_PER_CELL (1, 2, 4, 8; set from the particles per
 * cell at upload): lanes of a wave that fold one cell's particles in the FP32
 * cells kernels. It changes only the FP32 summation order.
 * MPM_OPT_RELAYOUT_EVERY (FP32, >= 0): physically reorder the particle slots
 * into cell order every k steps (0 = never); resets the interval counter.
 * MPM_OPT_GRAPH_BLOCK (default 0 = off): mpm_step records a block of this
 * many steps into a HIP graph once and relaunches it (no launch gaps); used
 * when timing is off, the load ramp is over, no samples are logged and the
 * block is a multiple of twice the relayout interval. Uploads, options and
 * surrogate changes drop the recorded graph.
 * MPM_OPT_FUSED_COUNT (FP32, default 1): the update kernel counts particles
 * per cell for the next binning, which then skips its count pass.
 * MPM_OPT_CELLS_PREFETCH (FP32, default 1): the P2G cells kernel requests the
 * next chunk's particle indices and the next group's cell ranges while a
 * chunk's fields load. MPM_OPT_CELLS_GRID (FP32, default 0 = automatic):
 * workgroups of the cells kernels. Neither changes results.
 * MPM_OPT_X_MERGE (FP32, default 1): the cells kernels add the corner
 * partials of x-neighbouring cells of one group before storing them (half
 * the partials written and read for the tent functions, about a third for
 * the quadratic B-spline). It changes the FP32 summation order, which
 * then depends on how the non-empty cells fall into groups: 0 makes every
 * node sum independent of the grouping (as needed for bit-equal results of
 * scenes packed into one grid and run alone). MPM_OPT_SURROGATE_FUSED
 * (default 1): the constitutive surrogate's network runs inside the update
 * kernel on each wave's own candidate rows; 0 queues the candidates and runs
 * them through a separate network pass. Both give the same results. */
enum {
    MPM_OPT_FULL_NODE_OUTPUT = 1,
    MPM_OPT_LANES_PER_CELL = 2,
    MPM_OPT_RELAYOUT_EVERY = 3,
    MPM_OPT_GRAPH_BLOCK = 4,
    MPM_OPT_FUSED_COUNT = 5,
    MPM_OPT_CELLS_PREFETCH = 6,
    MPM_OPT_CELLS_GRID = 7,
    MPM_OPT_X_MERGE = 8,
    MPM_OPT_SURROGATE_FUSED = 9
};
int mpm_set_option(mpm_engine *eng, int32_t key, int64_t value);
/* Reference particle shifting (DeltaCorrectionParticleShifting,
 * module/material_point.cpp:174-227, applied in update_point_coord_and_stress
 * var_trans_solid_explicit.cpp:159-176), both engines (the FP32 engine forms
 * the excess volume from its FP32 P2G volume and sums in FP64). on = 1
 * computes the nodal volume of the filled grid (MakNodalVol,
 * module/mesh.cpp:160-202) for npxye[3] points per cell and applies the
 * correction in every later step; on = 0 turns it off (the default). The two
 * global sums of the correction are taken in blocks of 256 consecutive
 * entries, then over the blocks (single-rank order; the reference sums per
 * MPI rank). */
int mpm_set_shifting(mpm_engine *eng, int32_t on, const int32_t *npxye);
/* Nodal volume used by the shifting ([nodec]). */
int mpm_download_nodal_volume(mpm_engine *eng, double *nvol);

/* TPIC / APIC particle matrices, host [n][3][3] row-major: cmat is the TPIC
 * vel_grad or the APIC vel_Bmat, dinv the APIC inv_Dmat (NULL to skip).
 * They start at zero on upload (solid_material_point.cpp:36-43); the upload
 * restores them from a restart (solid_data_io.cpp:88-95). */
int mpm_download_affine(mpm_engine *eng, double *cmat, double *dinv);
int mpm_upload_affine(mpm_engine *eng, const double *cmat, const double *dinv);

/* Implicit-path transfers of the reference's implicit solid step, FP64-exact
 * engine, PIC and FLIP (MPM_ERR_UNSUPPORTED otherwise); the implicit solve
 * itself (PETSc there) is the caller's.
 *
 * Particle accelerations, host [n][3] (MaterialPoint::accel,
 * module/material_point.h:113-121); zero after an upload until set.
 *
 * mpm_implicit_p2g replaces SolidMaterialPoint::Particle2Node
 * (module/solid/implicit/var_trans_solid_implicit.cpp:20-80): node mass,
 * volume, momentum sum N m v (VelP2G) and force sum N m a (AccelP2G,
 * module/map_and_interpolate.cpp:12-49) in the reference's summation order;
 * nvel = nmome / nmass, naccel = nforce / nmass where nmass > mtol; velocity
 * BCs set to fbc * facl, acceleration BCs zeroed (module/bc.cpp:45-67).
 * Needs a current binning (mpm_bin). Results: mpm_download_nodes.
 *
 * mpm_implicit_g2p replaces SolidMaterialPoint::Node2Particle (:106-167)
 * for the converged nodal displacement increment ndispl ([u|v|w] blocks of
 * nodec): the Newmark predictor of the node velocity and acceleration
 * (PredictNewmarkBetaVelAndAccel + CommitNodalKinematics,
 * module/nonlinear_time_integration.cpp:59-77) with nb_para the six values of
 * NewmarkBetaParaSet (:43-57); per particle dF = I + sum dN (x) ndispl
 * (ComputeDeltaDefGrad, module/material_point.cpp:127-151), F = dF F, J,
 * V = J V0 and the stress update; for non-rigid materials the displacement,
 * acceleration and (PIC) velocity gathers; CommitParticleKinematics (:79-96)
 * with gamma_nb. Particle shifting and F-bar must be off. Invalidates the
 * binning. */
int mpm_upload_accel(mpm_engine *eng, const double *accel);
int mpm_download_accel(mpm_engine *eng, double *accel);
int mpm_implicit_p2g(mpm_engine *eng, double facl);
int mpm_implicit_g2p(mpm_engine *eng, const double *ndispl, const double *nb_para, double gamma_nb);

/* F-bar projection (cal_def_grad_bar, module/solid/explicit/fbar_projection.cpp:18-94;
 * Fbar_flag, off by default in the reference), both engines. After
 * the deformation-gradient update, N (Jbar det dF) V is gathered to the
 * nodes, divided by nvof, interpolated back as Jbar; dF is scaled by
 * cbrt(Jbar / (Jbar_old det dF)) and Fbar = dF Fbar_old. Stress then uses
 * Fbar, the scaled dF, the true J and Jbar (solid_material_point.cpp:136-152);
 * the volume stays J V0. The reference passes delta_def_grad in a layout
 * its declaration does not accept; this is the projection with
 * delta_def_grad[p][i][j] = dF_ij. Turning it on (or uploading particles
 * while on) starts from Fbar = I, Jbar = 1. The FP32 engine keeps Fbar - I
 * and Jbar - 1 in FP32, averages (Jbar det dF - 1) V at the nodes (nodes
 * with nvof <= mtol take Jbar = 0, as the reference), runs the projection in
 * its stencil kernels (the tent functions included) and refuses it while the
 * constitutive surrogate or per-material time steps are on. */
int mpm_set_fbar(mpm_engine *eng, int32_t on);
/* Fbar [n][3][3] and Jbar [n] (NULL to skip); the upload restores a restart. */
int mpm_download_fbar(mpm_engine *eng, double *def_grad_bar, double *det_def_grad_bar);
int mpm_upload_fbar(mpm_engine *eng, const double *def_grad_bar, const double *det_def_grad_bar);
/* Per-material time steps, no counterpart in the reference (one dt per run
 * there): each particle advances by dt[matid] (nmat values, > 0) in place of
 * the step parameters' dt, which batched scenes packed into one grid use to
 * keep their own CFL steps (one material per scene). FP32 engine with the
 * tent functions and PIC / FLIP only (MPM_ERR_UNSUPPORTED otherwise); the
 * node velocity v + dt a of PIC's G2P is then formed per particle. NULL
 * restores the common dt. */
int mpm_set_material_dt(mpm_engine *eng, const double *dt);
/* Enables/disables per-kernel HIP-event timing inside mpm_step (off by default). */
int mpm_set_timing(mpm_engine *eng, int32_t enabled);

/* MakSf on the device for a batch (module/shape_function.cpp:8-119):
 * cell[n], xyg[n][3] -> nc[n][nen], sf[n][nen], dsf[n][nen][3], FP64,
 * with nen = (idimc0+1)(idimc1+1)(idimc2+1). Host buffers in and out. */
int mpm_maksf_batch(const mpm_grid_desc *grid, int64_t n, const int32_t *cell, const double *xyg, int32_t *nc,
                    double *sf, double *dsf);

/* Stable cell-key sort on the device (the binning kernel alone):
 * xyg[n][3] host FP64 -> cell_of[n] (-1 outside the grid), order[n] =
 * particle indices sorted by (cell, index). */
int mpm_bin_positions(const mpm_grid_desc *grid, int64_t n, const double *xyg, int32_t *cell_of, int32_t *order);

/* Allocation-free variant of mpm_bin_positions; no counterpart in the
 * reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * same stable cell-key sort, with every intermediate in a fixed device
 * array of the library (about 21 MB per GPU, held until the process ends),
 * and no hipMalloc, hipFree, memset or synchronisation.
 * xyg: DEVICE pointer to [n][3] FP64 positions; cell_of, order: DEVICE
 * pointers to [n] int32 outputs, same meaning as in mpm_bin_positions.
 * stream: the hipStream_t to enqueue on (NULL = the null stream); the call
 * returns once the work is enqueued. n <= MPM_NOALLOC_MAX_POSITIONS, else
 * MPM_ERR_INVALID_ARGUMENT. All calls on one GPU share that GPU's array:
 * calls that overlap on one GPU (other streams, other threads) race and must
 * be serialised by the caller. */
#define MPM_NOALLOC_MAX_POSITIONS 1048576
int mpm_bin_positions_noalloc(const mpm_grid_desc *grid, int64_t n, const double *xyg, int32_t *cell_of,
                              int32_t *order, void *stream);

/* Measurement helper with no counterpart in the reference: copies nbytes
 * device-to-device iters times with a plain 16-byte-per-lane copy kernel (one
 * word per thread) and returns the achieved bandwidth (bytes read + written
 * per second / 1e9). Gives the measured HBM ceiling next to the vendor peak,
 * and the reference kernel for scaling rocprof byte counters. */
int mpm_bandwidth_probe(int64_t nbytes, int32_t iters, double *gbps);
/* Measurement helper with no counterpart in the reference: the stream mix of
 * the FP32 particle update over n particles (15 four-byte fields read, 13 of
 * them written back and 6 more written, each field its own array, one
 * particle per thread, non-temporal stores), run iters times over four buffer
 * sets in turn (4 x 84 n bytes of device memory, more than the last-level
 * cache) and returned as achieved GB/s of bytes read + written. The ceiling of
 * a kernel that moves its bytes this way. n in [256, 2^28]. */
int mpm_stream_mix_probe(int64_t n, int32_t iters, double *gbps);

#ifdef __cplusplus
}
#endif

#endif /* MPM_CAPI_H */

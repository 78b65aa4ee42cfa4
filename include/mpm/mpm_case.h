/*
 * Reference case files (single rank): C ABI.
 *
 * Reads and writes the solid-case files of the reference's data generator so
 * that its cases run on the engine without conversion:
 *   input.txt     first case block, layout of SolidGenerator::LoadInput
 *                 (data/generate/solid/solid_generator.cpp:31-92)
 *   griddata.txt  DataGenerator::WriteGridDataFile (data/generate/data_generator.cpp:41-46):
 *                 global mesh header (WriteGlobalMeshHeader,
 *                 data/generate/output_util.cpp:16-37) followed by the u, v, w
 *                 BC lists (BoundaryCondition::BCOutput, module/bc.cpp:30-43),
 *                 read as BoundaryCondition::BCInput does
 *   spdata.txt    point count, coordinates, then "id matid mass vol0" per point
 *                 (solid_generator.cpp:199-206, OutputVector module/data_io.h:40-60)
 * The per-rank files of the MPI partitioner (data/divide) are not read.
 */
#ifndef MPM_CASE_H
#define MPM_CASE_H

#include <stdint.h>

#include "mpm_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mpm_case mpm_case;

typedef struct mpm_case_desc {
    mpm_grid_desc grid;     /* xymin, xymax, xyelem, idimc */
    mpm_step_params step;   /* dt, mtol, bb, scheme (solswitch), nlstep; precision/relayout unset */
    int32_t ista, iend, iout, rstflag;
    int32_t npxye[3];
    int32_t nmat;
    double rho;
    int64_t n;              /* points */
    int32_t ibc[3];         /* u, v, w BC list lengths */
    int32_t pad;
} mpm_case_desc;

/* Any file name may be NULL (that part stays empty). */
int mpm_case_read(const char *input_txt, const char *griddata_txt, const char *spdata_txt, mpm_case **out);
int mpm_case_describe(const mpm_case *c, mpm_case_desc *out);
int mpm_case_materials(const mpm_case *c, mpm_material_desc *mats, int32_t max);
/* Pointers into the case's own arrays (valid until mpm_case_free); coord,
 * mass, vol0, matid are set, the rest NULL. */
int mpm_case_particles(const mpm_case *c, mpm_particles_view *view);
int mpm_case_bc(const mpm_case *c, int32_t comp, const int32_t **nbc, const double **fbc, int32_t *ibc);
/* Writes griddata.txt / spdata.txt in the generator's formats. */
int mpm_case_write(const mpm_case *c, const char *griddata_txt, const char *spdata_txt);
/* Engine for the case: grid, materials, BCs and points uploaded; with
 * shifting = 1 the reference's particle shifting is on,
 * with the case's points per cell for the nodal volume. */
int mpm_case_engine(const mpm_case *c, int32_t precision, int32_t relayout_every, int32_t shifting,
                    mpm_engine **out);
void mpm_case_free(mpm_case *c);

/* Restart files (SolidMaterialPointBase::RestartOutput / RestartInput,
 * module/solid/solid_data_io.cpp:61-141; one rank's <pointfile><rank>_re.txt):
 * count line; coordinates; "id matid surf mass vol0 vol" per point; vel;
 * accel; stress; def_grad; then vel_grad, accel_grad (TPIC) or vel_Bmat,
 * accel_Bmat, inv_Dmat (APIC). det F is read back as vol / vol0. The
 * explicit step keeps no particle acceleration: it is written as zero and
 * ignored on reading, as are the acceleration matrices.
 * digits = 0 writes the reference's layout exactly (setw(15), scientific,
 * 6 digits: 7 significant digits, as the reference's own restarts);
 * digits = 17 widens the fields to 26 characters so that a restart
 * continues bit for bit. */
typedef struct mpm_restart mpm_restart;
int mpm_restart_save(const char *path, const mpm_particles_view *p, int32_t scheme, const double *cmat,
                     const double *dinv, const int32_t *id, const int32_t *surf, int32_t digits);
int mpm_restart_load(const char *path, int32_t scheme, mpm_restart **out);
/* Pointers into the restart's arrays (valid until mpm_restart_free): every
 * field of the view but trac_force and eps_p; cmat/dinv [n][3][3] or NULL
 * for the scheme; id, surf [n]. */
int mpm_restart_view(const mpm_restart *r, mpm_particles_view *p, const double **cmat, const double **dinv,
                     const int32_t **id, const int32_t **surf);
void mpm_restart_free(mpm_restart *r);
/* Engine-level: download and save / load and upload (matrices included). */
int mpm_restart_write(mpm_engine *eng, const char *path, int32_t digits);
int mpm_restart_read(mpm_engine *eng, const char *path);

#ifdef __cplusplus
}
#endif

#endif /* MPM_CASE_H */

/*
 * Batched ensembles of independent scenes: C ABI.
 *
 * No counterpart in the reference (one scene per MPI job there). Many small
 * engines (each one scene, any grid size and particle count) advance
 * together: a block of steps of every scene is captured once into a single
 * HIP graph, with one branch per scene (fork/join across the scenes' own
 * streams), and the graph is relaunched for as many blocks as requested.
 * The autotuner picks the per-size-bucket kernel shape (lanes per cell of
 * the P2G / MUSL cell kernels, particle relayout interval) by timing short
 * eager runs of each candidate.
 *
 * Constraints of a captured block: the load ramp must be over (facl = 1,
 * i.e. istep0 > nlstep of every scene), no sample logging, and in FP32 the
 * block length must be a multiple of twice every scene's relayout interval
 * (so that each replay starts from the same particle buffer).
 */
#ifndef MPM_ENSEMBLE_H
#define MPM_ENSEMBLE_H

#include <stdint.h>

#include "mpm_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mpm_ensemble mpm_ensemble;

typedef struct mpm_tune_result {
    int32_t bucket;
    int32_t lanes_per_cell;
    int32_t relayout_every;
    int32_t scenes;
    double ms_per_step;       /* best candidate, whole bucket */
    double ms_per_step_worst; /* worst candidate, whole bucket */
} mpm_tune_result;

/* engines: n engines with particles uploaded (the caller keeps ownership and
 * must destroy them after the ensemble); bucket: per-engine group id used by
 * the autotuner (may be NULL: one bucket). */
int mpm_ensemble_create(mpm_engine *const *engines, int32_t n, const int32_t *bucket, mpm_ensemble **out);
int mpm_ensemble_destroy(mpm_ensemble *ens);

/* Advance every scene by nsteps without a graph (one stream per scene). */
int mpm_ensemble_step(mpm_ensemble *ens, int32_t istep0, int32_t nsteps, double *elapsed_ms);

/* Capture steps istep0 .. istep0 + block_steps - 1 of all scenes into one graph. */
int mpm_ensemble_capture(mpm_ensemble *ens, int32_t istep0, int32_t block_steps);
/* Launch the captured graph nlaunch times (each launch advances every scene
 * by block_steps; facl stays 1); checks every scene's error flags after. */
int mpm_ensemble_launch(mpm_ensemble *ens, int32_t nlaunch, double *elapsed_ms);

/* For each bucket, time trial_steps eager steps of that bucket's scenes for
 * each candidate (lanes per cell in {1, 2, 4} x relayout interval in
 * {0, 8, 16}) and keep the fastest. The trial steps advance the scenes (they
 * are part of the simulation): steps istep0 .. istep0 + 9 * trial_steps - 1.
 * results: up to max_results entries, one per bucket; *nresults set. */
int mpm_ensemble_autotune(mpm_ensemble *ens, int32_t istep0, int32_t trial_steps, mpm_tune_result *results,
                          int32_t max_results, int32_t *nresults);

#ifdef __cplusplus
}
#endif

#endif /* MPM_ENSEMBLE_H */

/*
 * Learned constitutive surrogate: C ABI.
 *
 * No counterpart in the reference (module/solid/constitutive_model.cpp has no
 * plasticity, no fluid equation of state and no learned model). One network
 * stands in for one part of the engine's stress update, chosen by the
 * engine's materials: the plastic corrector of the J2 radial return when a
 * material is MPM_MODEL_J2, else the Tait pressure of MPM_MODEL_TAIT
 * materials (a pressure closure). Elastic J2 trial states and every other
 * model keep the exact update.
 *
 * J2 rows:
 *   inputs  (13) : strain increment (6, engineering shear, Voigt xx,yy,zz,
 *                  yz,xz,xy), elastic strain of the old stress C^-1 sigma_n (6),
 *                  equivalent plastic strain eps_p (1)
 *   outputs (7)  : plastic stress correction sigma_trial - sigma_{n+1} (6) and
 *                  plastic strain increment (1)
 *   A prediction is accepted only when the discrete return-mapping equations
 *   hold within `tol` (relative to the yield stress): the yield condition at
 *   the predicted state, and the radial-return stress for the predicted
 *   plastic increment. Otherwise that particle takes the exact return
 *   mapping. The network runs inside the update kernel on each wave's own
 *   candidates (MPM_OPT_SURROGATE_FUSED = 1, the default) or in a pass over
 *   the candidates queued by the update kernel (0); both give the same
 *   stresses.
 * Tait rows (pressure closure p(J, rate)):
 *   inputs  (2 of 13): J - 1 and the volumetric strain rate tr(dF - I) / dt
 *   output  (1 of 7) : pressure
 *   The exact Tait pressure is the fallback and the check: a prediction is
 *   accepted where |p_net - p| < tol (|p| + 1e-6 B), B = rho0 c^2 / gamma,
 *   and then replaces p in the stress.
 * Network: 16 -> 64 -> 64 -> 16, ReLU, BF16 weights/activations, FP32
 * accumulation on MFMA; unused inputs and outputs are zero.
 *
 * Parameter vector (FP32, MPM_MLP_NPARAMS values):
 *   [w1 64x16 | b1 64 | w2 64x64 | b2 64 | w3 16x64 | b3 16], weights [out][in].
 * Scaling vector (FP32, 64 values): [in_mu 16 | in_isd 16 | out_mu 16 | out_sd 16];
 * the network sees (x - in_mu) * in_isd and its output y maps to y*out_sd + out_mu.
 * Sample rows are 16 floats (inputs or targets, zero padded).
 */
#ifndef MPM_SURROGATE_H
#define MPM_SURROGATE_H

#include <stdint.h>

#include "mpm_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MPM_MLP_NPARAMS 6288
#define MPM_MLP_NSCALING 64
#define MPM_MLP_ROW 16

typedef struct mpm_train_params {
    int32_t epochs;  /* passes over the samples */
    int32_t batch;   /* samples per Adam step (multiple of 64 recommended) */
    float lr;        /* initial Adam learning rate */
    float lr_final;  /* cosine decay to this rate at the last step */
    uint32_t seed;   /* shuffling and (engine fit) weight initialisation */
    int32_t pad;
} mpm_train_params;

typedef struct mpm_surrogate_stats {
    int64_t samples;     /* samples logged (<= capacity) */
    int64_t logged_steps;
    int64_t evaluated;   /* plastic trial states sent through the network */
    int64_t fallbacks;   /* of those, rejected by the residual check */
    float loss;          /* final training loss (normalised units) */
    float tol;
    int64_t offered;     /* states selected for the log since collect (samples = min(offered, capacity)) */
} mpm_surrogate_stats;

/* Standalone network calls on host buffers (device memory is allocated per
 * call). n rows of MPM_MLP_ROW floats. */
int mpm_mlp_forward(const float *params, const float *scaling, const float *in, int64_t n, float *out);
/* Loss L = 1/(2n) sum over rows and the 7 used outputs of (y - t)^2 in
 * normalised units, and dL/dparams (MPM_MLP_NPARAMS floats). */
int mpm_mlp_gradient(const float *params, const float *scaling, const float *in, const float *tgt, int64_t n,
                     float *grad, float *loss);
/* Adam training on the GPU; params is updated in place; loss_per_epoch
 * (optional) receives tp->epochs values. */
int mpm_mlp_train(float *params, const float *scaling, const float *in, const float *tgt, int64_t n,
                  const mpm_train_params *tp, float *loss_per_epoch);

/* Engine hooks (FP32 engine with the tent functions and PIC / FLIP, and a J2
 * or Tait material; MPM_ERR_UNSUPPORTED otherwise).
 * collect: while on, every mpm_step logs the J2 plastic trial states (or the
 * Tait states) of a
 * deterministic subset of particles (a hash of particle id and step below
 * `rate`) into a device buffer of `capacity` samples. on = 0 stops logging
 * and keeps the samples; a new on = 1 discards them. */
int mpm_surrogate_collect(mpm_engine *e, int32_t on, int64_t capacity, double rate);
/* Sorts the samples by (step, particle id), derives the scaling, initialises
 * and trains the network on the GPU. */
int mpm_surrogate_fit(mpm_engine *e, const mpm_train_params *tp, float *final_loss);
/* Route J2 plastic trial states (or Tait pressures) through the network
 * (on = 1) with tolerance tol; on = 0 restores the exact update. */
int mpm_surrogate_enable(mpm_engine *e, int32_t on, double tol);
int mpm_surrogate_stats_get(mpm_engine *e, mpm_surrogate_stats *out);
int mpm_surrogate_reset_counters(mpm_engine *e);
int mpm_surrogate_get_params(mpm_engine *e, float *params, float *scaling);
int mpm_surrogate_set_params(mpm_engine *e, const float *params, const float *scaling);
/* Copies up to max logged samples (in the fit order once fitted) */
int mpm_surrogate_download_samples(mpm_engine *e, float *in, float *tgt, int64_t max, int64_t *n);

#ifdef __cplusplus
}
#endif

#endif /* MPM_SURROGATE_H */

// Host-side holder of one surrogate network on the device and its training
// loop (Adam over shuffled mini-batches, all arithmetic on the GPU).
#pragma once

#include <vector>

#include "mpm_common.h"
#include "mpm_mlp.h"

#include "../../include/mpm/mpm_surrogate.h"

namespace mpm {

int set_error_msg(int code, const char *msg);
int exception_status();

struct MlpDev {
    float *master = nullptr;   // FP32 parameters [kMlpNParams]
    float *m1 = nullptr, *m2 = nullptr, *grad = nullptr;
    float *scaling = nullptr;  // [64]
    uint16_t *wbf = nullptr;   // BF16 copies of w1 | w2 | w3
    float *partial = nullptr, *loss_partial = nullptr, *loss_hist = nullptr;
    uint32_t *order = nullptr;
    int partial_waves = 0;
    int64_t order_cap = 0, loss_cap = 0;

    int init();
    void release();
    MlpParams view() const;
    int ensure_train(int64_t n, int batch, int64_t steps);
};

// Kaiming-uniform weights, zero biases, from a seed (host)
void mlp_init_params(float *params, uint32_t seed);
int mlp_set(MlpDev &d, const float *params, const float *scaling, hipStream_t s);
int mlp_get(const MlpDev &d, float *params, float *scaling, hipStream_t s);
// scaling from the samples (device rows), deterministic FP64 reduction
int mlp_fit_scaling(const float *d_in, const float *d_tgt, int64_t n, float *scaling, hipStream_t s);
// trains d (its current parameters are the starting point); epoch_loss gets tp.epochs values
int mlp_train(MlpDev &d, const float *d_in, const float *d_tgt, int64_t n, const mpm_train_params &tp,
              std::vector<float> &epoch_loss, hipStream_t s);

}  // namespace mpm

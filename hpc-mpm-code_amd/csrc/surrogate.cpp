// Surrogate network: device holder, scaling, Adam training loop and the
// standalone C-ABI calls of include/mpm/mpm_surrogate.h.
#include "surrogate_host.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace mpm {

namespace {

uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t gcd64(uint64_t a, uint64_t b) {
    while (b) {
        const uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

template <typename T>
int dmalloc(T **p, size_t count) {
    *p = nullptr;
    hipError_t e = hipMalloc((void **)p, (count ? count : 1) * sizeof(T));
    if (e != hipSuccess) return hip_fail(e, "hipMalloc", __FILE__, __LINE__);
    return 0;
}

template <typename T>
void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

}  // namespace

int MlpDev::init() {
    int rc;
    if ((rc = dmalloc(&master, kMlpNParams)) || (rc = dmalloc(&m1, kMlpNParams)) || (rc = dmalloc(&m2, kMlpNParams)) ||
        (rc = dmalloc(&grad, kMlpNParams)) || (rc = dmalloc(&scaling, 64)) ||
        (rc = dmalloc(&wbf, (size_t)kMlpHid * kMlpIn + kMlpHid * kMlpHid + kMlpOut * kMlpHid)))
        return rc;
    return 0;
}

void MlpDev::release() {
    dfree(master);
    dfree(m1);
    dfree(m2);
    dfree(grad);
    dfree(scaling);
    dfree(wbf);
    dfree(partial);
    dfree(loss_partial);
    dfree(loss_hist);
    dfree(order);
    partial_waves = 0;
    order_cap = loss_cap = 0;
}

MlpParams MlpDev::view() const {
    MlpParams p;
    p.w1 = wbf;
    p.w2 = wbf + kMlpHid * kMlpIn;
    p.w3 = wbf + kMlpHid * kMlpIn + kMlpHid * kMlpHid;
    p.b1 = master + kMlpOffB1;
    p.b2 = master + kMlpOffB2;
    p.b3 = master + kMlpOffB3;
    p.in_mu = scaling;
    p.in_isd = scaling + 16;
    p.out_mu = scaling + 32;
    p.out_sd = scaling + 48;
    return p;
}

static int tiles_per_wave_for(int64_t rows) {
    const int64_t tiles = cdiv64(rows, 64);
    return (int)std::max<int64_t>(1, cdiv64(tiles, 1024));
}

int MlpDev::ensure_train(int64_t n, int batch, int64_t steps) {
    int rc;
    const int tpw = tiles_per_wave_for(batch);
    const int waves = cdiv(mlp_grad_waves(batch, tpw), 4) * 4;
    if (waves > partial_waves) {
        dfree(partial);
        dfree(loss_partial);
        if ((rc = dmalloc(&partial, (size_t)waves * kMlpNParams)) || (rc = dmalloc(&loss_partial, waves))) return rc;
        partial_waves = waves;
    }
    if (n > order_cap) {
        dfree(order);
        if ((rc = dmalloc(&order, n))) return rc;
        order_cap = n;
    }
    if (steps > loss_cap) {
        dfree(loss_hist);
        if ((rc = dmalloc(&loss_hist, steps))) return rc;
        loss_cap = steps;
    }
    return 0;
}

void mlp_init_params(float *params, uint32_t seed) {
    uint64_t st = 0xA5A5000000000000ull ^ seed;
    auto uni = [&](float bound) {
        const double u = (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
        return (float)((2.0 * u - 1.0) * bound);
    };
    std::memset(params, 0, sizeof(float) * kMlpNParams);
    const float b1 = std::sqrt(6.f / kMlpInUsed), b2 = std::sqrt(6.f / kMlpHid), b3 = std::sqrt(3.f / kMlpHid);
    for (int j = 0; j < kMlpHid; j++)
        for (int k = 0; k < kMlpInUsed; k++) params[kMlpOffW1 + j * kMlpIn + k] = uni(b1);
    for (int i = 0; i < kMlpHid * kMlpHid; i++) params[kMlpOffW2 + i] = uni(b2);
    for (int o = 0; o < kMlpOutUsed; o++)
        for (int k = 0; k < kMlpHid; k++) params[kMlpOffW3 + o * kMlpHid + k] = uni(b3);
}

int mlp_set(MlpDev &d, const float *params, const float *scaling, hipStream_t s) {
    MPM_HIP_CHECK(hipMemcpyAsync(d.master, params, sizeof(float) * kMlpNParams, hipMemcpyHostToDevice, s));
    MPM_HIP_CHECK(hipMemcpyAsync(d.scaling, scaling, sizeof(float) * 64, hipMemcpyHostToDevice, s));
    launch_mlp_publish(s, d.master, d.view());
    MPM_HIP_CHECK(hipGetLastError());
    MPM_HIP_CHECK(hipStreamSynchronize(s));
    return 0;
}

int mlp_get(const MlpDev &d, float *params, float *scaling, hipStream_t s) {
    if (params)
        MPM_HIP_CHECK(hipMemcpyAsync(params, d.master, sizeof(float) * kMlpNParams, hipMemcpyDeviceToHost, s));
    if (scaling) MPM_HIP_CHECK(hipMemcpyAsync(scaling, d.scaling, sizeof(float) * 64, hipMemcpyDeviceToHost, s));
    MPM_HIP_CHECK(hipStreamSynchronize(s));
    return 0;
}

int mlp_fit_scaling(const float *d_in, const float *d_tgt, int64_t n, float *scaling, hipStream_t s) {
    if (n <= 0) return set_error_msg(MPM_ERR_INVALID_ARGUMENT, "no samples to fit");
    double *part = nullptr;
    float *shift = nullptr;
    int rc;
    if ((rc = dmalloc(&part, (size_t)kMomentBlocks * 32)) || (rc = dmalloc(&shift, 16))) {
        dfree(part);
        return rc;
    }
    std::vector<double> hp((size_t)kMomentBlocks * 32);
    float sh[16];
    const float *rows[2] = {d_in, d_tgt};
    for (int w = 0; w < 2 && !rc; w++) {
        hipError_t e = hipMemcpyAsync(sh, rows[w], sizeof sh, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e == hipSuccess) e = hipMemcpyAsync(shift, sh, sizeof sh, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) {
            launch_moments(s, rows[w], n, shift, part);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(hp.data(), part, sizeof(double) * 32 * kMomentBlocks,
                                                hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            rc = hip_fail(e, "mlp_fit_scaling", __FILE__, __LINE__);
            break;
        }
        for (int f = 0; f < 16; f++) {
            double s1 = 0.0, s2 = 0.0;
            for (int b = 0; b < kMomentBlocks; b++) {
                s1 += hp[((size_t)b * 16 + f) * 2];
                s2 += hp[((size_t)b * 16 + f) * 2 + 1];
            }
            const double m = s1 / (double)n;
            const double var = std::max(0.0, s2 / (double)n - m * m);
            const double sd = std::sqrt(var);
            const double mean = (double)sh[f] + m;
            if (w == 0) {
                scaling[f] = (float)mean;
                scaling[16 + f] = (sd > 0.0 && f < kMlpInUsed) ? (float)(1.0 / sd) : 0.f;
            } else {
                scaling[32 + f] = (float)mean;
                scaling[48 + f] = (sd > 0.0) ? (float)sd : 1.f;
            }
        }
    }
    dfree(part);
    dfree(shift);
    return rc;
}

int mlp_train(MlpDev &d, const float *d_in, const float *d_tgt, int64_t n, const mpm_train_params &tp,
              std::vector<float> &epoch_loss, hipStream_t s) {
    if (n <= 0 || n >= (1ll << 31)) return set_error_msg(MPM_ERR_INVALID_ARGUMENT, "sample count out of range");
    if (tp.epochs <= 0 || tp.batch <= 0 || !(tp.lr > 0.f) || !(tp.lr_final >= 0.f))
        return set_error_msg(MPM_ERR_INVALID_ARGUMENT, "invalid training parameters");
    const int batch = (int)std::min<int64_t>(tp.batch, n);
    const int64_t per_epoch = cdiv64(n, batch);
    const int64_t total = per_epoch * tp.epochs;
    int rc = d.ensure_train(n, batch, total);
    if (rc) return rc;
    MPM_HIP_CHECK(hipMemsetAsync(d.m1, 0, sizeof(float) * kMlpNParams, s));
    MPM_HIP_CHECK(hipMemsetAsync(d.m2, 0, sizeof(float) * kMlpNParams, s));
    const MlpParams p = d.view();
    uint64_t st = 0x7F4A7C15ull ^ ((uint64_t)tp.seed << 20);
    int64_t t = 0;
    for (int ep = 0; ep < tp.epochs; ep++) {
        uint64_t a = 1, c = 0;
        if (n > 1) {
            a = splitmix(st) % (uint64_t)n;
            if (a == 0) a = 1;
            while (gcd64(a, (uint64_t)n) != 1) a = (a + 1) % (uint64_t)n == 0 ? 1 : a + 1;
            c = splitmix(st) % (uint64_t)n;
        }
        launch_affine_perm(s, d.order, n, a, c);
        for (int64_t b = 0; b < per_epoch; b++) {
            const int nb = (int)std::min<int64_t>(batch, n - b * batch);
            const double frac = total > 1 ? (double)t / (double)(total - 1) : 1.0;
            const float lr = (float)(tp.lr_final + 0.5 * (tp.lr - tp.lr_final) * (1.0 + std::cos(M_PI * frac)));
            launch_mlp_grad(s, p, d_in, d_tgt, d.order + b * batch, nb, tiles_per_wave_for(nb), d.partial,
                            d.loss_partial, d.grad, d.loss_hist + t);
            t++;
            launch_mlp_adam(s, d.master, d.m1, d.m2, d.grad, p, lr, (int)std::min<int64_t>(t, 1 << 30));
        }
        MPM_HIP_CHECK(hipGetLastError());
    }
    std::vector<float> lh((size_t)total);
    MPM_HIP_CHECK(hipMemcpyAsync(lh.data(), d.loss_hist, sizeof(float) * total, hipMemcpyDeviceToHost, s));
    MPM_HIP_CHECK(hipStreamSynchronize(s));
    epoch_loss.assign(tp.epochs, 0.f);
    for (int ep = 0; ep < tp.epochs; ep++) {
        double acc = 0.0;
        for (int64_t b = 0; b < per_epoch; b++) {
            const int64_t nb = std::min<int64_t>(batch, n - b * batch);
            acc += (double)lh[(size_t)(ep * per_epoch + b)] * (double)nb;
        }
        epoch_loss[ep] = (float)(acc / (double)n);
    }
    return 0;
}

}  // namespace mpm

using namespace mpm;

namespace {

struct Scratch {
    hipStream_t s = nullptr;
    std::vector<void *> bufs;
    ~Scratch() {
        for (void *p : bufs) (void)hipFree(p);
        if (s) (void)hipStreamDestroy(s);
    }
    template <typename T>
    int alloc(T **p, size_t count) {
        int rc = dmalloc(p, count);
        if (!rc) bufs.push_back((void *)*p);
        return rc;
    }
    int stream() {
        MPM_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        return 0;
    }
};

int check_rows(const void *a, const void *b, int64_t n) {
    if (n < 0 || n >= (1ll << 31)) return set_error_msg(MPM_ERR_INVALID_ARGUMENT, "row count out of range");
    if (n > 0 && (!a || !b)) return set_error_msg(MPM_ERR_INVALID_ARGUMENT, "null argument");
    return 0;
}

}  // namespace

extern "C" {

int mpm_mlp_forward(const float *params, const float *scaling, const float *in, int64_t n, float *out) try {
    if (!params || !scaling) return set_error_msg(MPM_ERR_INVALID_ARGUMENT, "null argument");
    int rc = check_rows(in, out, n);
    if (rc || n == 0) return rc;
    Scratch sc;
    MlpDev d;
    float *din = nullptr, *dout = nullptr;
    if ((rc = sc.stream()) || (rc = d.init()) || (rc = sc.alloc(&din, (size_t)n * 16)) ||
        (rc = sc.alloc(&dout, (size_t)n * 16)) || (rc = mlp_set(d, params, scaling, sc.s))) {
        d.release();
        return rc;
    }
    hipError_t e = hipMemcpyAsync(din, in, sizeof(float) * 16 * n, hipMemcpyHostToDevice, sc.s);
    if (e == hipSuccess) {
        launch_mlp_infer(sc.s, d.view(), din, dout, (int)n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, sizeof(float) * 16 * n, hipMemcpyDeviceToHost, sc.s);
    if (e == hipSuccess) e = hipStreamSynchronize(sc.s);
    d.release();
    if (e != hipSuccess) return hip_fail(e, "mpm_mlp_forward", __FILE__, __LINE__);
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_mlp_gradient(const float *params, const float *scaling, const float *in, const float *tgt, int64_t n,
                     float *grad, float *loss) try {
    if (!params || !scaling || !grad) return set_error_msg(MPM_ERR_INVALID_ARGUMENT, "null argument");
    int rc = check_rows(in, tgt, n);
    if (rc) return rc;
    if (n == 0) return set_error_msg(MPM_ERR_INVALID_ARGUMENT, "no rows");
    Scratch sc;
    MlpDev d;
    float *din = nullptr, *dtg = nullptr, *dloss = nullptr;
    if ((rc = sc.stream()) || (rc = d.init()) || (rc = sc.alloc(&din, (size_t)n * 16)) ||
        (rc = sc.alloc(&dtg, (size_t)n * 16)) || (rc = sc.alloc(&dloss, 1)) ||
        (rc = d.ensure_train(n, (int)n, 1)) || (rc = mlp_set(d, params, scaling, sc.s))) {
        d.release();
        return rc;
    }
    hipError_t e = hipMemcpyAsync(din, in, sizeof(float) * 16 * n, hipMemcpyHostToDevice, sc.s);
    if (e == hipSuccess) e = hipMemcpyAsync(dtg, tgt, sizeof(float) * 16 * n, hipMemcpyHostToDevice, sc.s);
    if (e == hipSuccess) {
        launch_mlp_grad(sc.s, d.view(), din, dtg, nullptr, (int)n, tiles_per_wave_for(n), d.partial, d.loss_partial,
                        d.grad, dloss);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(grad, d.grad, sizeof(float) * kMlpNParams, hipMemcpyDeviceToHost, sc.s);
    float lh = 0.f;
    if (e == hipSuccess) e = hipMemcpyAsync(&lh, dloss, sizeof(float), hipMemcpyDeviceToHost, sc.s);
    if (e == hipSuccess) e = hipStreamSynchronize(sc.s);
    d.release();
    if (e != hipSuccess) return hip_fail(e, "mpm_mlp_gradient", __FILE__, __LINE__);
    if (loss) *loss = lh;
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_mlp_train(float *params, const float *scaling, const float *in, const float *tgt, int64_t n,
                  const mpm_train_params *tp, float *loss_per_epoch) try {
    if (!params || !scaling || !tp) return set_error_msg(MPM_ERR_INVALID_ARGUMENT, "null argument");
    int rc = check_rows(in, tgt, n);
    if (rc) return rc;
    if (n == 0) return set_error_msg(MPM_ERR_INVALID_ARGUMENT, "no rows");
    Scratch sc;
    MlpDev d;
    float *din = nullptr, *dtg = nullptr;
    if ((rc = sc.stream()) || (rc = d.init()) || (rc = sc.alloc(&din, (size_t)n * 16)) ||
        (rc = sc.alloc(&dtg, (size_t)n * 16)) || (rc = mlp_set(d, params, scaling, sc.s))) {
        d.release();
        return rc;
    }
    hipError_t e = hipMemcpyAsync(din, in, sizeof(float) * 16 * n, hipMemcpyHostToDevice, sc.s);
    if (e == hipSuccess) e = hipMemcpyAsync(dtg, tgt, sizeof(float) * 16 * n, hipMemcpyHostToDevice, sc.s);
    if (e != hipSuccess) {
        d.release();
        return hip_fail(e, "mpm_mlp_train", __FILE__, __LINE__);
    }
    std::vector<float> el;
    rc = mlp_train(d, din, dtg, n, *tp, el, sc.s);
    if (!rc) rc = mlp_get(d, params, nullptr, sc.s);
    d.release();
    if (rc) return rc;
    if (loss_per_epoch) std::memcpy(loss_per_epoch, el.data(), sizeof(float) * el.size());
    return 0;
} catch (...) {
    return mpm::exception_status();
}

}  // extern "C"

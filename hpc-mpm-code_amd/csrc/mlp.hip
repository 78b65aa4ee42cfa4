// Learned constitutive surrogate: a 3-layer MLP 16 -> 64 -> 64 -> 16 with
// ReLU (13 inputs and 7 outputs are used; the rest are zero padding), BF16
// weights and activations, FP32 accumulation on the matrix cores
// (v_mfma_f32_16x16x16bf16_1k). No reference counterpart.
//
// Every wave works on a tile of 64 samples held in LDS. A 16x16 output block
// of a layer is one chain of MFMAs over K in steps of 16. Operand layout of
// the 16x16x16 MFMA (wave64): lane l supplies A[row l%16][k 4(l/16)..+3] and
// B[k 4(l/16)..+3][col l%16], and receives D[row 4(l/16)+r][col l%16],
// r = 0..3. Weights are stored [out][in] so both operands are 8-byte
// contiguous reads.
//
// Training: each wave accumulates its gradient in registers over its tiles,
// writes one partial per wave, a second kernel adds the partials in wave
// order (deterministic), and an Adam kernel updates the FP32 master weights
// and their BF16 copies.
#include "mpm_common.h"
#include "mpm_kernels.h"
#include "mpm_mlp.h"

namespace mpm {

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4f mfma16(v4s a, v4s b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// LDS written by some lanes of a wave and read by others: order the accesses
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ uint16_t f2bf(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// A operand: rows row0.., k0.. of a row-major [rows][ld] bf16 matrix
__device__ __forceinline__ v4s lds_a(const uint16_t *M, int ld, int row0, int k0, int lane) {
    return *(const v4s *)(M + (row0 + (lane & 15)) * ld + k0 + 4 * (lane >> 4));
}
// B operand from a matrix stored [n][k] row-major (i.e. B^T), n0.., k0..
__device__ __forceinline__ v4s lds_bt(const uint16_t *Mt, int ld, int n0, int k0, int lane) {
    return *(const v4s *)(Mt + (n0 + (lane & 15)) * ld + k0 + 4 * (lane >> 4));
}
// A operand from a matrix stored [k][m] (i.e. A^T), strided gather of 4 values
__device__ __forceinline__ v4s lds_at(const uint16_t *Mt, int ld, int row0, int k0, int lane) {
    const int m = row0 + (lane & 15), k = k0 + 4 * (lane >> 4);
    v4s r;
    r.x = (short)Mt[(k + 0) * ld + m];
    r.y = (short)Mt[(k + 1) * ld + m];
    r.z = (short)Mt[(k + 2) * ld + m];
    r.w = (short)Mt[(k + 3) * ld + m];
    return r;
}
// B operand from a matrix stored [k][n] row-major, strided gather
__device__ __forceinline__ v4s lds_b(const uint16_t *M, int ld, int n0, int k0, int lane) {
    const int n = n0 + (lane & 15), k = k0 + 4 * (lane >> 4);
    v4s r;
    r.x = (short)M[(k + 0) * ld + n];
    r.y = (short)M[(k + 1) * ld + n];
    r.z = (short)M[(k + 2) * ld + n];
    r.w = (short)M[(k + 3) * ld + n];
    return r;
}

// Y[64][N] = act(X[64][K] W^T + b), W stored [N][K]; Y written to LDS (bf16)
// and, when yf != null, the FP32 pre-activation values to yf[64][N].
template <int K, int N, bool RELU>
__device__ __forceinline__ void layer_fwd(const uint16_t *X, const uint16_t *W, const float *b, uint16_t *Y, float *yf,
                                          int lane) {
#pragma unroll
    for (int mi = 0; mi < 4; mi++) {
#pragma unroll
        for (int ni = 0; ni < N / 16; ni++) {
            v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kk = 0; kk < K / 16; kk++)
                acc = mfma16(lds_a(X, K, mi * 16, kk * 16, lane), lds_bt(W, K, ni * 16, kk * 16, lane), acc);
            const int col = ni * 16 + (lane & 15);
            const float bb = b[col];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = mi * 16 + 4 * (lane >> 4) + r;
                float v = acc[r] + bb;
                if (yf) yf[row * N + col] = v;
                if (RELU) v = fmaxf(v, 0.f);
                if (Y) Y[row * N + col] = f2bf(v);
            }
        }
    }
}

// Normalised input row of sample i (raw features, per-feature affine map).
__device__ __forceinline__ void stage_inputs(uint16_t *x, const float *raw, const MlpParams &p, int nvalid, int lane) {
    // lane = sample row
#pragma unroll
    for (int f = 0; f < kMlpIn; f++) {
        float v = 0.f;
        if (lane < nvalid && f < kMlpInUsed) v = (raw[f] - p.in_mu[f]) * p.in_isd[f];
        x[lane * kMlpIn + f] = f2bf(v);
    }
}

// ------------------------------------------------------------------ inference
// One sample per lane; the same wave routine the fused constitutive update uses.
__global__ void __launch_bounds__(256) k_mlp_infer(MlpParams p, const float *__restrict__ in, float *__restrict__ out,
                                                   int n) {
    __shared__ MlpInferSmem sw;
    __shared__ __align__(16) uint16_t xs[4][kMlpWaveScratchBytes / 2];
    mlp_load_infer_smem(sw, p);
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int s0 = (blockIdx.x * 4 + w) * 64; s0 < n; s0 += gridDim.x * 256) {
        const int i = s0 + lane;
        float xn[kMlpIn];
#pragma unroll
        for (int f = 0; f < kMlpIn; f++)
            xn[f] = (i < n && f < kMlpInUsed) ? (in[(size_t)i * kMlpIn + f] - p.in_mu[f]) * p.in_isd[f] : 0.f;
        float y[kMlpOutUsed];
        mlp_wave_infer(sw, xn, y, xs[w], lane);
        if (i < n) {
#pragma unroll
            for (int o = 0; o < kMlpOut; o++)
                out[(size_t)i * kMlpOut + o] = (o < kMlpOutUsed) ? y[o] * p.out_sd[o] + p.out_mu[o] : 0.f;
        }
    }
}

// ------------------------------------------------------------------ training
struct TrainLds {
    uint16_t w1[kMlpHid * kMlpIn];
    uint16_t w2[kMlpHid * kMlpHid];
    uint16_t w3[kMlpOut * kMlpHid];
    float b1[kMlpHid], b2[kMlpHid], b3[kMlpOut];
};

struct TrainWave {
    uint16_t x[64 * kMlpIn];
    uint16_t h1[64 * kMlpHid];
    uint16_t h2[64 * kMlpHid];
    uint16_t d[64 * kMlpHid];    // gradient w.r.t. the current layer's pre-activation
    uint16_t dy[64 * kMlpOut];
    float y[64 * kMlpOut];
};

// dW[N][K] += dA^T X over the 64 rows: A operand = dA^T ([N][64], stored as dA
// [64][N]), B operand = X ([64][K]); acc blocks [N/16][K/16].
template <int N, int K>
__device__ __forceinline__ void grad_w(const uint16_t *dA, const uint16_t *X, v4f (*acc)[K / 16], int lane) {
#pragma unroll
    for (int ni = 0; ni < N / 16; ni++)
#pragma unroll
        for (int ki = 0; ki < K / 16; ki++)
#pragma unroll
            for (int kk = 0; kk < 4; kk++)  // batch of 64 in 4 steps of 16
                acc[ni][ki] = mfma16(lds_at(dA, N, ni * 16, kk * 16, lane), lds_b(X, K, ki * 16, kk * 16, lane),
                                     acc[ni][ki]);
}

// dX[64][K] = dA[64][N] W[N][K]; B operand = W as [k=N][n=K] row-major
template <int N, int K>
__device__ __forceinline__ void grad_x(const uint16_t *dA, const uint16_t *W, const uint16_t *Hmask, uint16_t *dX,
                                       int lane) {
#pragma unroll
    for (int mi = 0; mi < 4; mi++)
#pragma unroll
        for (int ki = 0; ki < K / 16; ki++) {
            v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int nn = 0; nn < N / 16; nn++)
                acc = mfma16(lds_a(dA, N, mi * 16, nn * 16, lane), lds_b(W, K, ki * 16, nn * 16, lane), acc);
            const int col = ki * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = mi * 16 + 4 * (lane >> 4) + r;
                // ReLU: pass the gradient where the forward activation was positive
                const float v = (bf2f(Hmask[row * K + col]) > 0.f) ? acc[r] : 0.f;
                dX[row * K + col] = f2bf(v);
            }
        }
}

// column sums of a [64][N] bf16 matrix into FP32 (bias gradients), added to g
template <int N>
__device__ __forceinline__ void grad_b(const uint16_t *dA, float *g, int lane) {
    for (int c = lane; c < N; c += 64) {
        float s = 0.f;
        for (int r = 0; r < 64; r++) s += bf2f(dA[r * N + c]);
        g[c] += s;
    }
}

__global__ void __launch_bounds__(256) k_mlp_grad(MlpParams p, const float *__restrict__ in,
                                                  const float *__restrict__ tgt, const uint32_t *__restrict__ order,
                                                  int n, int tiles_per_wave, float *__restrict__ partial,
                                                  float *__restrict__ loss_partial) {
    __shared__ TrainLds sw;
    __shared__ TrainWave wb[4];
    __shared__ float gb[4][kMlpHid + kMlpHid + kMlpOut];
    for (int i = threadIdx.x; i < kMlpHid * kMlpIn; i += blockDim.x) sw.w1[i] = p.w1[i];
    for (int i = threadIdx.x; i < kMlpHid * kMlpHid; i += blockDim.x) sw.w2[i] = p.w2[i];
    for (int i = threadIdx.x; i < kMlpOut * kMlpHid; i += blockDim.x) sw.w3[i] = p.w3[i];
    for (int i = threadIdx.x; i < kMlpHid; i += blockDim.x) {
        sw.b1[i] = p.b1[i];
        sw.b2[i] = p.b2[i];
    }
    for (int i = threadIdx.x; i < kMlpOut; i += blockDim.x) sw.b3[i] = p.b3[i];
    for (int i = threadIdx.x; i < 4 * (2 * kMlpHid + kMlpOut); i += blockDim.x) (&gb[0][0])[i] = 0.f;
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wave = blockIdx.x * 4 + w;
    TrainWave &b = wb[w];
    float *gbw = gb[w];
    v4f g1[kMlpHid / 16][kMlpIn / 16], g2[kMlpHid / 16][kMlpHid / 16], g3[kMlpOut / 16][kMlpHid / 16];
#pragma unroll
    for (int i = 0; i < kMlpHid / 16; i++) {
        for (int j = 0; j < kMlpIn / 16; j++) g1[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
        for (int j = 0; j < kMlpHid / 16; j++) g2[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int j = 0; j < kMlpHid / 16; j++) g3[0][j] = v4f{0.f, 0.f, 0.f, 0.f};
    float loss = 0.f;
    const float inv_n = 1.f / (float)n;
    for (int t = 0; t < tiles_per_wave; t++) {
        const int s0 = (wave * tiles_per_wave + t) * 64;
        if (s0 >= n) break;
        const int nvalid = min(64, n - s0);
        const uint32_t sidx = (lane < nvalid) ? (order ? order[s0 + lane] : (uint32_t)(s0 + lane)) : 0u;
        float raw[kMlpIn];
#pragma unroll
        for (int f = 0; f < kMlpIn; f++) raw[f] = (lane < nvalid) ? in[(size_t)sidx * kMlpIn + f] : 0.f;
        stage_inputs(b.x, raw, p, nvalid, lane);
        wave_lds_sync();
        layer_fwd<kMlpIn, kMlpHid, true>(b.x, sw.w1, sw.b1, b.h1, nullptr, lane);
        wave_lds_sync();
        layer_fwd<kMlpHid, kMlpHid, true>(b.h1, sw.w2, sw.b2, b.h2, nullptr, lane);
        wave_lds_sync();
        layer_fwd<kMlpHid, kMlpOut, false>(b.h2, sw.w3, sw.b3, nullptr, b.y, lane);
        wave_lds_sync();
        // loss = 1/(2n) sum over samples and used outputs of (y - t)^2 in normalised units
#pragma unroll
        for (int o = 0; o < kMlpOut; o++) {
            float d = 0.f;
            if (lane < nvalid && o < kMlpOutUsed) {
                const float tn = (tgt[(size_t)sidx * kMlpOut + o] - p.out_mu[o]) / p.out_sd[o];
                d = b.y[lane * kMlpOut + o] - tn;
                loss += 0.5f * d * d * inv_n;
                d *= inv_n;
            }
            b.dy[lane * kMlpOut + o] = f2bf(d);
        }
        wave_lds_sync();
        grad_b<kMlpOut>(b.dy, gbw + 2 * kMlpHid, lane);
        grad_w<kMlpOut, kMlpHid>(b.dy, b.h2, g3, lane);
        grad_x<kMlpOut, kMlpHid>(b.dy, sw.w3, b.h2, b.d, lane);  // dA2
        wave_lds_sync();
        grad_b<kMlpHid>(b.d, gbw + kMlpHid, lane);
        grad_w<kMlpHid, kMlpHid>(b.d, b.h1, g2, lane);
        // dA1 = (dA2 W2) masked by h1 > 0; reuse h2 as the output buffer (h2 no longer needed)
        grad_x<kMlpHid, kMlpHid>(b.d, sw.w2, b.h1, b.h2, lane);
        wave_lds_sync();
        grad_b<kMlpHid>(b.h2, gbw, lane);
        grad_w<kMlpHid, kMlpIn>(b.h2, b.x, g1, lane);
        wave_lds_sync();
    }
    // write this wave's partial gradient: [w1 | b1 | w2 | b2 | w3 | b3]
    float *gp = partial + (size_t)wave * kMlpNParams;
    const int q4 = 4 * (lane >> 4), c16 = lane & 15;
#pragma unroll
    for (int ni = 0; ni < kMlpHid / 16; ni++) {
#pragma unroll
        for (int r = 0; r < 4; r++) gp[kMlpOffW1 + (ni * 16 + q4 + r) * kMlpIn + c16] = g1[ni][0][r];
#pragma unroll
        for (int ki = 0; ki < kMlpHid / 16; ki++)
#pragma unroll
            for (int r = 0; r < 4; r++) gp[kMlpOffW2 + (ni * 16 + q4 + r) * kMlpHid + ki * 16 + c16] = g2[ni][ki][r];
    }
#pragma unroll
    for (int ki = 0; ki < kMlpHid / 16; ki++)
#pragma unroll
        for (int r = 0; r < 4; r++) gp[kMlpOffW3 + (q4 + r) * kMlpHid + ki * 16 + c16] = g3[0][ki][r];
    for (int i = lane; i < kMlpHid; i += 64) {
        gp[kMlpOffB1 + i] = gbw[i];
        gp[kMlpOffB2 + i] = gbw[kMlpHid + i];
    }
    for (int i = lane; i < kMlpOut; i += 64) gp[kMlpOffB3 + i] = gbw[2 * kMlpHid + i];
    // wave loss
    for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o, 64);
    if (lane == 0) loss_partial[wave] = loss;
}

// g[i] = sum over waves in order of partial[w][i]
__global__ void k_mlp_reduce(const float *__restrict__ partial, int nwaves, float *__restrict__ g,
                             const float *__restrict__ loss_partial, float *__restrict__ loss) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kMlpNParams) {
        float s = 0.f;
        for (int w = 0; w < nwaves; w++) s += partial[(size_t)w * kMlpNParams + i];
        g[i] = s;
    }
    if (i == 0 && loss) {
        float s = 0.f;
        for (int w = 0; w < nwaves; w++) s += loss_partial[w];
        *loss = s;
    }
}

// Adam on the FP32 master parameters; refreshes the BF16 weight copies.
__global__ void k_mlp_adam(float *__restrict__ master, float *__restrict__ m1, float *__restrict__ m2,
                           const float *__restrict__ g, MlpParams p, float lr, float beta1, float beta2, float eps,
                           float bc1, float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kMlpNParams) return;
    const float gi = g[i];
    const float a = beta1 * m1[i] + (1.f - beta1) * gi;
    const float b = beta2 * m2[i] + (1.f - beta2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    const float v = master[i] - lr * (a / bc1) / (sqrtf(b / bc2) + eps);
    master[i] = v;
    mlp_publish_param(p, i, v);
}

__global__ void k_mlp_publish(const float *__restrict__ master, MlpParams p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kMlpNParams) mlp_publish_param(p, i, master[i]);
}

__global__ void k_affine_perm(uint32_t *__restrict__ order, int64_t n, uint64_t a, uint64_t c) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) order[i] = (uint32_t)(((unsigned __int128)a * (uint64_t)i + c) % (uint64_t)n);
}

__global__ void k_gather_rows(const float4 *__restrict__ src, const uint32_t *__restrict__ idx, int64_t n,
                              float4 *__restrict__ dst) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 4 * n) return;
    const int64_t i = t >> 2;
    dst[t] = src[(int64_t)idx[i] * 4 + (t & 3)];
}

// 256 threads: column f = t & 15, row phase t >> 4; fixed-order block reduction
__global__ void __launch_bounds__(256) k_moments(const float *__restrict__ rows, int64_t n,
                                                 const float *__restrict__ shift, double *__restrict__ partial) {
    __shared__ double red[2][256];
    const int f = threadIdx.x & 15, ph = threadIdx.x >> 4;
    const float sh = shift[f];
    double s1 = 0.0, s2 = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * 16 + ph; r < n; r += (int64_t)gridDim.x * 16) {
        const double v = (double)rows[r * 16 + f] - (double)sh;
        s1 += v;
        s2 += v * v;
    }
    red[0][threadIdx.x] = s1;
    red[1][threadIdx.x] = s2;
    __syncthreads();
    if (threadIdx.x < 16) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < 16; k++) {
            a += red[0][k * 16 + f];
            b += red[1][k * 16 + f];
        }
        partial[((size_t)blockIdx.x * 16 + f) * 2] = a;
        partial[((size_t)blockIdx.x * 16 + f) * 2 + 1] = b;
    }
}

}  // namespace

__device__ void mlp_publish_param(const MlpParams &p, int i, float v) {
    uint16_t *w1 = const_cast<uint16_t *>(p.w1), *w2 = const_cast<uint16_t *>(p.w2), *w3 = const_cast<uint16_t *>(p.w3);
    float *b1 = const_cast<float *>(p.b1), *b2 = const_cast<float *>(p.b2), *b3 = const_cast<float *>(p.b3);
    if (i < kMlpOffB1) w1[i - kMlpOffW1] = f2bf(v);
    else if (i < kMlpOffW2) b1[i - kMlpOffB1] = v;
    else if (i < kMlpOffB2) w2[i - kMlpOffW2] = f2bf(v);
    else if (i < kMlpOffW3) b2[i - kMlpOffB2] = v;
    else if (i < kMlpOffB3) w3[i - kMlpOffW3] = f2bf(v);
    else b3[i - kMlpOffB3] = v;
}

void launch_mlp_infer(hipStream_t s, const MlpParams &p, const float *in, float *out, int n) {
    if (n <= 0) return;
    int blocks = cdiv(cdiv(n, 64), 4);
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(256), 0, s, p, in, out, n);
}

int mlp_grad_waves(int n, int tiles_per_wave) { return cdiv(cdiv(n, 64), tiles_per_wave); }

void launch_mlp_grad(hipStream_t s, const MlpParams &p, const float *in, const float *tgt, const uint32_t *order, int n,
                     int tiles_per_wave, float *partial, float *loss_partial, float *g, float *loss) {
    const int nwaves = mlp_grad_waves(n, tiles_per_wave);
    const int blocks = cdiv(nwaves, 4);
    hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(256), 0, s, p, in, tgt, order, n, tiles_per_wave, partial,
                       loss_partial);
    // waves of the last block beyond nwaves wrote zero partials; include them (zeros add nothing)
    hipLaunchKernelGGL(k_mlp_reduce, dim3(cdiv(kMlpNParams, 256)), dim3(256), 0, s, (const float *)partial,
                       blocks * 4, g, (const float *)loss_partial, loss);
}

void launch_mlp_publish(hipStream_t s, const float *master, const MlpParams &p) {
    hipLaunchKernelGGL(k_mlp_publish, dim3(cdiv(kMlpNParams, 256)), dim3(256), 0, s, master, p);
}

void launch_affine_perm(hipStream_t s, uint32_t *order, int64_t n, uint64_t a, uint64_t c) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_affine_perm, dim3((unsigned)cdiv64(n, 256)), dim3(256), 0, s, order, n, a, c);
}

void launch_gather_rows(hipStream_t s, const float *src, const uint32_t *idx, int64_t n, float *dst) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)cdiv64(4 * n, 256)), dim3(256), 0, s, (const float4 *)src, idx, n,
                       (float4 *)dst);
}

void launch_moments(hipStream_t s, const float *rows, int64_t n, const float *shift, double *partial) {
    hipLaunchKernelGGL(k_moments, dim3(kMomentBlocks), dim3(256), 0, s, rows, n, shift, partial);
}

void launch_mlp_adam(hipStream_t s, float *master, float *m1, float *m2, const float *g, const MlpParams &p, float lr,
                     int t) {
    const float beta1 = 0.9f, beta2 = 0.999f;
    const float bc1 = 1.f - powf(beta1, (float)t), bc2 = 1.f - powf(beta2, (float)t);
    hipLaunchKernelGGL(k_mlp_adam, dim3(cdiv(kMlpNParams, 256)), dim3(256), 0, s, master, m1, m2, g, p, lr, beta1,
                       beta2, 1e-8f, bc1, bc2);
}

}  // namespace mpm

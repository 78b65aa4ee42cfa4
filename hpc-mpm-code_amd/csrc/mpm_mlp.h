// Shapes and parameter layout of the constitutive-surrogate MLP.
#pragma once

#include <stdint.h>

namespace mpm {

constexpr int kMlpIn = 16;       // padded input width (13 used)
constexpr int kMlpInUsed = 13;
constexpr int kMlpHid = 64;
constexpr int kMlpOut = 16;      // padded output width (7 used)
constexpr int kMlpOutUsed = 7;

// flat FP32 parameter vector: [w1 (64x16) | b1 (64) | w2 (64x64) | b2 (64) | w3 (16x64) | b3 (16)],
// weights row-major [out][in]
constexpr int kMlpOffW1 = 0;
constexpr int kMlpOffB1 = kMlpOffW1 + kMlpHid * kMlpIn;
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpHid;
constexpr int kMlpOffB2 = kMlpOffW2 + kMlpHid * kMlpHid;
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpHid;
constexpr int kMlpOffB3 = kMlpOffW3 + kMlpOut * kMlpHid;
constexpr int kMlpNParams = kMlpOffB3 + kMlpOut;

// Device view: BF16 weights (bit patterns), FP32 biases, feature scaling
// (x_norm = (x - in_mu) * in_isd; y = y_norm * out_sd + out_mu).
struct MlpParams {
    const uint16_t *w1, *w2, *w3;
    const float *b1, *b2, *b3;
    const float *in_mu, *in_isd, *out_mu, *out_sd;
};

__device__ void mlp_publish_param(const MlpParams &p, int i, float v);

typedef short mlp_v4s __attribute__((ext_vector_type(4)));
typedef float mlp_v4f __attribute__((ext_vector_type(4)));
typedef uint32_t mlp_v2u __attribute__((ext_vector_type(2)));
typedef float mlp_v2f __attribute__((ext_vector_type(2)));
typedef __bf16 mlp_v2bf __attribute__((ext_vector_type(2)));

// Weights of one network staged in LDS by a whole block (bf16 bit patterns,
// [out][in] row-major) and the FP32 biases.
struct MlpSmem {
    uint16_t w1[kMlpHid * kMlpIn];
    uint16_t w2[kMlpHid * kMlpHid];
    uint16_t w3[kMlpOut * kMlpHid];
    float b1[kMlpHid], b2[kMlpHid], b3[kMlpOut];
};

// The same weights staged for inference in MFMA-fragment order, each lane's
// A operand contiguous so that a wave reads a fragment with one
// conflict-free LDS read. Layer 1 (K = 16) runs on the 16x16x16 BF16 MFMA:
// fragment hb holds, for lane l, row 16 hb + l % 16, columns 4 (l / 16) .. + 3.
// Layers 2 and 3 (K = 64) run on gfx950's 16x16x32 BF16 MFMA as two 32-wide
// K blocks. Their K order is permuted so that the accumulator fragments of
// the previous layer are the B operand as they stand: in K block kb, lane
// group q = l / 16 supplies slots 8 q .. 8 q + 7 = hidden units
// 16 (2 kb) + 4 q .. + 3 and 16 (2 kb + 1) + 4 q .. + 3, which are exactly
// the rows that lane holds of hidden blocks 2 kb and 2 kb + 1; the weights'
// columns follow the same permutation.
struct __attribute__((aligned(16))) MlpInferSmem {
    uint16_t f1[4][64][4];
    uint16_t f2[4][2][64][8];  // [output block][K block][lane][8]
    uint16_t f3[2][64][8];     // [K block][lane][8]
    float b1[kMlpHid], b2[kMlpHid], b3[kMlpOut];
};

// Per-wave exchange buffer: 64 samples x 16 bf16 inputs, reused for 64 x 8
// FP32 outputs.
constexpr int kMlpWaveScratchBytes = 64 * kMlpIn * 2;

__device__ __forceinline__ uint16_t mlp_f2bf(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    return (uint16_t)(u >> 16);
}

// two floats to packed bf16 (a in the low half), round to nearest even: one
// v_cvt_pk_bf16_f32 on gfx950; equal to mlp_f2bf for finite values
__device__ __forceinline__ uint32_t mlp_pk_bf16(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((mlp_v2f){a, b}, mlp_v2bf));
}

// four floats to the bf16 operand fragment of one MFMA
__device__ __forceinline__ mlp_v4s mlp_frag_bf16(float a, float b, float c, float d) {
    return __builtin_bit_cast(mlp_v4s, (mlp_v2u){mlp_pk_bf16(a, b), mlp_pk_bf16(c, d)});
}

// ReLU of an accumulator fragment, as the bf16 operand fragment of the next
// layer: rounded first, then the negative halves cleared by a packed 16-bit
// integer max (a bf16 with its sign bit set is a negative int16, and the
// order of non-negative bf16 patterns is their integer order), the same
// values as max(x, 0) before rounding
__device__ __forceinline__ mlp_v4s mlp_relu_frag_bf16(mlp_v4f acc) {
    typedef short v2s __attribute__((ext_vector_type(2)));
    const v2s z = {0, 0};
    const v2s lo = __builtin_elementwise_max(__builtin_bit_cast(v2s, mlp_pk_bf16(acc[0], acc[1])), z);
    const v2s hi = __builtin_elementwise_max(__builtin_bit_cast(v2s, mlp_pk_bf16(acc[2], acc[3])), z);
    return __builtin_bit_cast(mlp_v4s, (mlp_v2u){__builtin_bit_cast(uint32_t, lo), __builtin_bit_cast(uint32_t, hi)});
}

__device__ __forceinline__ void mlp_load_smem(MlpSmem &s, const MlpParams &p) {
    for (int i = threadIdx.x; i < kMlpHid * kMlpIn / 4; i += blockDim.x)
        ((uint2 *)s.w1)[i] = ((const uint2 *)p.w1)[i];
    for (int i = threadIdx.x; i < kMlpHid * kMlpHid / 4; i += blockDim.x)
        ((uint2 *)s.w2)[i] = ((const uint2 *)p.w2)[i];
    for (int i = threadIdx.x; i < kMlpOut * kMlpHid / 4; i += blockDim.x)
        ((uint2 *)s.w3)[i] = ((const uint2 *)p.w3)[i];
    for (int i = threadIdx.x; i < kMlpHid; i += blockDim.x) {
        s.b1[i] = p.b1[i];
        s.b2[i] = p.b2[i];
    }
    for (int i = threadIdx.x; i < kMlpOut; i += blockDim.x) s.b3[i] = p.b3[i];
}

// hidden unit of slot j (0..7) of lane group q in K block kb (the permutation above)
__device__ __forceinline__ int mlp_kperm(int kb, int q, int j) { return 16 * (2 * kb + (j >> 2)) + 4 * q + (j & 3); }

__device__ __forceinline__ void mlp_load_infer_smem(MlpInferSmem &s, const MlpParams &p) {
    for (int u = threadIdx.x; u < 4 * 64; u += blockDim.x) {  // layer 1: row 16 hb + c16 of [64][16]
        const int f = u >> 6, l = u & 63, c16 = l & 15, q4 = 4 * (l >> 4);
        *(uint2 *)s.f1[f][l] = *(const uint2 *)(p.w1 + (16 * f + c16) * kMlpIn + q4);
    }
    for (int u = threadIdx.x; u < (8 + 2) * 64; u += blockDim.x) {
        const int f = u >> 6, l = u & 63, c16 = l & 15, q = l >> 4;
        const int kb = f < 8 ? (f & 1) : f - 8;
        const uint16_t *row = f < 8 ? p.w2 + (16 * (f >> 1) + c16) * kMlpHid : p.w3 + c16 * kMlpHid;
        uint16_t *dst = f < 8 ? s.f2[f >> 1][kb][l] : s.f3[kb][l];
        // two runs of 4 consecutive columns
        *(uint2 *)dst = *(const uint2 *)(row + mlp_kperm(kb, q, 0));
        *(uint2 *)(dst + 4) = *(const uint2 *)(row + mlp_kperm(kb, q, 4));
    }
    for (int i = threadIdx.x; i < kMlpHid; i += blockDim.x) {
        s.b1[i] = p.b1[i];
        s.b2[i] = p.b2[i];
    }
    for (int i = threadIdx.x; i < kMlpOut; i += blockDim.x) s.b3[i] = p.b3[i];
}

typedef __bf16 mlp_v8bf __attribute__((ext_vector_type(8)));
typedef short mlp_v8s __attribute__((ext_vector_type(8)));

// B operand of a 32-wide K block from two 16-row accumulator fragments
__device__ __forceinline__ mlp_v8bf mlp_cat_frag(mlp_v4s lo, mlp_v4s hi) {
    const mlp_v8s v = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    return __builtin_bit_cast(mlp_v8bf, v);
}

__device__ __forceinline__ void mlp_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Forward pass of the 64 samples staged in xs ([64][16] bf16 normalised
// inputs) on the BF16 MFMA (16x16x16 for the first layer, 16x16x32 for the
// others); on return xs holds [64][8] FP32
// normalised outputs. The layers run transposed (hidden x samples) so that
// the accumulator fragment of one layer (lane l: rows 4(l/16)+r, column
// l%16) is exactly the B operand fragment of the next; only the inputs and
// outputs pass through LDS. Must be called by all 64 lanes of the wave after
// the rows are visible to the wave.
__device__ __forceinline__ void mlp_wave_run(const MlpInferSmem &s, uint16_t *xs, int lane) {
    const int c16 = lane & 15, q4 = 4 * (lane >> 4);
    mlp_v4s xb[4];
#pragma unroll
    for (int si = 0; si < 4; si++) xb[si] = *(const mlp_v4s *)(xs + (si * 16 + c16) * kMlpIn + q4);
    mlp_wave_sync();
    float *ys = (float *)xs;  // [64][8]
#pragma unroll
    for (int si = 0; si < 4; si++) {
        mlp_v4s h1[4], h2[4];
        // each accumulator starts at the bias of its rows
#pragma unroll
        for (int hb = 0; hb < 4; hb++) {
            const mlp_v4s a = *(const mlp_v4s *)s.f1[hb][lane];
            const mlp_v4f b0 = *(const mlp_v4f *)(s.b1 + hb * 16 + q4);
            h1[hb] = mlp_relu_frag_bf16(__builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, xb[si], b0, 0, 0, 0));
        }
#pragma unroll
        for (int hb = 0; hb < 4; hb++) {
            mlp_v4f acc = *(const mlp_v4f *)(s.b2 + hb * 16 + q4);
#pragma unroll
            for (int kb = 0; kb < 2; kb++) {
                const mlp_v8bf a = *(const mlp_v8bf *)s.f2[hb][kb][lane];
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, mlp_cat_frag(h1[2 * kb], h1[2 * kb + 1]), acc, 0, 0, 0);
            }
            h2[hb] = mlp_relu_frag_bf16(acc);
        }
        mlp_v4f acc = *(const mlp_v4f *)(s.b3 + q4);
#pragma unroll
        for (int kb = 0; kb < 2; kb++) {
            const mlp_v8bf a = *(const mlp_v8bf *)s.f3[kb][lane];
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, mlp_cat_frag(h2[2 * kb], h2[2 * kb + 1]), acc, 0, 0, 0);
        }
        if (q4 < 8) *(mlp_v4f *)(ys + (si * 16 + c16) * 8 + q4) = acc;
    }
    mlp_wave_sync();
}

// row of 16 normalised inputs as bf16 into a staging row
__device__ __forceinline__ void mlp_stage_row(uint16_t *row, const float xn[kMlpIn]) {
    uint32_t pk[kMlpIn / 2];
#pragma unroll
    for (int f = 0; f < kMlpIn / 2; f++)
        pk[f] = mlp_pk_bf16(xn[2 * f], xn[2 * f + 1]);
    uint4 *r4 = (uint4 *)row;
    r4[0] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
    r4[1] = make_uint4(pk[4], pk[5], pk[6], pk[7]);
}

// one sample per lane: xn this lane's normalised inputs, y its normalised outputs
__device__ __forceinline__ void mlp_wave_infer(const MlpInferSmem &s, const float xn[kMlpIn], float y[kMlpOutUsed],
                                               uint16_t *xs, int lane) {
    mlp_stage_row(xs + lane * kMlpIn, xn);
    mlp_wave_sync();
    mlp_wave_run(s, xs, lane);
    const float *ys = (const float *)xs;
#pragma unroll
    for (int o = 0; o < kMlpOutUsed; o++) y[o] = ys[lane * 8 + o];
    mlp_wave_sync();
}

void launch_mlp_infer(hipStream_t s, const MlpParams &p, const float *in, float *out, int n);
int mlp_grad_waves(int n, int tiles_per_wave);
// partial: [cdiv(waves,4)*4][kMlpNParams], loss_partial: same number of waves
void launch_mlp_grad(hipStream_t s, const MlpParams &p, const float *in, const float *tgt, const uint32_t *order, int n,
                     int tiles_per_wave, float *partial, float *loss_partial, float *g, float *loss);
void launch_mlp_adam(hipStream_t s, float *master, float *m1, float *m2, const float *g, const MlpParams &p, float lr,
                     int t);
void launch_mlp_publish(hipStream_t s, const float *master, const MlpParams &p);
// order[i] = (a i + c) mod n, a coprime to n
void launch_affine_perm(hipStream_t s, uint32_t *order, int64_t n, uint64_t a, uint64_t c);
// dst row i = src row idx[i] (rows of 16 floats)
void launch_gather_rows(hipStream_t s, const float *src, const uint32_t *idx, int64_t n, float *dst);
// per-block FP64 sums of (x - shift) and (x - shift)^2 of each of the 16 columns;
// partial: [kMomentBlocks][16][2]
constexpr int kMomentBlocks = 256;
void launch_moments(hipStream_t s, const float *rows, int64_t n, const float *shift, double *partial);

}  // namespace mpm

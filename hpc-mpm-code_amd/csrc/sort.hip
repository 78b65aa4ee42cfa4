// Cell binning on the device: FP64 cell keys, a stable LSD radix sort of
// (cell key, particle slot) pairs, and per-cell particle ranges.
//
// Stability is what reproduces the reference's traversal order: the linked
// lists of MaterialPoint::MeshPointLinklist (module/material_point.cpp:52-77)
// append particles to their cell in ascending index order, so "cells
// ascending, then particle index ascending" is exactly a stable sort by cell
// of the index sequence.
//
// Radix sort design (wave64, gfx950): 8-bit digits, 256-thread blocks over
// tiles of 4096 keys. Per pass: a histogram kernel (wave-aggregated LDS
// counts), a device-wide exclusive scan of the digit-major histogram table,
// and a scatter kernel that ranks each key stably inside its tile with
// 8 ballots per wave (lanes holding the same digit) plus per-wave digit
// counts in LDS. Keys that are already nearly sorted -- the usual case from
// one time step to the next -- leave most lanes of a wave on one digit, so
// the wave aggregation keeps LDS traffic small.
#include "mpm_common.h"
#include "mpm_kernels.h"

namespace mpm {

namespace {

constexpr int kBlock = 256;
constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems;  // 4096
constexpr int kWaves = kBlock / 64;

__device__ inline uint64_t lanemask_lt(int lane) { return (lane == 0) ? 0ull : (~0ull >> (64 - lane)); }

// Lanes of the calling wave whose 8-bit digit equals this lane's digit,
// restricted to lanes with valid != 0.
__device__ inline uint64_t match_digit8(bool valid, uint32_t d) {
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
        bool bit = (d >> b) & 1u;
        uint64_t bal = __ballot(valid && bit);
        peers &= bit ? bal : ~bal;
    }
    return peers;
}

__device__ inline uint32_t wave_incl_scan(uint32_t v, int lane) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}

// Block-wide exclusive scan of one value per thread; returns the exclusive
// prefix and writes the block total to *total.
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t *total) {
    __shared__ uint32_t wsum[kWaves];
    int t = threadIdx.x, lane = t & 63, w = t >> 6;
    uint32_t inc = wave_incl_scan(v, lane);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kWaves; i++) {
        uint32_t s = wsum[i];
        if (i < w) pre += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return pre + inc - v;
}

__global__ void __launch_bounds__(kBlock) k_scan_reduce(const uint32_t *__restrict__ in, uint32_t *__restrict__ bsum,
                                                         int n) {
    int base = blockIdx.x * kTile;
    uint32_t s = 0;
    for (int i = threadIdx.x; i < kTile; i += kBlock) {
        int idx = base + i;
        if (idx < n) s += in[idx];
    }
    uint32_t tot;
    block_excl_scan(s, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Exclusive scan of a tile of kTile entries in place; adds *offset_src (if
// non-null, indexed by block) to every entry.
__global__ void __launch_bounds__(kBlock) k_scan_tile(uint32_t *__restrict__ data, int n,
                                                       const uint32_t *__restrict__ offset_src) {
    int base = blockIdx.x * kTile + threadIdx.x * kItems;
    uint32_t v[kItems];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < kItems; i++) {
        int idx = base + i;
        v[i] = (idx < n) ? data[idx] : 0u;
        s += v[i];
    }
    uint32_t tot;
    uint32_t pre = block_excl_scan(s, &tot);
    if (offset_src) pre += offset_src[blockIdx.x];
#pragma unroll
    for (int i = 0; i < kItems; i++) {
        int idx = base + i;
        if (idx < n) data[idx] = pre;
        pre += v[i];
    }
}

__global__ void __launch_bounds__(kBlock) k_rs_hist(const uint32_t *__restrict__ keys, int n, int shift,
                                                     uint32_t *__restrict__ hist, int nblocks) {
    __shared__ uint32_t h[256];
    int t = threadIdx.x, lane = t & 63;
    h[t] = 0;
    __syncthreads();
    int base = blockIdx.x * kTile;
#pragma unroll 4
    for (int r = 0; r < kItems; r++) {
        int idx = base + r * kBlock + t;
        bool valid = idx < n;
        uint32_t d = valid ? ((keys[idx] >> shift) & 255u) : 0u;
        uint64_t peers = match_digit8(valid, d);
        if (valid && (peers & lanemask_lt(lane)) == 0) atomicAdd(&h[d], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    hist[t * nblocks + blockIdx.x] = h[t];
}

__global__ void __launch_bounds__(kBlock) k_rs_scatter(const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                        uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                        int n, int shift, const uint32_t *__restrict__ offs,
                                                        int nblocks, int vals_identity) {
    __shared__ uint32_t goff[256];
    __shared__ uint32_t wcnt[kWaves][256];
    int t = threadIdx.x, lane = t & 63, w = t >> 6;
    goff[t] = offs[t * nblocks + blockIdx.x];
#pragma unroll
    for (int i = 0; i < kWaves; i++) wcnt[i][t] = 0;
    __syncthreads();
    int base = blockIdx.x * kTile;
    for (int r = 0; r < kItems; r++) {
        int idx = base + r * kBlock + t;
        bool valid = idx < n;
        uint32_t key = valid ? kin[idx] : 0u;
        uint32_t val = valid ? (vals_identity ? (uint32_t)idx : vin[idx]) : 0u;
        uint32_t d = (key >> shift) & 255u;
        uint64_t peers = match_digit8(valid, d);
        uint32_t rank = (uint32_t)__popcll(peers & lanemask_lt(lane));
        if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pre = goff[d];
            for (int i = 0; i < w; i++) pre += wcnt[i][d];
            uint32_t dest = pre + rank;
            kout[dest] = key;
            vout[dest] = val;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int i = 0; i < kWaves; i++) {
            add += wcnt[i][t];
            wcnt[i][t] = 0;
        }
        goff[t] += add;
        __syncthreads();
    }
}

__global__ void k_cell_ranges(const uint32_t *__restrict__ skeys, int n, int nelem, uint32_t *__restrict__ cstart,
                              uint32_t *__restrict__ cend) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    uint32_t k = skeys[s];
    if (k >= (uint32_t)nelem) return;
    if (s == 0 || skeys[s - 1] != k) cstart[k] = (uint32_t)s;
    if (s == n - 1 || skeys[s + 1] != k) cend[k] = (uint32_t)(s + 1);
}

__global__ void k_fill_u32(uint32_t *__restrict__ p, size_t n, uint32_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void k_keys_f64(GridGeom g, const double *__restrict__ x, size_t ld, int n, uint32_t *__restrict__ keys,
                           uint32_t *__restrict__ err) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    uint32_t k = cell_key_f64(g, x[p], x[ld + p], x[2 * ld + p]);
    keys[p] = k;
    if (k == (uint32_t)g.nelem) atomicOr(err, kErrOutOfGrid);
}


// Plain copy of n 16-byte words (HBM bandwidth probe).
__global__ void __launch_bounds__(256) k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

// The particle-update stream mix (HBM probe): per particle, 15 FP32 fields
// read, 13 of them written back and 6 more written, each field its own array
// of n values (structure of arrays, as the engine's particle state), stores
// non-temporal. Launched over one particle per thread.
constexpr int kMixFields = 21, kMixRead = 15, kMixInPlace = 13;
__global__ void __launch_bounds__(256) k_stream_mix(float *__restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v[kMixRead];
#pragma unroll
    for (int f = 0; f < kMixRead; f++) v[f] = p[f * n + i];
    float s = 0.f;
#pragma unroll
    for (int f = 0; f < kMixRead; f++) s += v[f];
#pragma unroll
    for (int f = 0; f < kMixInPlace; f++) __builtin_nontemporal_store(v[f] * 0.5f + s, &p[f * n + i]);
#pragma unroll
    for (int f = kMixRead; f < kMixFields; f++) __builtin_nontemporal_store(s, &p[f * n + i]);
}

}  // namespace

size_t stream_mix_floats(size_t n) { return (size_t)kMixFields * n; }
double stream_mix_bytes(size_t n) { return 4.0 * (double)(kMixRead + kMixInPlace + (kMixFields - kMixRead)) * (double)n; }
void launch_stream_mix(hipStream_t s, float *p, size_t n) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_stream_mix, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
}

void launch_copy16(hipStream_t s, const void *src, void *dst, size_t n16) {
    // one 16-byte word per thread: blocks dispatched in address order keep the
    // accesses in flight within one window of memory (6.2 TB/s for 1 GiB on
    // MI355X, against 4.6-5.4 TB/s when each thread strides over the buffer)
    size_t nb = (n16 + 255) / 256;
    int blocks = (int)(nb > (size_t)0x7FFFFFFF ? 0x7FFFFFFF : nb);
    hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, s, (const uint4 *)src, (uint4 *)dst, n16);
}

size_t radix_hist_entries(int n) { return (size_t)256 * (size_t)cdiv(n > 0 ? n : 1, kTile); }

int excl_scan_u32(hipStream_t s, uint32_t *data, int n, uint32_t *bsum) {
    if (n <= 0) return 0;
    int nb = cdiv(n, kTile);
    if (nb == 1) {
        hipLaunchKernelGGL(k_scan_tile, dim3(1), dim3(kBlock), 0, s, data, n, (const uint32_t *)nullptr);
        return (int)hipGetLastError();
    }
    if (nb > kTile) return -1;  // > 16.7M entries: not needed by the engine
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kBlock), 0, s, (const uint32_t *)data, bsum, n);
    hipLaunchKernelGGL(k_scan_tile, dim3(1), dim3(kBlock), 0, s, bsum, nb, (const uint32_t *)nullptr);
    hipLaunchKernelGGL(k_scan_tile, dim3(nb), dim3(kBlock), 0, s, data, n, (const uint32_t *)bsum);
    return (int)hipGetLastError();
}

int radix_sort_pairs(hipStream_t s, SortBuffers &b, const uint32_t *keys_in, int n, int nbits, uint32_t **keys_out,
                     uint32_t **vals_out) {
    // pass 0 reads keys_in with identity values; passes then alternate
    // between (keys, vals) and (keys_alt, vals_alt)
    if (n <= 0 || nbits <= 0) {
        *keys_out = b.keys;
        *vals_out = b.vals;
        return 0;
    }
    int nb = cdiv(n, kTile);
    int hist_n = 256 * nb;
    const uint32_t *kin = keys_in, *vin = nullptr;
    uint32_t *kout = b.keys, *vout = b.vals;
    for (int shift = 0; shift < nbits; shift += 8) {
        hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(kBlock), 0, s, kin, n, shift, b.hist, nb);
        int rc = excl_scan_u32(s, b.hist, hist_n, b.bsum);
        if (rc) return rc;
        hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(kBlock), 0, s, kin, vin, kout, vout, n, shift,
                           (const uint32_t *)b.hist, nb, vin == nullptr ? 1 : 0);
        kin = kout;
        vin = vout;
        if (kout == b.keys) {
            kout = b.keys_alt;
            vout = b.vals_alt;
        } else {
            kout = b.keys;
            vout = b.vals;
        }
    }
    *keys_out = const_cast<uint32_t *>(kin);
    *vals_out = const_cast<uint32_t *>(vin);
    return (int)hipGetLastError();
}

void launch_fill_u32(hipStream_t s, uint32_t *p, size_t n, uint32_t v) {
    if (n == 0) return;
    int blocks = (int)((n + 255) / 256);
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, s, p, n, v);
}

void launch_cell_ranges(hipStream_t s, const uint32_t *skeys, int n, int nelem, uint32_t *cstart, uint32_t *cend) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cell_ranges, dim3(cdiv(n, 256)), dim3(256), 0, s, skeys, n, nelem, cstart, cend);
}

void launch_keys_f64(hipStream_t s, const GridGeom &g, const double *x, size_t ld, int n, uint32_t *keys,
                     uint32_t *err) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_keys_f64, dim3(cdiv(n, 256)), dim3(256), 0, s, g, x, ld, n, keys, err);
}


}  // namespace mpm

// Allocation-free stable cell-key sort: the kernels of mpm_bin_positions
// with every intermediate buffer (keys, radix-sort ping-pong buffers, digit
// histograms, scan block sums) in one file-scope device array.
//
// The array is a host-side symbol: its device address differs per GPU and
// is looked up once per device with hipGetSymbolAddress; only that address
// is ever handed to kernels. Every call on a device shares the device's
// array, so calls that overlap on one GPU (two streams, two threads) race
// and must be serialised by the caller.
#include "mpm_common.h"
#include "mpm_kernels.h"

#include <mutex>

namespace mpm {

namespace {

constexpr int kTileKeys = 4096;  // keys per radix-sort block (sort.hip)
constexpr int kMaxDevices = 64;

// Layout in 32-bit words for n <= kNoAllocMaxPositions:
//   keys | sort keys | sort vals | sort keys_alt | sort vals_alt | hist | bsum
constexpr size_t kHistWords = (size_t)256 * (kNoAllocMaxPositions / kTileKeys);
constexpr size_t kScratchWords = (size_t)5 * kNoAllocMaxPositions + kHistWords + kTileKeys;

__device__ __align__(256) uint32_t scratch[kScratchWords];

std::mutex g_mu;
uint32_t *g_ptr[kMaxDevices] = {};

// Device address of this GPU's scratch array (looked up on the first call
// with that device current).
int scratch_ptr(uint32_t **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= kMaxDevices) return (int)hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        g_ptr[dev] = (uint32_t *)p;
    }
    *out = g_ptr[dev];
    return 0;
}

// Cell keys of AoS FP64 positions xyg[n][3] (nelem for a point outside).
__global__ void k_keys_aos_f64(GridGeom g, const double *__restrict__ xyg, int n, uint32_t *__restrict__ keys) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    keys[p] = cell_key_f64(g, xyg[3 * (size_t)p], xyg[3 * (size_t)p + 1], xyg[3 * (size_t)p + 2]);
}

__global__ void k_bin_outputs(const uint32_t *__restrict__ keys, const uint32_t *__restrict__ order, int n,
                              uint32_t nelem, int32_t *__restrict__ cell_of, int32_t *__restrict__ order_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = keys[i];
    cell_of[i] = k >= nelem ? -1 : (int32_t)k;
    order_out[i] = (int32_t)order[i];
}

int bit_len(uint32_t v) {
    int b = 0;
    while (v) {
        b++;
        v >>= 1;
    }
    return b;
}

}  // namespace

int noalloc_bin_positions(const GridGeom &g, int n, const double *xyg, int32_t *cell_of, int32_t *order,
                          hipStream_t s) {
    if (n < 0 || n > kNoAllocMaxPositions) return (int)hipErrorInvalidValue;
    if (n == 0) return 0;
    uint32_t *base = nullptr;
    int rc = scratch_ptr(&base);
    if (rc) return rc;
    const size_t N = kNoAllocMaxPositions;
    uint32_t *keys = base;
    SortBuffers sb{};
    sb.keys = base + N;
    sb.vals = base + 2 * N;
    sb.keys_alt = base + 3 * N;
    sb.vals_alt = base + 4 * N;
    sb.hist = base + 5 * N;
    sb.bsum = base + 5 * N + kHistWords;
    hipLaunchKernelGGL(k_keys_aos_f64, dim3(cdiv(n, 256)), dim3(256), 0, s, g, xyg, n, keys);
    uint32_t *sk = nullptr, *sv = nullptr;
    rc = radix_sort_pairs(s, sb, keys, n, bit_len((uint32_t)g.nelem), &sk, &sv);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bin_outputs, dim3(cdiv(n, 256)), dim3(256), 0, s, (const uint32_t *)keys,
                       (const uint32_t *)sv, n, (uint32_t)g.nelem, cell_of, order);
    return (int)hipGetLastError();
}

}  // namespace mpm

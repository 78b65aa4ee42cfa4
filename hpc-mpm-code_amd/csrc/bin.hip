// Per-step cell binning of the engine: a counting sort by cell key whose
// result is exactly the stable sort by (cell, slot) -- the traversal order of
// the reference's per-cell linked lists (module/material_point.cpp:52-77),
// where slot order is particle-index order.
//
//   1. count      wave-aggregated atomic increments of a per-cell counter
//                 (one atomic per run of equal keys in a wave)
//   2. scan       one exclusive scan of the packed pair (non-empty, count)
//                 per cell gives each cell's start offset and, for non-empty
//                 cells, its index in the active-cell list
//   3. finalize   writes cstart/cend (cend starts as a cursor at cstart), the
//                 active-cell list, the active node-tile flags, and re-zeroes
//                 the counters for the next step -- or keeps them, when the
//                 next update kernel adjusts them for the particles that
//                 change cell (kept counts)
//   4. scatter    wave-aggregated atomic cursor per run; lanes of a run get
//                 consecutive positions in slot order. A run that does not
//                 land at its cell's start means the cell received several
//                 runs (whose relative order depends on scheduling): that
//                 cell is queued once for
//   5. fix-up     sorting its positions by slot index, which restores the
//                 deterministic stable order.
//
// Integer atomics only; counts, offsets and the final order do not depend on
// the order in which waves run.
#include "mpm_common.h"
#include "mpm_kernels.h"

namespace mpm {

namespace {

constexpr int kBlock = 256;
constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems;
constexpr int kWaves = kBlock / 64;

__global__ void __launch_bounds__(256) k_bin_count(const uint32_t *__restrict__ keys, int n, int nelem,
                                                    uint32_t *__restrict__ cnt, uint32_t *__restrict__ tiles,
                                                    int ntiles, BinCounters *ctr) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    for (int i = gid; i < ntiles; i += gridDim.x * blockDim.x) tiles[i] = 0u;
    if (gid == 0) {
        ctr->n_active = 0u;
        ctr->n_fix = 0u;
    }
    for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const int p = base + threadIdx.x;
        const uint32_t k = (p < n) ? keys[p] : 0xFFFFFFFFu;
        int len, head;
        bool h = wave_runs(k, lane, &len, &head);
        if (h && k < (uint32_t)nelem) atomicAdd(&cnt[k], (uint32_t)len);
    }
}

__device__ __forceinline__ uint64_t wave_incl_scan64(uint64_t v, int lane) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint64_t u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}

__device__ __forceinline__ uint64_t block_excl_scan64(uint64_t v, uint64_t *total) {
    __shared__ uint64_t wsum[kWaves];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    uint64_t inc = wave_incl_scan64(v, lane);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint64_t pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kWaves; i++) {
        uint64_t s = wsum[i];
        if (i < w) pre += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return pre + inc - v;
}

__device__ __forceinline__ uint64_t packed(uint32_t c) { return ((uint64_t)(c != 0u) << 32) | (uint64_t)c; }

// Per-cell arrays are padded to a multiple of kTile cells (padding counts
// stay zero), so tiles are processed with 16-byte vector accesses: round r
// of a tile covers cells base + 1024 r + 4 t .. + 3.
__device__ __forceinline__ uint64_t packed4(uint4 v) { return packed(v.x) + packed(v.y) + packed(v.z) + packed(v.w); }

__global__ void __launch_bounds__(kBlock) k_bin_reduce(const uint32_t *__restrict__ cnt,
                                                        uint64_t *__restrict__ bsum) {
    const uint4 *c4 = (const uint4 *)(cnt + (size_t)blockIdx.x * kTile);
    uint64_t s = 0;
#pragma unroll
    for (int r = 0; r < 4; r++) s += packed4(c4[r * kBlock + threadIdx.x]);
    uint64_t tot;
    block_excl_scan64(s, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kBlock) k_bin_top(uint64_t *__restrict__ bsum, int nb, BinCounters *ctr,
                                                     uint32_t *__restrict__ astart, int n, uint32_t *err) {
    const int base = threadIdx.x * kItems;
    uint64_t v[kItems];
    uint64_t s = 0;
#pragma unroll
    for (int i = 0; i < kItems; i++) {
        v[i] = (base + i < nb) ? bsum[base + i] : 0ull;
        s += v[i];
    }
    uint64_t tot;
    uint64_t pre = block_excl_scan64(s, &tot);
#pragma unroll
    for (int i = 0; i < kItems; i++) {
        if (base + i < nb) bsum[base + i] = pre;
        pre += v[i];
    }
    if (threadIdx.x == 0) {
        ctr->stamp += 1u;
        astart[tot >> 32] = (uint32_t)tot;
        // more counted particles than exist: nothing of this binning may be
        // used (the later kernels would index past the permutation)
        const bool bad = (uint64_t)(uint32_t)tot > (uint64_t)n;
        ctr->bad = bad ? 1u : 0u;
        ctr->n_active = bad ? 0u : (uint32_t)(tot >> 32);
        if (bad) atomicOr(err, kErrBinCounts);
    }
}

__device__ __forceinline__ void mark_cell(const GridGeom &g, const TileGeom &tg, uint32_t *tiles, int c) {
    const int nex = g.nex[0], nexy = g.nex[0] * g.nex[1];
    int iz = c / nexy;
    int iy = (c - iz * nexy) / nex;
    int ix = c - iz * nexy - iy * nex;
    // node tiles of the cell's corners (benign identical stores)
#pragma unroll
    for (int k = 0; k < 2; k++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int ii = 0; ii < 2; ii++) {
                int tx = (ix + ii) / kNodeTileX, ty = (iy + j) / kNodeTileY, tz = (iz + k) / kNodeTileZ;
                tiles[tx + tg.ntx * (ty + tg.nty * tz)] = 1u;
            }
}

__global__ void __launch_bounds__(kBlock) k_bin_finalize(GridGeom g, uint32_t *__restrict__ cnt,
                                                          const uint64_t *__restrict__ bsum,
                                                          uint32_t *__restrict__ cstart, uint32_t *__restrict__ cend,
                                                          uint32_t *__restrict__ active, uint32_t *__restrict__ astart,
                                                          uint32_t *__restrict__ tiles, TileGeom tg,
                                                          uint32_t *__restrict__ occ, int keep) {
    const size_t tbase = (size_t)blockIdx.x * kTile;
    uint4 *c4 = (uint4 *)(cnt + tbase);
    uint4 *s4 = (uint4 *)(cstart + tbase);
    uint4 *e4 = (uint4 *)(cend + tbase);
    uint64_t run = bsum[blockIdx.x];
    for (int r = 0; r < 4; r++) {
        const int idx = r * kBlock + threadIdx.x;
        const uint4 v = c4[idx];
        uint64_t tot;
        uint64_t pre = block_excl_scan64(packed4(v), &tot) + run;
        run += tot;
        const uint32_t vv[4] = {v.x, v.y, v.z, v.w};
        uint32_t st[4];
        bool any = false;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            st[q] = (uint32_t)pre;
            if (vv[q]) {
                const int c = (int)(tbase + 4 * idx + q);
                active[(uint32_t)(pre >> 32)] = (uint32_t)c;
                astart[(uint32_t)(pre >> 32)] = st[q];
                mark_cell(g, tg, tiles, c);
                any = true;
            }
            pre += packed(vv[q]);
        }
        const uint4 so = make_uint4(st[0], st[1], st[2], st[3]);
        s4[idx] = so;
        e4[idx] = so;
        if (any && !keep) c4[idx] = make_uint4(0u, 0u, 0u, 0u);
        // occupancy bits: 8 consecutive lanes hold the 32 cells of one word
        const int lane = threadIdx.x & 63;
        uint32_t w = ((vv[0] ? 1u : 0u) | (vv[1] ? 2u : 0u) | (vv[2] ? 4u : 0u) | (vv[3] ? 8u : 0u))
                     << (4 * (lane & 7));
        w |= (uint32_t)__shfl_xor((int)w, 1, 64);
        w |= (uint32_t)__shfl_xor((int)w, 2, 64);
        w |= (uint32_t)__shfl_xor((int)w, 4, 64);
        if ((lane & 7) == 0) occ[(tbase + 4 * (size_t)idx) >> 5] = w;
    }
}

__global__ void __launch_bounds__(256) k_bin_scatter(const uint32_t *__restrict__ keys, int n, int nelem,
                                                      const uint32_t *__restrict__ cstart, uint32_t *__restrict__ cend,
                                                      uint32_t *__restrict__ perm, uint32_t *__restrict__ stamp_arr,
                                                      uint32_t *__restrict__ fixlist, BinCounters *ctr) {
    const int lane = threadIdx.x & 63;
    const uint32_t stamp = ctr->stamp;
    if (ctr->bad) return;
    for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const int p = base + threadIdx.x;
        const uint32_t k = (p < n) ? keys[p] : 0xFFFFFFFFu;
        int len, head;
        bool h = wave_runs(k, lane, &len, &head);
        uint32_t b = 0;
        const bool valid = k < (uint32_t)nelem;
        if (h && valid) {
            b = atomicAdd(&cend[k], (uint32_t)len);
            if (b != cstart[k]) {
                // this cell received more than one run: queue it once for the fix-up
                if (atomicExch(&stamp_arr[k], stamp) != stamp) fixlist[atomicAdd(&ctr->n_fix, 1u)] = k;
            }
        }
        uint32_t bb = (uint32_t)__shfl((int)b, head, 64);
        if (valid) perm[bb + (uint32_t)(lane - head)] = (uint32_t)p;
    }
}

// Sorts the positions of each queued cell by slot index.
__global__ void __launch_bounds__(256) k_bin_fixup(const uint32_t *__restrict__ fixlist, const BinCounters *ctr,
                                                    const uint32_t *__restrict__ cstart,
                                                    const uint32_t *__restrict__ cend, uint32_t *__restrict__ perm,
                                                    uint32_t *__restrict__ err) {
    if (ctr->bad) return;
    const uint32_t nfix = ctr->n_fix;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nfix; i += gridDim.x * blockDim.x) {
        const uint32_t c = fixlist[i];
        const uint32_t s0 = cstart[c], m = cend[c] - s0;
        uint32_t *a = perm + s0;
        if (m <= 16u) {
            uint32_t v[16];
#pragma unroll
            for (int q = 0; q < 16; q++) v[q] = (q < (int)m) ? a[q] : 0xFFFFFFFFu;
#pragma unroll
            for (int r = 0; r < 16; r++) {
#pragma unroll
                for (int q = (r & 1); q + 1 < 16; q += 2) {
                    uint32_t x = v[q], y = v[q + 1];
                    v[q] = min(x, y);
                    v[q + 1] = max(x, y);
                }
            }
#pragma unroll
            for (int q = 0; q < 16; q++)
                if (q < (int)m) a[q] = v[q];
        } else if (m <= 32u) {
            uint32_t v[32];
#pragma unroll
            for (int q = 0; q < 32; q++) v[q] = (q < (int)m) ? a[q] : 0xFFFFFFFFu;
            // odd-even transposition network on 32 registers (sentinels stay at the end)
#pragma unroll
            for (int r = 0; r < 32; r++) {
#pragma unroll
                for (int q = (r & 1); q + 1 < 32; q += 2) {
                    uint32_t x = v[q], y = v[q + 1];
                    v[q] = min(x, y);
                    v[q + 1] = max(x, y);
                }
            }
#pragma unroll
            for (int q = 0; q < 32; q++)
                if (q < (int)m) a[q] = v[q];
        } else if (m <= 4096u) {
            for (uint32_t q = 1; q < m; q++) {
                uint32_t x = a[q];
                uint32_t r = q;
                while (r > 0 && a[r - 1] > x) {
                    a[r] = a[r - 1];
                    r--;
                }
                a[r] = x;
            }
        } else {
            atomicOr(err, kErrCellOverflow);
        }
    }
}

}  // namespace

TileGeom make_tile_geom(const GridGeom &g) {
    TileGeom t;
    t.ntx = cdiv(g.nxc[0], kNodeTileX);
    t.nty = cdiv(g.nxc[1], kNodeTileY);
    t.ntz = cdiv(g.nxc[2], kNodeTileZ);
    t.ntiles = t.ntx * t.nty * t.ntz;
    return t;
}

size_t bin_scan_blocks(int nelem) { return (size_t)cdiv(nelem, kTile); }
size_t bin_cell_capacity(int nelem) { return bin_scan_blocks(nelem) * kTile; }

void launch_bin_count(hipStream_t s, const GridGeom &g, const TileGeom &tg, const uint32_t *keys, int n,
                      BinBuffers &b) {
    int pblocks = cdiv(n > 0 ? n : 1, 256);
    if (pblocks > 8192) pblocks = 8192;
    hipLaunchKernelGGL(k_bin_count, dim3(pblocks), dim3(256), 0, s, keys, n, g.nelem, b.cnt, b.tiles, tg.ntiles, b.ctr);
}

int launch_bin(hipStream_t s, const GridGeom &g, const TileGeom &tg, const uint32_t *keys, int n, BinBuffers &b,
               uint32_t *err, bool counted, bool keep) {
    const int nelem = g.nelem;
    const int nb = cdiv(nelem, kTile);
    if (nb > kTile) return -1;
    int pblocks = cdiv(n > 0 ? n : 1, 256);
    if (pblocks > 8192) pblocks = 8192;
    if (!counted)
        hipLaunchKernelGGL(k_bin_count, dim3(pblocks), dim3(256), 0, s, keys, n, nelem, b.cnt, b.tiles, tg.ntiles,
                           b.ctr);
    hipLaunchKernelGGL(k_bin_reduce, dim3(nb), dim3(kBlock), 0, s, (const uint32_t *)b.cnt, b.bsum);
    hipLaunchKernelGGL(k_bin_top, dim3(1), dim3(kBlock), 0, s, b.bsum, nb, b.ctr, b.astart, n, err);
    hipLaunchKernelGGL(k_bin_finalize, dim3(nb), dim3(kBlock), 0, s, g, b.cnt, (const uint64_t *)b.bsum, b.cstart,
                       b.cend, b.active, b.astart, b.tiles, tg, b.occ, keep ? 1 : 0);
    if (n > 0) {
        hipLaunchKernelGGL(k_bin_scatter, dim3(pblocks), dim3(256), 0, s, keys, n, nelem, (const uint32_t *)b.cstart,
                           b.cend, b.perm, b.stamp, b.fixlist, b.ctr);
        hipLaunchKernelGGL(k_bin_fixup, dim3(256), dim3(256), 0, s, (const uint32_t *)b.fixlist,
                           (const BinCounters *)b.ctr, (const uint32_t *)b.cstart, (const uint32_t *)b.cend, b.perm,
                           err);
    }
    return (int)hipGetLastError();
}

}  // namespace mpm

// Per-step cell binning of the engine: a counting sort by cell key whose
// result is exactly the stable sort by (cell, slot) -- the traversal order of
// the reference's per-cell linked lists (module/material_point.cpp:52-77),
// where slot order is particle-index order.
//
//   1. count      wave-aggregated atomic increments of a per-cell counter
//                 (one atomic per run of equal keys in a wave)
//   2. reduce     the packed pair (non-empty, count) summed per tile of 4096
//                 cells
//   3. finalize   per tile: the sum of the tile totals before it, then one
//                 exclusive scan of the packed pairs gives each cell's start
//                 offset and, for non-empty cells, its index in the
//                 active-cell list; writes cstart/cend (cend starts as a
//                 cursor at cstart), the active-cell list, the active
//                 node-tile flags, and re-zeroes the counters for the next
//                 step -- or keeps them, when the next update kernel adjusts
//                 them for the particles that change cell (kept counts).
//                 A tile that is empty and was empty at the previous binning
//                 is left as it is: its cells still have cstart == cend and
//                 clear occupancy bits, which is all an empty cell needs.
//   4. scatter    wave-aggregated atomic cursor per run; lanes of a run get
//                 consecutive positions in slot order. A run that does not
//                 land at its cell's start means the cell received several
//                 runs (whose relative order depends on scheduling): the cell
//                 is marked with this binning's stamp (a plain store), and
//   5. fix-up     one thread per active cell sorts the positions of a marked
//                 cell by slot index, which restores the deterministic stable
//                 order. No shared queue counter: when many cells are marked
//                 (particles out of cell order), appends to one counter
//                 serialised the scatter.
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
static_assert(kTile == kBinTile, "a scan block covers one binning tile");
constexpr int kWaves = kBlock / 64;

__global__ void __launch_bounds__(256) k_bin_count(const uint32_t *__restrict__ keys, int n, int nelem,
                                                    uint32_t *__restrict__ cnt, uint32_t *__restrict__ tiles,
                                                    int ntiles, BinCounters *ctr) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    for (int i = gid; i < ntiles; i += gridDim.x * blockDim.x) tiles[i] = 0u;
    if (gid == 0) {
        ctr->n_active = 0u;
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

// sum over the block (every thread gets it)
__device__ __forceinline__ uint64_t block_sum64(uint64_t v) {
    __shared__ uint64_t wsum[kWaves];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) wsum[w] = v;
    __syncthreads();
    uint64_t tot = 0;
#pragma unroll
    for (int i = 0; i < kWaves; i++) tot += wsum[i];
    __syncthreads();
    return tot;
}

// Node tiles touched by the corners of cell (ix, iy, iz): tile ranges
// [tx0, tx1] x [ty0, ty1] x [tz0, tz1]; consecutive cells of a row share them
// except where the row crosses a tile boundary.
struct TileSet {
    int tx0, tx1, ty0, ty1, tz0, tz1;
    __device__ bool operator==(const TileSet &o) const {
        return tx0 == o.tx0 && tx1 == o.tx1 && ty0 == o.ty0 && ty1 == o.ty1 && tz0 == o.tz0 && tz1 == o.tz1;
    }
};

// (a cell's stencil spans nodes ix .. ix + idimc; with idimc <= 2 and tiles of
// at least 2 nodes per axis that is at most two tiles per axis)
__device__ __forceinline__ TileSet tile_set(const GridGeom &g, int ix, int iy, int iz) {
    return TileSet{ix / kNodeTileX, (ix + g.idimc[0]) / kNodeTileX, iy / kNodeTileY,
                   (iy + g.idimc[1]) / kNodeTileY, iz / kNodeTileZ, (iz + g.idimc[2]) / kNodeTileZ};
}

__device__ __forceinline__ void mark_tiles(const TileGeom &tg, uint32_t *tiles, const TileSet &t) {
    // (benign identical stores)
#pragma unroll
    for (int k = 0; k < 2; k++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int ii = 0; ii < 2; ii++) {
                const int tx = ii ? t.tx1 : t.tx0, ty = j ? t.ty1 : t.ty0, tz = k ? t.tz1 : t.tz0;
                tiles[tx + tg.ntx * (ty + tg.nty * tz)] = 1u;
            }
}

__global__ void __launch_bounds__(kBlock) k_bin_finalize(GridGeom g, uint32_t *__restrict__ cnt,
                                                          const uint64_t *__restrict__ bsum, int nb,
                                                          uint32_t *__restrict__ tprev, BinCounters *ctr, int n,
                                                          uint32_t *err, uint32_t *__restrict__ cstart,
                                                          uint32_t *__restrict__ cend,
                                                          uint32_t *__restrict__ active, uint32_t *__restrict__ astart,
                                                          uint32_t *__restrict__ tiles, TileGeom tg,
                                                          uint32_t *__restrict__ occ, uint32_t *__restrict__ gs,
                                                          uint32_t gsize, uint32_t *__restrict__ gs2,
                                                          uint32_t gsize2, int keep) {
    const int b = blockIdx.x;
    const uint32_t was = tprev[b];
    const uint64_t mine = bsum[b];
    // a tile left as it is needs no prefix (only the last block forms the totals)
    if (mine == 0ull && was == 0u && b != nb - 1) return;  // uniform over the block
    // tile totals before this tile, and of all tiles
    uint64_t before = 0, all = 0;
    for (int i = threadIdx.x; i < nb; i += kBlock) {
        const uint64_t v = bsum[i];
        all += v;
        before += i < b ? v : 0ull;
    }
    before = block_sum64(before);
    all = block_sum64(all);
    if (b == nb - 1 && threadIdx.x == 0) {
        ctr->stamp += 1u;
        astart[all >> 32] = (uint32_t)all;
        // more counted particles than exist: nothing of this binning may be
        // used (the later kernels would index past the permutation)
        const bool bad = (uint64_t)(uint32_t)all > (uint64_t)n;
        ctr->bad = bad ? 1u : 0u;
        ctr->n_active = bad ? 0u : (uint32_t)(all >> 32);
        if (bad) atomicOr(err, kErrBinCounts);
    }
    if (mine == 0ull && was == 0u) return;  // uniform over the block
    if (threadIdx.x == 0) tprev[b] = mine != 0ull ? 1u : 0u;
    // thread t owns cells c0 .. c0 + 15
    const size_t c0 = (size_t)b * kTile + (size_t)kItems * threadIdx.x;
    uint4 *c4 = (uint4 *)(cnt + c0);
    uint32_t cc[kItems];
#pragma unroll
    for (int r = 0; r < kItems / 4; r++) {
        const uint4 v = c4[r];
        cc[4 * r] = v.x;
        cc[4 * r + 1] = v.y;
        cc[4 * r + 2] = v.z;
        cc[4 * r + 3] = v.w;
    }
    uint64_t own = 0;
#pragma unroll
    for (int q = 0; q < kItems; q++) own += packed(cc[q]);
    uint64_t tot;
    uint64_t pre = block_excl_scan64(own, &tot) + before;
    const int nex = g.nex[0], ney = g.nex[1], nexy = nex * ney;
    int iz = (int)(c0 / (size_t)nexy);
    int iy = (int)((c0 - (size_t)iz * nexy) / (size_t)nex);
    int ix = (int)(c0 - (size_t)iz * nexy - (size_t)iy * nex);
    TileSet last{-1, -1, -1, -1, -1, -1};
    uint32_t st[kItems];
    uint32_t bits = 0, gbits = 0, gbits2 = 0;
#pragma unroll
    for (int q = 0; q < kItems; q++) {
        st[q] = (uint32_t)pre;
        if (cc[q]) {
            const uint32_t a = (uint32_t)(pre >> 32);
            active[a] = (uint32_t)(c0 + q);
            astart[a] = st[q];
            bits |= 1u << q;
            if (a % gsize == 0u) gbits |= 1u << q;
            if (gsize2 && a % gsize2 == 0u) gbits2 |= 1u << q;
            const TileSet t = tile_set(g, ix, iy, iz);
            if (!(t == last)) {
                mark_tiles(tg, tiles, t);
                last = t;
            }
        }
        pre += packed(cc[q]);
        if (++ix == nex) {
            ix = 0;
            if (++iy == ney) {
                iy = 0;
                iz++;
            }
        }
    }
    uint4 *s4 = (uint4 *)(cstart + c0);
    uint4 *e4 = (uint4 *)(cend + c0);
#pragma unroll
    for (int r = 0; r < kItems / 4; r++) {
        const uint4 so = make_uint4(st[4 * r], st[4 * r + 1], st[4 * r + 2], st[4 * r + 3]);
        s4[r] = so;
        e4[r] = so;
        if (bits && !keep) c4[r] = make_uint4(0u, 0u, 0u, 0u);
    }
    // occupancy and group-start bits: two threads per 32-cell word
    static_assert(kItems == 16, "two threads per occupancy word");
    uint32_t w = bits << (16 * (threadIdx.x & 1));
    uint32_t gw = gbits << (16 * (threadIdx.x & 1));
    uint32_t gw2 = gbits2 << (16 * (threadIdx.x & 1));
    w |= (uint32_t)__shfl_xor((int)w, 1, 64);
    gw |= (uint32_t)__shfl_xor((int)gw, 1, 64);
    gw2 |= (uint32_t)__shfl_xor((int)gw2, 1, 64);
    if ((threadIdx.x & 1) == 0) {
        occ[c0 >> 5] = w;
        gs[c0 >> 5] = gw;
        if (gsize2) gs2[c0 >> 5] = gw2;
    }
}

// Each wave places 128 consecutive slots as two halves of 64 whose loads and
// atomics are in flight together; a run of equal keys that crosses from the
// first half into the second takes one atomic for both parts.
__global__ void __launch_bounds__(256) k_bin_scatter(const uint32_t *__restrict__ keys, int n, int nelem,
                                                      const uint32_t *__restrict__ cstart, uint32_t *__restrict__ cend,
                                                      uint32_t *__restrict__ perm, uint32_t *__restrict__ stamp_arr,
                                                      const BinCounters *ctr) {
    const int lane = threadIdx.x & 63;
    const uint32_t stamp = ctr->stamp;
    if (ctr->bad) return;
    const uint32_t un = (uint32_t)nelem;
    for (int wbase = (blockIdx.x * (int)blockDim.x + (int)threadIdx.x - lane) * 2; wbase < n;
         wbase += gridDim.x * blockDim.x * 2) {
        const int p0 = wbase + lane, p1 = wbase + 64 + lane;
        const uint32_t k0 = (p0 < n) ? keys[p0] : 0xFFFFFFFFu;
        const uint32_t k1 = (p1 < n) ? keys[p1] : 0xFFFFFFFFu;
        int len0, head0, len1, head1;
        const bool h0 = wave_runs(k0, lane, &len0, &head0);
        const bool h1 = wave_runs(k1, lane, &len1, &head1);
        const bool cross = (uint32_t)__shfl((int)k0, 63, 64) == (uint32_t)__shfl((int)k1, 0, 64);
        const int head_last0 = __shfl(head0, 63, 64);  // head of the run that reaches the end of half 0
        const int len_first1 = __shfl(len1, 0, 64);
        // both halves' atomics (and the cell starts they are compared with)
        // are issued before either result is used
        uint32_t b0 = 0, b1 = 0, s0 = 0, s1 = 0;
        const bool a0 = h0 && k0 < un;
        const bool a1 = h1 && k1 < un && !(cross && lane == 0);
        if (a0) {
            const int add = len0 + ((cross && lane == head_last0) ? len_first1 : 0);
            b0 = atomicAdd(&cend[k0], (uint32_t)add);
            s0 = cstart[k0];
        }
        if (a1) {
            b1 = atomicAdd(&cend[k1], (uint32_t)len1);
            s1 = cstart[k1];
        }
        // (an empty asm holds the compares until both atomics are issued: the
        // compiler would otherwise move the first compare, and its wait,
        // into the first branch)
        asm volatile("" : "+v"(b0), "+v"(s0), "+v"(b1), "+v"(s1));
        if (a0 && b0 != s0) stamp_arr[k0] = stamp;  // several runs: the fix-up sorts this cell
        if (a1 && b1 != s1) stamp_arr[k1] = stamp;
        const uint32_t bb0 = (uint32_t)__shfl((int)b0, head0, 64);
        // the crossing run continues after its 64 - head_last0 slots in half 0
        const uint32_t cont = (uint32_t)__shfl((int)b0, head_last0, 64) + (uint32_t)(64 - head_last0);
        const uint32_t bb1 = (cross && head1 == 0) ? cont : (uint32_t)__shfl((int)b1, head1, 64);
        if (k0 < un) perm[bb0 + (uint32_t)(lane - head0)] = (uint32_t)p0;
        if (k1 < un) perm[bb1 + (uint32_t)(lane - head1)] = (uint32_t)p1;
    }
}

// Sorts the positions of each cell marked by this binning by slot index.
__global__ void __launch_bounds__(256) k_bin_fixup(const uint32_t *__restrict__ active,
                                                    const uint32_t *__restrict__ stamp_arr, const BinCounters *ctr,
                                                    const uint32_t *__restrict__ cstart,
                                                    const uint32_t *__restrict__ cend, uint32_t *__restrict__ perm,
                                                    uint32_t *__restrict__ err) {
    if (ctr->bad) return;
    const uint32_t na = ctr->n_active, stamp = ctr->stamp;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < na; i += gridDim.x * blockDim.x) {
        const uint32_t c = active[i];
        if (stamp_arr[c] != stamp) continue;
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

void launch_bin_reduce(hipStream_t s, const GridGeom &g, BinBuffers &b) {
    hipLaunchKernelGGL(k_bin_reduce, dim3(cdiv(g.nelem, kTile)), dim3(kBlock), 0, s, (const uint32_t *)b.cnt, b.bsum);
}

void launch_bin_count(hipStream_t s, const GridGeom &g, const TileGeom &tg, const uint32_t *keys, int n,
                      BinBuffers &b) {
    int pblocks = cdiv(n > 0 ? n : 1, 256);
    if (pblocks > 8192) pblocks = 8192;
    hipLaunchKernelGGL(k_bin_count, dim3(pblocks), dim3(256), 0, s, keys, n, g.nelem, b.cnt, b.tiles, tg.ntiles, b.ctr);
    launch_bin_reduce(s, g, b);
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
    if (!counted) hipLaunchKernelGGL(k_bin_reduce, dim3(nb), dim3(kBlock), 0, s, (const uint32_t *)b.cnt, b.bsum);
    hipLaunchKernelGGL(k_bin_finalize, dim3(nb), dim3(kBlock), 0, s, g, b.cnt, (const uint64_t *)b.bsum, nb, b.tprev,
                       b.ctr, n, err, b.cstart, b.cend, b.active, b.astart, b.tiles, tg, b.occ, b.gs, b.gsize > 0 ? b.gsize : 1u,
                       b.gs2, b.gsize2, keep ? 1 : 0);
    if (n > 0) {
        int sblocks = cdiv(n, 512);
        if (sblocks > 8192) sblocks = 8192;
        hipLaunchKernelGGL(k_bin_scatter, dim3(sblocks), dim3(256), 0, s, keys, n, nelem, (const uint32_t *)b.cstart,
                           b.cend, b.perm, b.stamp, (const BinCounters *)b.ctr);
        int fblocks = cdiv(nelem, 256);
        if (fblocks > 4096) fblocks = 4096;
        hipLaunchKernelGGL(k_bin_fixup, dim3(fblocks), dim3(256), 0, s, (const uint32_t *)b.active,
                           (const uint32_t *)b.stamp, (const BinCounters *)b.ctr, (const uint32_t *)b.cstart,
                           (const uint32_t *)b.cend, b.perm, err);
    }
    return (int)hipGetLastError();
}

}  // namespace mpm

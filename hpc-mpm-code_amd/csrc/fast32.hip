// FP32 explicit-MPM step kernels for gfx950 (wave64, 160 KB LDS per CU).
//
// Data layout (HBM, structure of arrays, one slot per particle): position as
// the cell key plus an FP32 offset inside the cell in cell units
// (x = xmin + dx (i + r), r in [0, 1), resolution 2^-24 of a cell),
// velocity, mass, reference volume, the displacement gradient H = F - I (so
// small strains do not cancel against the identity), Cauchy stress, and FP64
// det F and volume. Nodes are FP32 [u|v|w] blocks like the reference's node
// vectors.
//
// P2G is deterministic without float atomics: a 64-thread workgroup owns a
// row segment of 64 consecutive cells along x. Their particles are one
// contiguous range of the cell-sorted order, which is staged through LDS in
// chunks with coalesced loads; each thread then folds the particles of its
// own cell, in sorted order, into the 8 corner partials of that cell
// (registers). A node pass adds the partials of the <= 8 cells around each
// node in ascending cell order with Neumaier-compensated FP32 sums and
// applies the grid update (velocity, acceleration, BCs) in the same pass.
// The MUSL remap uses the same structure, with the G2P velocity of every
// particle computed from the 8 node values of its cell held in registers.
#include "mpm_common.h"
#include "mpm_kernels.h"

#include <algorithm>

namespace mpm {

namespace {

constexpr int kSeg = 64;      // cells per row segment = threads per workgroup
constexpr int kChunk = 256;   // particles staged per LDS pass

// q = a / n, rem = a % n for a < 2^24 via the float reciprocal, corrected to exact
__device__ __forceinline__ int fdivmod(int a, int n, float inv_n, int &rem) {
    int q = (int)((float)a * inv_n);
    int r = a - q * n;
    if (r < 0) {
        q--;
        r += n;
    } else if (r >= n) {
        q++;
        r -= n;
    }
    rem = r;
    return q;
}

// Neumaier (improved Kahan) compensated sum.
__device__ __forceinline__ void neu_add(float &s, float &c, float x) {
    float t = s + x;
    if (fabsf(s) >= fabsf(x))
        c += (s - t) + x;
    else
        c += (x - t) + s;
    s = t;
}

// A group is 64 consecutive entries of the active-cell list (one per thread).
// Because the list is in ascending cell order and skips only empty cells, the
// particles of a group are one contiguous range [S, E) of the sorted order.
struct CellInfo {
    int ix, iy, iz, c;
    uint32_t cs, ce;
};

// A group is G = 64 / L consecutive entries of the active-cell list; L lanes
// work on each cell (lane = cell-in-group * L + sub). Because the list is in
// ascending cell order and skips only empty cells, the particles of a group
// are one contiguous range [S, E) of the sorted order.
template <int L>
__device__ __forceinline__ void group_setup(const GridGeom &g, const CellLists &L_, uint32_t n_active, int grp,
                                            CellInfo &ci, uint32_t &S, uint32_t &E) {
    constexpr int G = kSeg / L;
    const int t = threadIdx.x;
    const uint32_t a0 = (uint32_t)grp * G;
    const uint32_t a = a0 + t / L;
    const uint32_t a_end = min(a0 + (uint32_t)G, n_active);
    // uniform: scalar loads, which do not hold up the vector-memory counter
    S = ((const __attribute__((address_space(4))) uint32_t *)L_.astart)[a0];
    E = ((const __attribute__((address_space(4))) uint32_t *)L_.astart)[a_end];
    ci.c = -1;
    ci.cs = ci.ce = 0u;
    ci.ix = ci.iy = ci.iz = 0;
    if (a < n_active) {
        ci.c = (int)L_.active[a];
        ci.cs = L_.astart[a];
        ci.ce = L_.astart[a + 1];
        const int nexy = g.nex[0] * g.nex[1];
        ci.iz = ci.c / nexy;
        const int r = ci.c - ci.iz * nexy;
        ci.iy = r / g.nex[0];
        ci.ix = r - ci.iy * g.nex[0];
    }
}

// first position >= lo of the arithmetic sequence cs + sub + k L
template <int L>
__device__ __forceinline__ uint32_t first_pos(uint32_t cs, uint32_t lo, int sub) {
    uint32_t f = cs + (uint32_t)sub;
    if (f < lo) f += ((lo - f + L - 1) / L) * L;
    return f;
}

// Sums NV values over the L lanes of a cell and leaves lane sub with the
// totals of entries k L + sub at v[k], k < NV / L. Each level exchanges half
// of the remaining entries with the lane O apart; a total is
// ((a0 + a1) + (a2 + a3)) ... in lane order, as a xor butterfly would give.
template <int O, int L, int N>
struct ReduceScatter {
    static __device__ __forceinline__ void run(float *v, int sub) {
        const bool hi = (sub & O) != 0;
#pragma unroll
        for (int k = 0; k < N / 2; k++) {
            const float keep = hi ? v[2 * k + 1] : v[2 * k];
            const float send = hi ? v[2 * k] : v[2 * k + 1];
            v[k] = keep + __shfl_xor(send, O, 64);
        }
        ReduceScatter<O * 2, L, N / 2>::run(v, sub);
    }
};
template <int L, int N>
struct ReduceScatter<L, L, N> {
    static __device__ __forceinline__ void run(float *, int) {}
};

// ------------------------------------------------------------------ P2G
// Particle fields are gathered (through the sorted order) straight from HBM
// into LDS with gfx950's global_load_lds_dword: lane l of a wave writes the
// dword for chunk particle 64*r + l at stage block r. Nothing is held in
// VGPRs while the loads are in flight, so a wave keeps all of a chunk's loads
// outstanding. Each field's 64-particle blocks are S = 64 + L dwords apart
// for L lanes per cell: lane (cell c, sub s) of a 16-point-per-cell group
// reads particle 16 c + s + L k, which lands in block c / (64 / 16) at
// bank (L (c / 4) + 16 (c % 4) + s + L k) mod 64 -- a different bank for
// every lane of the wave (likewise for 8 points per cell).
constexpr int kBlk = kChunk / 64;
template <int S>
__device__ __forceinline__ int st_idx(int f, int i) { return (f * kBlk + (i >> 6)) * S + (i & 63); }

// cache policy of the LDS DMA: 0, or kNT for a stream read once per step
// (non-temporal, the gfx9 SLC / gfx940+ NT bit), which keeps L2 for the data
// the next kernels re-read
constexpr int kNT = 2;
template <int S, int CPOL = 0>
__device__ __forceinline__ void lds_gather(const void *field_base, uint32_t slot, float *stage, int f, int r) {
    __builtin_amdgcn_global_load_lds((const uint32_t *)field_base + slot, stage + st_idx<S>(f, 64 * r), 4, 0, CPOL);
}

// wait for every outstanding vector-memory load (including LDS DMA) of this wave
__device__ __forceinline__ void wait_vm_all() { __builtin_amdgcn_s_waitcnt(0); }

// wait until at most the kBlk most recent vector-memory loads of this wave
// are in flight (vmcnt(4); loads complete in issue order), leaving the slot
// prefetch of the next chunk outstanding
static_assert(kBlk == 4, "the wait below encodes vmcnt(4)");
__device__ __forceinline__ void wait_vm_keep_slots() { __builtin_amdgcn_s_waitcnt(0x0F74); }

// A cells-kernel block is one wave: LDS written by its own DMA is visible
// after the vmcnt wait, and its own LDS accesses execute in order, so a
// compiler barrier (no hardware barrier, whose fence would wait for every
// outstanding load) orders the stage buffer's reuse.
static_assert(kSeg == 64, "cells kernels assume one wave per block");
__device__ __forceinline__ void wave_fence() { asm volatile("" ::: "memory"); }

// slot indices of the chunk starting at sorted position base (nch particles)
__device__ __forceinline__ void load_slots(const CellLists &L, uint32_t base, int nch, uint32_t slot[kBlk]) {
#pragma unroll
    for (int r = 0; r < kBlk; r++) slot[r] = (64 * r + (int)threadIdx.x < nch) ? L.perm[base + 64 * r + threadIdx.x] : 0u;
}

// the same for every lane, positions clamped into [base, end): exactly kBlk
// load instructions, whatever the chunk length
__device__ __forceinline__ void load_slots_all(const CellLists &L, uint32_t base, uint32_t end, uint32_t slot[kBlk]) {
#pragma unroll
    for (int r = 0; r < kBlk; r++) slot[r] = L.perm[min(base + 64u * r + threadIdx.x, end - 1u)];
}

// PF: software prefetch. The next group's cell ranges are requested at the
// start of a group and the next chunk's slot indices while the current
// chunk's fields are in flight, so that a chunk waits for one memory round
// trip instead of three (setup -> slots -> fields).
template <int LPC, bool DIAG, bool TRAC, bool PF>
__global__ void __launch_bounds__(kSeg) k32_p2g_cells(GridGeom g, PartF32 P, CellLists L, NodeTables T,
                                                       float *__restrict__ part) {
    constexpr int G = kSeg / LPC;
    // fields: 0-2 r, 3-5 v, 6 m, 7 V, 8-13 sigma, [14-16 traction]
    constexpr int NF = 14 + (TRAC ? 3 : 0);
    constexpr int SS = 64 + LPC;
    __shared__ float stage[NF * kBlk * SS];
    const uint32_t n_active = L.ctr->n_active;
    const int ngroups = (int)((n_active + G - 1) / G);
    const int sub = threadIdx.x % LPC;
    const int t = threadIdx.x;
    const size_t ld = P.ld;
    const float inv0 = T.inv_dx[0], inv1 = T.inv_dx[1], inv2 = T.inv_dx[2];
    int grp = blockIdx.x;
    if (grp >= ngroups) return;
    CellInfo si;
    uint32_t S, E;
    group_setup<LPC>(g, L, n_active, grp, si, S, E);
    uint32_t slot[kBlk];
    if (PF) load_slots(L, S, (int)min((uint32_t)kChunk, E - S), slot);
    for (; grp < ngroups;) {
        const int grp_n = grp + (int)gridDim.x;
        CellInfo si_n;
        uint32_t S_n = 0, E_n = 0;
        if (PF && grp_n < ngroups) group_setup<LPC>(g, L, n_active, grp_n, si_n, S_n, E_n);
        // mass and momentum per corner; internal force as moments of -V sigma
        // against the weight products (FX[a][j + 2k] = sum S_ax wy wz, FY over
        // wx wz, FZ over wx wy) so that one multiply-add per particle covers
        // all corners; corners take them with the sign of dN/dx at the end
        float acc[8][4];
        float FX[3][4], FY[3][4], FZ[3][4];
        float tr[8][3];
        float vof[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
#pragma unroll
            for (int f = 0; f < 4; f++) acc[q][f] = 0.f;
#pragma unroll
            for (int a = 0; a < 3; a++) tr[q][a] = 0.f;
            vof[q] = 0.f;
        }
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int u = 0; u < 4; u++) FX[a][u] = FY[a][u] = FZ[a][u] = 0.f;

        for (uint32_t base = S; base < E; base += kChunk) {
            const int nch = (int)min((uint32_t)kChunk, E - base);
            if (!PF) load_slots(L, base, nch, slot);
#pragma unroll
            for (int r = 0; r < kBlk; r++) {
                if (64 * r + t < nch) {
                    const uint32_t sl = slot[r];
                    // P2G reads these fields once per step: non-temporal (the
                    // update, two kernels on, measured 4 us faster)
#pragma unroll
                    for (int a = 0; a < 3; a++) {
                        lds_gather<SS, kNT>(P.r + a * ld, sl, stage, a, r);
                        lds_gather<SS, kNT>(P.v + a * ld, sl, stage, 3 + a, r);
                    }
                    lds_gather<SS, kNT>(P.m, sl, stage, 6, r);
                    lds_gather<SS, kNT>(P.V32, sl, stage, 7, r);
#pragma unroll
                    for (int q = 0; q < 6; q++) lds_gather<SS, kNT>(P.sig + q * ld, sl, stage, 8 + q, r);
                    if (TRAC) {
#pragma unroll
                        for (int a = 0; a < 3; a++) lds_gather<SS, kNT>(P.trac + a * ld, sl, stage, 14 + a, r);
                    }
                }
            }
            // slots of the next chunk (later in this group or the first of the
            // next group) load while this chunk's fields are in flight
            uint32_t slot_n[kBlk];
            bool pf_out = false;
            if (PF) {
                const uint32_t nb = base + kChunk;
                if (nb < E) {
                    load_slots_all(L, nb, E, slot_n);
                    pf_out = true;
                } else if (grp_n < ngroups) {
                    load_slots_all(L, S_n, E_n, slot_n);
                    pf_out = true;
                }
            }
            if (pf_out)
                wait_vm_keep_slots();
            else
                wait_vm_all();
            wave_fence();
            const uint32_t lo = max(si.cs, base), hi = min(si.ce, base + (uint32_t)nch);
            for (uint32_t pos = first_pos<LPC>(si.cs, lo, sub); pos < hi; pos += LPC) {
                const int i = (int)(pos - base);
                const float fx = stage[st_idx<SS>(0, i)];
                const float fy = stage[st_idx<SS>(1, i)];
                const float fz = stage[st_idx<SS>(2, i)];
                const float m = stage[st_idx<SS>(6, i)];
                const float px = m * stage[st_idx<SS>(3, i)], py = m * stage[st_idx<SS>(4, i)], pz = m * stage[st_idx<SS>(5, i)];
                const float Vp = stage[st_idx<SS>(7, i)];
                const float nV = -Vp;
                // -V sigma (Voigt xx,yy,zz,yz,xz,xy) scaled by 1/dx of the derivative direction
                const float s0 = nV * stage[st_idx<SS>(8, i)], s1 = nV * stage[st_idx<SS>(9, i)];
                const float s2 = nV * stage[st_idx<SS>(10, i)], s3 = nV * stage[st_idx<SS>(11, i)];
                const float s4 = nV * stage[st_idx<SS>(12, i)], s5 = nV * stage[st_idx<SS>(13, i)];
                // force on a corner: f_a = sum_b (-V sigma_ab) dN/dx_b, dN/dx = (sx/dx wy wz, sy/dy wx wz, sz/dz wx wy)
                const float X0 = s0 * inv0, X1 = s5 * inv0, X2 = s4 * inv0;  // b = x: sigma_xx, sigma_yx, sigma_zx
                const float Y0 = s5 * inv1, Y1 = s1 * inv1, Y2 = s3 * inv1;  // b = y
                const float Z0 = s4 * inv2, Z1 = s3 * inv2, Z2 = s2 * inv2;  // b = z
                float tx = 0.f, ty = 0.f, tz = 0.f;
                if (TRAC) {
                    tx = stage[st_idx<SS>(14, i)];
                    ty = stage[st_idx<SS>(15, i)];
                    tz = stage[st_idx<SS>(16, i)];
                }
                const float wx[2] = {1.f - fx, fx}, wy[2] = {1.f - fy, fy}, wz[2] = {1.f - fz, fz};
                float wyz[2][2], wxz[2][2], wxy[2][2];
#pragma unroll
                for (int u = 0; u < 2; u++)
#pragma unroll
                    for (int w = 0; w < 2; w++) {
                        wyz[u][w] = wy[u] * wz[w];
                        wxz[u][w] = wx[u] * wz[w];
                        wxy[u][w] = wx[u] * wy[w];
                    }
                const float Xa[3] = {X0, X1, X2}, Ya[3] = {Y0, Y1, Y2}, Za[3] = {Z0, Z1, Z2};
#pragma unroll
                for (int a = 0; a < 3; a++)
#pragma unroll
                    for (int u = 0; u < 2; u++)
#pragma unroll
                        for (int w = 0; w < 2; w++) {
                            FX[a][u + 2 * w] += Xa[a] * wyz[u][w];
                            FY[a][u + 2 * w] += Ya[a] * wxz[u][w];
                            FZ[a][u + 2 * w] += Za[a] * wxy[u][w];
                        }
#pragma unroll
                for (int k = 0; k < 2; k++)
#pragma unroll
                    for (int j = 0; j < 2; j++)
#pragma unroll
                        for (int ii = 0; ii < 2; ii++) {
                            const int q = ii + 2 * j + 4 * k;
                            const float N = wx[ii] * wyz[j][k];
                            acc[q][0] += N * m;
                            acc[q][1] += N * px;
                            acc[q][2] += N * py;
                            acc[q][3] += N * pz;
                            if (TRAC) {
                                tr[q][0] += N * tx;
                                tr[q][1] += N * ty;
                                tr[q][2] += N * tz;
                            }
                            if (DIAG) vof[q] += N * Vp;
                        }
            }
            wave_fence();
            if (pf_out) {
#pragma unroll
                for (int r = 0; r < kBlk; r++) slot[r] = slot_n[r];
            }
        }
        // partial of entry q * 7 + f (f: m, p, force) and 56 + q (volume)
        constexpr int NV = DIAG ? 64 : 56;
        float v[NV];
#pragma unroll
        for (int k = 0; k < 2; k++)
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int ii = 0; ii < 2; ii++) {
                    const int q = ii + 2 * j + 4 * k;
#pragma unroll
                    for (int f = 0; f < 4; f++) v[q * 7 + f] = acc[q][f];
#pragma unroll
                    for (int a = 0; a < 3; a++) {
                        const float fx_ = ii ? FX[a][j + 2 * k] : -FX[a][j + 2 * k];
                        const float fy_ = j ? FY[a][ii + 2 * k] : -FY[a][ii + 2 * k];
                        const float fz_ = k ? FZ[a][ii + 2 * j] : -FZ[a][ii + 2 * j];
                        v[q * 7 + 4 + a] = (fx_ + fy_) + fz_ + (TRAC ? tr[q][a] : 0.f);
                    }
                    if (DIAG) v[56 + q] = vof[q];
                }
        ReduceScatter<1, LPC, NV>::run(v, sub);
        if (si.ce > si.cs) {
            const size_t ne = (size_t)g.nelem;
#pragma unroll
            for (int k = 0; k < NV / LPC; k++) part[(size_t)(k * LPC + sub) * ne + si.c] = v[k];
        }
        grp = grp_n;
        if (grp < ngroups) {
            if (PF) {
                si = si_n;
                S = S_n;
                E = E_n;
            } else {
                group_setup<LPC>(g, L, n_active, grp, si, S, E);
            }
        }
    }  // groups
}

__device__ __forceinline__ void node_ijk(const GridGeom &g, int n, int &i, int &j, int &k) {
    i = n % g.nxc[0];
    j = (n / g.nxc[0]) % g.nxc[1];
    k = n / (g.nxc[0] * g.nxc[1]);
}

__device__ __forceinline__ float safe_div(float p, float m, float m_small) {
    if (m >= m_small) return p / m;
    return (float)((double)p / (double)m);
}

// Node of this thread in the node tile of this block; false when the tile has
// no non-empty cell next to it (all threads of the block agree) or the node
// lies outside the grid.
__device__ __forceinline__ bool tile_node(const GridGeom &g, const CellLists &L, const TileGeom &tg, bool all_tiles,
                                          int &n, int &i, int &j, int &k) {
    const int tile = blockIdx.x;
    if (!all_tiles && L.tiles[tile] == 0u) return false;
    const int tx = tile % tg.ntx;
    const int ty = (tile / tg.ntx) % tg.nty;
    const int tz = tile / (tg.ntx * tg.nty);
    const int l = threadIdx.x;
    i = tx * kNodeTileX + (l % kNodeTileX);
    j = ty * kNodeTileY + (l / kNodeTileX) % kNodeTileY;
    k = tz * kNodeTileZ + l / (kNodeTileX * kNodeTileY);
    if (i >= g.nxc[0] || j >= g.nxc[1] || k >= g.nxc[2]) return false;
    n = i + g.nxc[0] * (j + g.nxc[1] * k);
    return true;
}

__device__ __forceinline__ bool cell_nonempty(const CellLists &L, int c) { return (L.occ[c >> 5] >> (c & 31)) & 1u; }

// The <= 8 cells around node (i, j, k) in ascending cell order: cell index
// (clamped into the grid when outside) and whether it holds particles. All
// of a node's partial loads can then be issued at once instead of one cell's
// loads per dependent round trip.
struct NodeCells {
    int c[8];
    int corner[8];
    bool on[8];
};
__device__ __forceinline__ bool node_cells8(const GridGeom &g, const CellLists &L, int i, int j, int k, NodeCells &nc) {
    bool any = false;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int di = q & 1, dj = (q >> 1) & 1, dk = q >> 2;  // q ascending = cell index ascending
        const int ci = i - 1 + di, cj = j - 1 + dj, ck = k - 1 + dk;
        const bool in = ci >= 0 && ci < g.nex[0] && cj >= 0 && cj < g.nex[1] && ck >= 0 && ck < g.nex[2];
        const int c = in ? ci + cj * g.nex[0] + ck * g.nex[0] * g.nex[1] : 0;
        nc.c[q] = c;
        nc.corner[q] = (1 - di) + 2 * (1 - dj) + 4 * (1 - dk);
        nc.on[q] = in && cell_nonempty(L, c);
        any = any || nc.on[q];
    }
    return any;
}

template <bool DIAG>
__global__ void __launch_bounds__(256) k32_p2g_nodes(GridGeom g, CellLists L, TileGeom tg,
                                                      const float *__restrict__ part, NodesF32 N, BcDense bc,
                                                      StepConsts sc) {
    int n, i, j, k;
    if (!tile_node(g, L, tg, DIAG, n, i, j, k)) return;
    const size_t ne = (size_t)g.nelem;
    constexpr int NQ = DIAG ? 8 : 7;
    float s[NQ], cp[NQ];
#pragma unroll
    for (int f = 0; f < NQ; f++) {
        s[f] = 0.f;
        cp[f] = 0.f;
    }
    NodeCells nc;
    const bool any = node_cells8(g, L, i, j, k, nc);
    if (any) {
        // loads only from cells that hold particles: a tile's nodes next to
        // empty cells fetch nothing for them
        float val[8][NQ];
#pragma unroll
        for (int q = 0; q < 8; q++) {
#pragma unroll
            for (int f = 0; f < NQ; f++) val[q][f] = 0.f;
            if (nc.on[q]) {
#pragma unroll
                for (int f = 0; f < 7; f++) val[q][f] = part[(size_t)(nc.corner[q] * 7 + f) * ne + nc.c[q]];
                if (DIAG) val[q][7] = part[(size_t)(56 + nc.corner[q]) * ne + nc.c[q]];
            }
        }
#pragma unroll
        for (int q = 0; q < 8; q++)
            if (nc.on[q]) {
#pragma unroll
                for (int f = 0; f < NQ; f++) neu_add(s[f], cp[f], val[q][f]);
            }
    }
    const int nodec = g.nodec;
    if (!any) {
        if (DIAG) {
            N.nm[n] = 0.f;
            N.nvof[n] = 0.f;
            for (int c = 0; c < 3; c++) {
                N.nv[c * nodec + n] = 0.f;
                N.na[c * nodec + n] = 0.f;
                N.np[c * nodec + n] = 0.f;
                N.nf[c * nodec + n] = 0.f;
                N.nvm[c * nodec + n] = 0.f;
            }
        }
        return;
    }
    float m = s[0] + cp[0];
    float p[3], f[3];
#pragma unroll
    for (int c = 0; c < 3; c++) p[c] = s[1 + c] + cp[1 + c];
    f[0] = (s[4] + cp[4]) + m * sc.bx;
    f[1] = (s[5] + cp[5]) + m * sc.by;
    f[2] = (s[6] + cp[6]) + m * sc.bz;
    float v[3] = {0.f, 0.f, 0.f}, a[3] = {0.f, 0.f, 0.f};
    if ((double)m > sc.mtol) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            v[c] = safe_div(p[c], m, sc.m_small);
            a[c] = safe_div(f[c], m, sc.m_small);
        }
    }
    uint8_t mk = bc.mask ? bc.mask[n] : 0;
    if (mk) {
#pragma unroll
        for (int c = 0; c < 3; c++)
            if (mk & (1 << c)) {
                v[c] = (float)(bc.val[c * nodec + n] * sc.facl);
                a[c] = 0.f;
            }
    }
    N.nm[n] = m;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        N.nv[c * nodec + n] = v[c];
        N.na[c * nodec + n] = a[c];
        // formed once here rather than from two loads per corner in each MUSL cell
        N.nvp[c * nodec + n] = v[c] + sc.dt * a[c];
    }
    if (DIAG) {
        float vof = s[7] + cp[7];
        if ((double)vof < sc.mtol) vof = 0.f;
        N.nvof[n] = vof;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            N.np[c * nodec + n] = p[c];
            N.nf[c * nodec + n] = f[c];
        }
    }
}

// ------------------------------------------------------ G2P velocity + MUSL
template <int LPC, bool FLIP>
__global__ void __launch_bounds__(kSeg) k32_musl_cells(GridGeom g, PartF32 P, CellLists L, NodeTables T, NodesF32 N,
                                                        StepConsts sc, float *__restrict__ part) {
    constexpr int G = kSeg / LPC;
    // staged fields: 0-2 r, 3 m, [4-6 old v for FLIP]; results: G2P velocity per particle
    constexpr int NF = 4 + (FLIP ? 3 : 0);
    constexpr int SS = 64 + LPC;
    __shared__ float stage[NF * kBlk * SS];
    __shared__ float vout[3 * kBlk * SS];
    const uint32_t n_active = L.ctr->n_active;
    const int ngroups = (int)((n_active + G - 1) / G);
    const int sub = threadIdx.x % LPC;
    const int t = threadIdx.x;
    const size_t ld = P.ld;
    const int nodec = g.nodec;
    for (int grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
        CellInfo si;
        uint32_t S, E;
        __syncthreads();
        group_setup<LPC>(g, L, n_active, grp, si, S, E);
        __syncthreads();
        const bool ne = si.ce > si.cs;

        // node values of this cell: PIC uses v + dt a (var_trans_solid_explicit.cpp:194-200),
        // FLIP uses a
        float nvl[8][3];
        if (ne) {
            const int n0 = si.ix + g.nxc[0] * si.iy + g.nxc[0] * g.nxc[1] * si.iz;
#pragma unroll
            for (int k = 0; k < 2; k++)
#pragma unroll
                for (int j = 0; j < 2; j++)
#pragma unroll
                    for (int i = 0; i < 2; i++) {
                        const int q = i + 2 * j + 4 * k;
                        const int nid = n0 + i + j * g.nxc[0] + k * g.nxc[0] * g.nxc[1];
#pragma unroll
                        for (int c = 0; c < 3; c++) {
                            nvl[q][c] = FLIP ? N.na[c * nodec + nid] : N.nvp[c * nodec + nid];
                        }
                    }
        } else {
#pragma unroll
            for (int q = 0; q < 8; q++)
#pragma unroll
                for (int c = 0; c < 3; c++) nvl[q][c] = 0.f;
        }

        float acc[8][3];
#pragma unroll
        for (int q = 0; q < 8; q++)
#pragma unroll
            for (int c = 0; c < 3; c++) acc[q][c] = 0.f;

        for (uint32_t base = S; base < E; base += kChunk) {
            const int nch = (int)min((uint32_t)kChunk, E - base);
            uint32_t slot[kBlk];
#pragma unroll
            for (int r = 0; r < kBlk; r++) slot[r] = (64 * r + t < nch) ? L.perm[base + 64 * r + t] : 0u;
#pragma unroll
            for (int r = 0; r < kBlk; r++) {
                if (64 * r + t < nch) {
                    const uint32_t sl = slot[r];
#pragma unroll
                    for (int a = 0; a < 3; a++) {
                        lds_gather<SS>(P.r + a * ld, sl, stage, a, r);
                        if (FLIP) lds_gather<SS>(P.v + a * ld, sl, stage, 4 + a, r);
                    }
                    lds_gather<SS>(P.m, sl, stage, 3, r);
                }
            }
            wait_vm_all();
            __syncthreads();
            const uint32_t lo = max(si.cs, base), hi = min(si.ce, base + (uint32_t)nch);
            for (uint32_t pos = first_pos<LPC>(si.cs, lo, sub); pos < hi; pos += LPC) {
                const int i = (int)(pos - base);
                const float fx = stage[st_idx<SS>(0, i)];
                const float fy = stage[st_idx<SS>(1, i)];
                const float fz = stage[st_idx<SS>(2, i)];
                const float m = stage[st_idx<SS>(3, i)];
                const float wx[2] = {1.f - fx, fx}, wy[2] = {1.f - fy, fy}, wz[2] = {1.f - fz, fz};
                float Nw[8];
                float vp[3] = {0.f, 0.f, 0.f};
#pragma unroll
                for (int k = 0; k < 2; k++)
#pragma unroll
                    for (int j = 0; j < 2; j++)
#pragma unroll
                        for (int ii = 0; ii < 2; ii++) {
                            const int q = ii + 2 * j + 4 * k;
                            Nw[q] = wx[ii] * (wy[j] * wz[k]);
#pragma unroll
                            for (int c = 0; c < 3; c++) vp[c] += Nw[q] * nvl[q][c];
                        }
                if (FLIP) {
                    vp[0] = stage[st_idx<SS>(4, i)] + sc.dt * vp[0];
                    vp[1] = stage[st_idx<SS>(5, i)] + sc.dt * vp[1];
                    vp[2] = stage[st_idx<SS>(6, i)] + sc.dt * vp[2];
                }
                vout[st_idx<SS>(0, i)] = vp[0];
                vout[st_idx<SS>(1, i)] = vp[1];
                vout[st_idx<SS>(2, i)] = vp[2];
                const float mv[3] = {m * vp[0], m * vp[1], m * vp[2]};
#pragma unroll
                for (int q = 0; q < 8; q++)
#pragma unroll
                    for (int c = 0; c < 3; c++) acc[q][c] += Nw[q] * mv[c];
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < kBlk; r++) {
                const int i = 64 * r + t;
                if (i < nch) {
                    const uint32_t sl = slot[r];
                    // next read a step later (non-temporal, as the update's stores)
                    __builtin_nontemporal_store(vout[st_idx<SS>(0, i)], &P.v[sl]);
                    __builtin_nontemporal_store(vout[st_idx<SS>(1, i)], &P.v[ld + sl]);
                    __builtin_nontemporal_store(vout[st_idx<SS>(2, i)], &P.v[2 * ld + sl]);
                }
            }
            __syncthreads();
        }
        float v[24];
#pragma unroll
        for (int q = 0; q < 8; q++)
#pragma unroll
            for (int c = 0; c < 3; c++) v[q * 3 + c] = acc[q][c];
        ReduceScatter<1, LPC, 24>::run(v, sub);
        if (ne) {
            const size_t nel = (size_t)g.nelem;
#pragma unroll
            for (int k = 0; k < 24 / LPC; k++) part[(size_t)(k * LPC + sub) * nel + si.c] = v[k];
        }
    }  // groups
}

template <bool DIAG>
__global__ void __launch_bounds__(256) k32_musl_nodes(GridGeom g, CellLists L, TileGeom tg,
                                                       const float *__restrict__ part, NodesF32 N, BcDense bc,
                                                       StepConsts sc) {
    int n, i, j, k;
    if (!tile_node(g, L, tg, DIAG, n, i, j, k)) return;
    const size_t ne = (size_t)g.nelem;
    float s[3] = {0.f, 0.f, 0.f}, cp[3] = {0.f, 0.f, 0.f};
    NodeCells nc;
    const bool any = node_cells8(g, L, i, j, k, nc);
    if (any) {
        float val[8][3];
#pragma unroll
        for (int q = 0; q < 8; q++) {
#pragma unroll
            for (int f = 0; f < 3; f++) val[q][f] = 0.f;
            if (nc.on[q]) {
#pragma unroll
                for (int f = 0; f < 3; f++) val[q][f] = part[(size_t)(nc.corner[q] * 3 + f) * ne + nc.c[q]];
            }
        }
#pragma unroll
        for (int q = 0; q < 8; q++)
            if (nc.on[q]) {
#pragma unroll
                for (int f = 0; f < 3; f++) neu_add(s[f], cp[f], val[q][f]);
            }
    }
    const int nodec = g.nodec;
    if (!any) {
        if (DIAG)
            for (int c = 0; c < 3; c++) N.nvm[c * nodec + n] = 0.f;
        return;
    }
    float m = N.nm[n];
    float v[3] = {0.f, 0.f, 0.f};
    float p[3] = {s[0] + cp[0], s[1] + cp[1], s[2] + cp[2]};
    if ((double)m > sc.mtol) {
#pragma unroll
        for (int c = 0; c < 3; c++) v[c] = safe_div(p[c], m, sc.m_small);
    }
    uint8_t mk = bc.mask ? bc.mask[n] : 0;
    if (mk) {
#pragma unroll
        for (int c = 0; c < 3; c++)
            if (mk & (1 << c)) v[c] = (float)(bc.val[c * nodec + n] * sc.facl);
    }
#pragma unroll
    for (int c = 0; c < 3; c++) N.nvm[c * nodec + n] = v[c];
    if (DIAG) {
#pragma unroll
        for (int c = 0; c < 3; c++) N.np[c * nodec + n] = p[c];
    }
}

// ------------------------------------------------------ particle update
// Stress update in FP32 from H = F - I and the increment D = dF - I
// (semantics of ConstitutiveModel::UpdateStress, constitutive_model.cpp:287-305,
// rewritten without cancellation against the identity).
// J2 plasticity with linear isotropic hardening, small-increment form:
// elastic predictor sigma_tr = sigma_n + C:de, radial return onto
// q = sy + H eps_p (Simo & Hughes 1998, Box 3.2).
struct J2Trial {
    float st[6];          // trial stress
    float pm, d0, d1, d2;  // trial pressure and normal deviators
    float q, f;           // von Mises stress and yield function of the trial state
    float mu, lmd, E, nu, sy, Hh, ep;
};

__device__ __forceinline__ void j2_trial(const double *p, const float D[9], const float s[6], float ep, J2Trial &t) {
    const float E = (float)p[0], nu = (float)p[1];
    t.E = E;
    t.nu = nu;
    t.sy = (float)p[2];
    t.Hh = (float)p[3];
    t.ep = ep;
    t.lmd = E * nu / ((1.f + nu) * (1.f - 2.f * nu));
    t.mu = E / (2.f * (1.f + nu));
    const float lmd = t.lmd, mu = t.mu;
    float de[6] = {D[0], D[4], D[8], D[5] + D[7], D[6] + D[2], D[1] + D[3]};
    float tr = de[0] + de[1] + de[2];
    t.st[0] = s[0] + lmd * tr + 2.f * mu * de[0];
    t.st[1] = s[1] + lmd * tr + 2.f * mu * de[1];
    t.st[2] = s[2] + lmd * tr + 2.f * mu * de[2];
    t.st[3] = s[3] + mu * de[3];
    t.st[4] = s[4] + mu * de[4];
    t.st[5] = s[5] + mu * de[5];
    t.pm = (t.st[0] + t.st[1] + t.st[2]) * (1.f / 3.f);
    t.d0 = t.st[0] - t.pm;
    t.d1 = t.st[1] - t.pm;
    t.d2 = t.st[2] - t.pm;
    float j2 = 0.5f * (t.d0 * t.d0 + t.d1 * t.d1 + t.d2 * t.d2) + t.st[3] * t.st[3] + t.st[4] * t.st[4] +
               t.st[5] * t.st[5];
    t.q = sqrtf(3.f * j2);
    t.f = t.q - (t.sy + t.Hh * ep);
}

// exact return (or the elastic trial state); eps_p receives the new value
__device__ __forceinline__ void j2_return(const J2Trial &t, float s[6], float *eps_p) {
    if (t.f <= 0.f) {
        for (int c = 0; c < 6; c++) s[c] = t.st[c];
    } else {
        float dg = t.f / (3.f * t.mu + t.Hh);
        float scl = 1.f - 3.f * t.mu * dg / t.q;
        s[0] = t.pm + t.d0 * scl;
        s[1] = t.pm + t.d1 * scl;
        s[2] = t.pm + t.d2 * scl;
        s[3] = t.st[3] * scl;
        s[4] = t.st[4] * scl;
        s[5] = t.st[5] * scl;
        if (eps_p) *eps_p = t.ep + dg;
    }
}

__device__ __forceinline__ void stress32(const MaterialDev &md, const float H[9], const float D[9], double J,
                                         float s[6], float *eps_p, float inv_dt) {
    const double *p = md.prop;
    switch (md.model) {
    case -1:
        for (int q = 0; q < 6; q++) s[q] = 0.f;
        break;
    case 0: {  // hypoelastic increment, dF - I = D
        float E = (float)p[0], nu = (float)p[1];
        float C0 = (E * (1.f - nu)) / ((1.f - 2.f * nu) * (1.f + nu));
        float C1 = (E * nu) / ((1.f - 2.f * nu) * (1.f + nu));
        float C2 = E / (2.f * (1.f + nu));
        float e0 = D[0], e1 = D[4], e2 = D[8];
        s[0] += C0 * e0 + C1 * (e1 + e2);
        s[1] += C0 * e1 + C1 * (e0 + e2);
        s[2] += C0 * e2 + C1 * (e0 + e1);
        s[3] += C2 * (D[5] + D[7]);
        s[4] += C2 * (D[6] + D[2]);
        s[5] += C2 * (D[1] + D[3]);
        break;
    }
    case 1: {  // Neo-Hookean: sigma = (mu (b - I) + lambda ln J I) / J, b - I = H + H^T + H H^T
        float E = (float)p[0], nu = (float)p[1];
        float lmd = nu * E / ((1.f + nu) * (1.f - 2.f * nu));
        float mu = E / (2.f * (1.f + nu));
        float lnj = log1pf((float)(J - 1.0));
        float invJ = (float)(1.0 / J);
        float B[6];
        // B = H + H^T + H H^T in Voigt order xx,yy,zz,yz,xz,xy
        B[0] = 2.f * H[0] + (H[0] * H[0] + H[1] * H[1] + H[2] * H[2]);
        B[1] = 2.f * H[4] + (H[3] * H[3] + H[4] * H[4] + H[5] * H[5]);
        B[2] = 2.f * H[8] + (H[6] * H[6] + H[7] * H[7] + H[8] * H[8]);
        B[3] = (H[5] + H[7]) + (H[3] * H[6] + H[4] * H[7] + H[5] * H[8]);
        B[4] = (H[2] + H[6]) + (H[0] * H[6] + H[1] * H[7] + H[2] * H[8]);
        B[5] = (H[1] + H[3]) + (H[0] * H[3] + H[1] * H[4] + H[2] * H[5]);
        float vol = lmd * lnj;
        s[0] = (mu * B[0] + vol) * invJ;
        s[1] = (mu * B[1] + vol) * invJ;
        s[2] = (mu * B[2] + vol) * invJ;
        s[3] = mu * B[3] * invJ;
        s[4] = mu * B[4] * invJ;
        s[5] = mu * B[5] * invJ;
        break;
    }
    case 2: {  // St.Venant-Kirchhoff: sigma = (mu (B + B^2) + lambda trE (I + B)) / J with b = I + B
        float E = (float)p[0], nu = (float)p[1];
        float lmd = nu * E / ((1.f + nu) * (1.f - 2.f * nu));
        float mu = E / (2.f * (1.f + nu));
        float invJ = (float)(1.0 / J);
        float Bm[9];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                Bm[3 * a + b] = H[3 * a + b] + H[3 * b + a] +
                                (H[3 * a] * H[3 * b] + H[3 * a + 1] * H[3 * b + 1] + H[3 * a + 2] * H[3 * b + 2]);
        float trE = 0.5f * (Bm[0] + Bm[4] + Bm[8]);
        float sg[9];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                float b2 = Bm[3 * a] * Bm[b] + Bm[3 * a + 1] * Bm[3 + b] + Bm[3 * a + 2] * Bm[6 + b];
                float eye = (a == b) ? 1.f : 0.f;
                sg[3 * a + b] = (mu * (Bm[3 * a + b] + b2) + lmd * trE * (eye + Bm[3 * a + b])) * invJ;
            }
        s[0] = sg[0];
        s[1] = sg[4];
        s[2] = sg[8];
        s[3] = sg[5];
        s[4] = sg[2];
        s[5] = sg[1];
        break;
    }
    case 3: {  // Mooney-Rivlin (constitutive_model.cpp:187-285)
        // The reference's form 2/J ((c1 + c2 I1) b - c2 b^2) - 2/(3J) (c1 I1 + 2 c2 I2) I + hp I
        // (b of the isochoric F) subtracts terms of size c1 + c2 down to the stress. Written with
        // B = b - I, its deviator is 2/J ((c1 + c2 (1 + tr B)) dev B - c2 dev B^2): the same
        // quantity with nothing cancelling in FP32.
        float c1 = (float)p[0], c2 = (float)p[1], bulk = (float)p[2];
        float Jf = (float)J;
        const float lj = log1pf((float)(J - 1.0));
        const float E = expm1f(-2.f / 3.f * lj);  // J^(-2/3) - 1
        float B[9];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                const float HHt = H[3 * a] * H[3 * b] + H[3 * a + 1] * H[3 * b + 1] + H[3 * a + 2] * H[3 * b + 2];
                B[3 * a + b] = (1.f + E) * (H[3 * a + b] + H[3 * b + a] + HHt) + (a == b ? E : 0.f);
            }
        float B2[9];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                B2[3 * a + b] = B[3 * a] * B[b] + B[3 * a + 1] * B[3 + b] + B[3 * a + 2] * B[6 + b];
        const float trB = B[0] + B[4] + B[8], trB2 = B2[0] + B2[4] + B2[8];
        const float k1 = c1 + c2 * (1.f + trB);
        const float hp = bulk * lj / Jf;
        float sg[9];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                const float dB = B[3 * a + b] - (a == b ? trB / 3.f : 0.f);
                const float dB2 = B2[3 * a + b] - (a == b ? trB2 / 3.f : 0.f);
                sg[3 * a + b] = 2.f * (k1 * dB - c2 * dB2) / Jf + (a == b ? hp : 0.f);
            }
        s[0] = sg[0];
        s[1] = sg[4];
        s[2] = sg[8];
        s[3] = sg[5];
        s[4] = sg[2];
        s[5] = sg[1];
        break;
    }
    case 20: {  // Tait equation of state + Newtonian viscosity; p without cancellation near J = 1
        const float rho0 = (float)p[0], c = (float)p[1], gam = (float)p[2], mu = (float)p[3];
        const float B = rho0 * c * c / gam;
        const float pr = B * expm1f(-gam * log1pf((float)(J - 1.0)));
        const float tr = (D[0] + D[4] + D[8]) * (1.f / 3.f);
        const float m2 = 2.f * mu * inv_dt, m1 = mu * inv_dt;
        s[0] = -pr + m2 * (D[0] - tr);
        s[1] = -pr + m2 * (D[4] - tr);
        s[2] = -pr + m2 * (D[8] - tr);
        s[3] = m1 * (D[5] + D[7]);
        s[4] = m1 * (D[6] + D[2]);
        s[5] = m1 * (D[1] + D[3]);
        break;
    }
    case 10: {  // J2 radial return, small-increment form
        J2Trial t;
        j2_trial(p, D, s, eps_p ? *eps_p : 0.f, t);
        j2_return(t, s, eps_p);
        break;
    }
    default: break;
    }
}

__device__ __forceinline__ uint32_t mix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

// 13 surrogate inputs: strain increment (engineering shear), elastic strain
// of the old stress, old equivalent plastic strain
__device__ __forceinline__ void j2_features(const J2Trial &t, const float D[9], const float s_old[6], float x[16]) {
    x[0] = D[0];
    x[1] = D[4];
    x[2] = D[8];
    x[3] = D[5] + D[7];
    x[4] = D[6] + D[2];
    x[5] = D[1] + D[3];
    const float tr = s_old[0] + s_old[1] + s_old[2];
    const float a = (1.f + t.nu) / t.E, b = t.nu / t.E, im = 1.f / t.mu;
#pragma unroll
    for (int c = 0; c < 3; c++) x[6 + c] = a * s_old[c] - b * tr;
#pragma unroll
    for (int c = 3; c < 6; c++) x[6 + c] = s_old[c] * im;
    x[12] = t.ep;
    x[13] = x[14] = x[15] = 0.f;
}

// Residual of the discrete return-mapping equations at a predicted state
// (stress s, plastic increment dg), relative to the yield stress:
// max(|q(s) - sy - H (ep + dg)|, max_c |s_c - s_rr(dg)_c|) / sy, where
// s_rr(dg) is the radial-return stress for that increment.
__device__ __forceinline__ float j2_residual(const J2Trial &t, const float s[6], float dg) {
    if (!(dg >= 0.f)) return 1e30f;
    const float pm = (s[0] + s[1] + s[2]) * (1.f / 3.f);
    const float e0 = s[0] - pm, e1 = s[1] - pm, e2 = s[2] - pm;
    const float j2 = 0.5f * (e0 * e0 + e1 * e1 + e2 * e2) + s[3] * s[3] + s[4] * s[4] + s[5] * s[5];
    const float q = sqrtf(3.f * j2);
    float r = fabsf(q - (t.sy + t.Hh * (t.ep + dg)));
    const float scl = 1.f - 3.f * t.mu * dg / t.q;
    const float rr[6] = {t.pm + t.d0 * scl, t.pm + t.d1 * scl, t.pm + t.d2 * scl,
                         t.st[3] * scl,     t.st[4] * scl,     t.st[5] * scl};
#pragma unroll
    for (int c = 0; c < 6; c++) r = fmaxf(r, fabsf(s[c] - rr[c]));
    return r / t.sy;
}

// tsum[t] += d (modulo 2^64) for the lanes with act set; one atomic per wave
// when all of them share the tile. Called by every lane of the wave.
__device__ __forceinline__ void tile_total_add(uint64_t *tsum, bool act, uint32_t t, unsigned long long d) {
    const unsigned long long m = __ballot(act);
    if (!m) return;
    const int f = __ffsll(m) - 1;
    const uint32_t t0 = (uint32_t)__shfl((int)t, f, 64);
    if (__ballot(act && t != t0) == 0ull) {
        unsigned long long sum = act ? d : 0ull;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
        if ((int)(threadIdx.x & 63) == f) atomicAdd((unsigned long long *)&tsum[t0], sum);
    } else if (act) {
        atomicAdd((unsigned long long *)&tsum[t], d);
    }
}

// det(I + H) in FP64 of an FP32 displacement gradient. No contraction, so
// that the update kernel and the J refresh give the same bits.
__device__ __forceinline__ double det_I_plus_H(const float H[9]) {
#pragma clang fp contract(off)
    double A[9];
#pragma unroll
    for (int q = 0; q < 9; q++) A[q] = (double)H[q] + ((q == 0 || q == 4 || q == 8) ? 1.0 : 0.0);
    return A[0] * (A[4] * A[8] - A[5] * A[7]) + A[1] * (A[5] * A[6] - A[3] * A[8]) +
           A[2] * (A[3] * A[7] - A[4] * A[6]);
}

template <int SUR>
__global__ void __launch_bounds__(256) k32_g2p_update(GridGeom g, PartF32 P, NodesF32 N, NodeTables T,
                                                       MaterialTable mt, StepConsts sc, uint32_t *err,
                                                       SurrogateArgs sa, BinFuse bf) {
    const int slot0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (bf.cnt) {
        for (int i = slot0; i < bf.ntiles; i += gridDim.x * blockDim.x) bf.tiles[i] = 0u;
        if (slot0 == 0) {
            bf.ctr->n_active = 0u;
        }
    }
    // Wave-wide steps (the network, the run-aggregated cell counts) need every
    // lane: no lane leaves early; invalid lanes compute on slot 0 / cell 0
    // and store nothing.
    bool valid = slot0 < P.n;
    const int slot = valid ? slot0 : 0;
    uint32_t key = P.key[slot];
    if (key >= (uint32_t)g.nelem) valid = false;
    if (!valid) key = 0;
    const size_t ld = P.ld;
    const int nodec = g.nodec;
    const int nexy = g.nex[0] * g.nex[1];
    int rem, ix, iy;
    const int iz = fdivmod((int)key, nexy, T.inv_nexy, rem);
    iy = fdivmod(rem, g.nex[0], T.inv_nex, ix);
    float r[3];
#pragma unroll
    for (int a = 0; a < 3; a++) r[a] = P.r[a * ld + slot];
    const float fx = r[0], fy = r[1], fz = r[2];
    const float wx[2] = {1.f - fx, fx}, wy[2] = {1.f - fy, fy}, wz[2] = {1.f - fz, fz};
    const float inv0 = T.inv_dx[0], inv1 = T.inv_dx[1], inv2 = T.inv_dx[2];
    const int n0 = ix + g.nxc[0] * iy + g.nxc[0] * g.nxc[1] * iz;
    float Dv[9];
#pragma unroll
    for (int q = 0; q < 9; q++) Dv[q] = 0.f;
    float dxp[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 2; k++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const int nid = n0 + i + j * g.nxc[0] + k * g.nxc[0] * g.nxc[1];
                const float v0 = N.nvm[nid], v1 = N.nvm[nodec + nid], v2 = N.nvm[2 * nodec + nid];
                const float wyz = wy[j] * wz[k];
                const float Nw = wx[i] * wyz;
                const float gx = (i ? inv0 : -inv0) * wyz;
                const float gy = (j ? inv1 : -inv1) * (wx[i] * wz[k]);
                const float gz = (k ? inv2 : -inv2) * (wx[i] * wy[j]);
                Dv[0] += v0 * gx;
                Dv[1] += v0 * gy;
                Dv[2] += v0 * gz;
                Dv[3] += v1 * gx;
                Dv[4] += v1 * gy;
                Dv[5] += v1 * gz;
                Dv[6] += v2 * gx;
                Dv[7] += v2 * gy;
                Dv[8] += v2 * gz;
                dxp[0] += Nw * v0;
                dxp[1] += Nw * v1;
                dxp[2] += Nw * v2;
            }
    float D[9];
#pragma unroll
    for (int q = 0; q < 9; q++) D[q] = sc.dt * Dv[q];
    // H_new = (I + D)(I + H) - I = H + D + D H
    float H[9], Hn[9];
#pragma unroll
    for (int q = 0; q < 9; q++) H[q] = __builtin_nontemporal_load(&P.H[q * ld + slot]);  // read once
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++)
            Hn[3 * a + b] = (H[3 * a + b] + D[3 * a + b]) +
                            (D[3 * a] * H[b] + D[3 * a + 1] * H[3 + b] + D[3 * a + 2] * H[6 + b]);
    const double J = det_I_plus_H(Hn);
    const double V = J * (double)P.V0[slot];
    // position: offset in cell units, carried into the cell index when it leaves [0, 1)
    int ci[3] = {ix, iy, iz};
    bool inside = true;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float ra = r[a] + (sc.dt * dxp[a]) * T.inv_dx[a];
        const float fl = floorf(ra);
        ra -= fl;  // exact for |fl| <= 1 (Sterbenz)
        ci[a] += (int)fl;
        if (ra >= 1.f) {  // -tiny + 1 rounded up to 1: the particle sits on the next cell's lower face
            ra = 0.f;
            ci[a] += 1;
        }
        r[a] = ra;
        inside = inside && ci[a] >= 0 && ci[a] < g.nex[a];
    }
    const int mat = P.matid ? P.matid[slot] : 0;
    const MaterialDev &md = mt.m[mat];
    // only the rate forms (linear elastic increment, J2) read the old stress
    const bool rate_form = (md.model == 0 || md.model == 10);
    float s[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (rate_form) {
#pragma unroll
        for (int q = 0; q < 6; q++) s[q] = P.sig[q * ld + slot];
    }
    float ep = (md.model == 10 && P.eps_p) ? P.eps_p[slot] : 0.f;
    if constexpr (SUR == 0) {
        stress32(md, Hn, D, J, s, &ep, 1.f / sc.dt);
    } else {
        const bool j2 = md.model == 10;
        J2Trial t;
        bool plastic = false;
        if (j2) {
            j2_trial(md.prop, D, s, ep, t);
            plastic = valid && t.f > 0.f;
        } else {
            stress32(md, Hn, D, J, s, &ep, 1.f / sc.dt);
        }
        const int lane = threadIdx.x & 63;
        if constexpr (SUR == 1) {
            float s_old[6];
#pragma unroll
            for (int c = 0; c < 6; c++) s_old[c] = s[c];
            if (j2) j2_return(t, s, &ep);
            const uint32_t id = (uint32_t)sa.pid[slot];
            const bool want = plastic && mix32(id * 0x9E3779B9u ^ mix32(sa.step_tag + 0x51ED27u)) <= sa.sel_threshold;
            const unsigned long long m = __ballot(want);
            if (m) {
                const int leader = __ffsll((unsigned long long)m) - 1;
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(&sa.ctr[2], (unsigned long long)__popcll(m));
                base = __shfl(base, leader, 64);
                if (want) {
                    const unsigned long long idx = base + __popcll(m & ((1ull << lane) - 1ull));
                    if (idx < sa.cap) {
                        float x[16];
                        j2_features(t, D, s_old, x);
                        float y[16];
#pragma unroll
                        for (int c = 0; c < 6; c++) y[c] = t.st[c] - s[c];
                        y[6] = ep - t.ep;
#pragma unroll
                        for (int c = 7; c < 16; c++) y[c] = 0.f;
                        float4 *xi = (float4 *)(sa.log_in + idx * 16), *yo = (float4 *)(sa.log_tgt + idx * 16);
#pragma unroll
                        for (int c = 0; c < 4; c++) {
                            xi[c] = make_float4(x[4 * c], x[4 * c + 1], x[4 * c + 2], x[4 * c + 3]);
                            yo[c] = make_float4(y[4 * c], y[4 * c + 1], y[4 * c + 2], y[4 * c + 3]);
                        }
                        sa.log_key[idx] = sa.step_tag * (uint32_t)P.n + id;
                    }
                }
            }
        } else {
            // exact update for every particle; the plastic trial states of this
            // block are compacted, in slot order, into the block's own queue
            // region for the network pass that follows this kernel
            float x[16];
            if (plastic) j2_features(t, D, s, x);
            if (j2) j2_return(t, s, &ep);
            __shared__ uint32_t wcnt[kSurRegion / 64];
            const int w = (int)(threadIdx.x >> 6);
            const unsigned long long m = __ballot(plastic);
            if (lane == 0) wcnt[w] = (uint32_t)__popcll(m);
            __syncthreads();
            uint32_t off = 0, tot = 0;
#pragma unroll
            for (int i = 0; i < kSurRegion / 64; i++) {
                off += i < w ? wcnt[i] : 0u;
                tot += wcnt[i];
            }
            if (threadIdx.x == 0) sa.bcount[blockIdx.x] = tot;
            {
                const uint32_t qi = blockIdx.x * (uint32_t)kSurRegion + off +
                                    (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                if (plastic && qi < sa.qcap) {
                    float4 *row = (float4 *)(sa.queue + (size_t)qi * kSurQueueRow);
                    row[0] = make_float4(__int_as_float(slot), x[0], x[1], x[2]);
                    row[1] = make_float4(x[3], x[4], x[5], x[6]);
                    row[2] = make_float4(x[7], x[8], x[9], x[10]);
                    row[3] = make_float4(x[11], x[12], t.st[0], t.st[1]);
                    row[4] = make_float4(t.st[2], t.st[3], t.st[4], t.st[5]);
                    row[5] = make_float4(t.E, t.nu, t.sy, t.Hh);
                }
            }
        }
    }
    const uint32_t nk = inside ? (uint32_t)(ci[0] + ci[1] * g.nex[0] + ci[2] * nexy) : (uint32_t)g.nelem;
    if (bf.cnt) {
        if (bf.delta) {
            // kept counts: only a particle that changes cell moves one count,
            // and the packed tile totals (non-empty cells, particles) follow
            const bool mover = valid && nk != key;
            const bool lands = mover && nk < (uint32_t)g.nelem;
            uint32_t o_from = 0u, o_to = 1u;
            if (mover) o_from = atomicSub(&bf.cnt[key], 1u);
            if (lands) o_to = atomicAdd(&bf.cnt[nk], 1u);
            const unsigned long long d_from = 1ull + (o_from == 1u ? (1ull << 32) : 0ull);
            const unsigned long long d_to = 1ull + (o_to == 0u ? (1ull << 32) : 0ull);
            tile_total_add(bf.tsum, mover, key >> kBinTileShift, 0ull - d_from);
            tile_total_add(bf.tsum, lands, nk >> kBinTileShift, d_to);
        } else {
            const uint32_t ck = valid ? nk : 0xFFFFFFFFu;
            int len, head;
            const bool h = wave_runs(ck, threadIdx.x & 63, &len, &head);
            if (h && ck < (uint32_t)g.nelem) atomicAdd(&bf.cnt[ck], (uint32_t)len);
        }
    }
    if (!valid) return;
    if (md.model == 10 && P.eps_p) P.eps_p[slot] = ep;
    // streamed out with non-temporal stores: the particle state is next read
    // a whole step later, and stores that stay dirty in L2 slow the next
    // kernels (P2G cells 77 -> 66 us on the 4M-particle cube)
#pragma unroll
    for (int q = 0; q < 9; q++) __builtin_nontemporal_store(Hn[q], &P.H[q * ld + slot]);
#pragma unroll
    for (int q = 0; q < 6; q++) __builtin_nontemporal_store(s[q], &P.sig[q * ld + slot]);
#pragma unroll
    for (int a = 0; a < 3; a++) __builtin_nontemporal_store(r[a], &P.r[a * ld + slot]);
    __builtin_nontemporal_store((float)V, &P.V32[slot]);
    if (!inside) {
        // frozen from here on: keep its FP64 det F and volume
        P.J[slot] = J;
        P.V[slot] = V;
    }
    if (nk != key) P.key[slot] = nk;  // most particles stay in their cell
    uint32_t e = 0;
    if (!inside) e |= kErrOutOfGrid;
    if (!(J > 0.0)) e |= kErrJacobian;
    if (!isfinite(r[0] + r[1] + r[2] + s[0] + s[1] + s[2])) e |= kErrNonFinite;
    if (e) atomicOr(err, e);
}

// queue row qi into registers (zeros when not live)
__device__ __forceinline__ void sur_load_row(const float *queue, uint32_t qi, bool live, float r[kSurQueueRow]) {
    if (live) {
        const float4 *row = (const float4 *)(queue + (size_t)qi * kSurQueueRow);
#pragma unroll
        for (int c = 0; c < kSurQueueRow / 4; c++) {
            const float4 v = row[c];
            r[4 * c] = v.x;
            r[4 * c + 1] = v.y;
            r[4 * c + 2] = v.z;
            r[4 * c + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int c = 0; c < kSurQueueRow; c++) r[c] = 0.f;
    }
}

// Network pass over the queued plastic J2 trial states of a step. A tile
// slot is (region, k): rows [64 k, 64 k + 64) of queue region k / 4 (the
// plastic states of one update block). The blocks are persistent (the
// network weights are staged in LDS once per block); each wave visits every
// nw-th tile slot, finds the non-empty ones of its next 64 slots with one
// load of the region counts per lane, and loads the rows of its next tile
// while the current tile runs through the network on the BF16 MFMA. A
// prediction is accepted where the return-mapping residual is within the
// tolerance; the exact result written by the update kernel stays otherwise.
// The evaluated / rejected tallies go to per-region counters: a region's
// tiles may be spread over several waves, so they are added atomically, but
// to distinct addresses.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k32_j2_surrogate(PartF32 P, SurrogateArgs sa) {
    __shared__ MlpInferSmem sw;
    __shared__ float scl[64];
    __shared__ __align__(16) uint16_t xs[4][kMlpWaveScratchBytes / 2];
    mlp_load_infer_smem(sw, sa.net);
    if (threadIdx.x < 64) scl[threadIdx.x] = sa.net_scaling[threadIdx.x];
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t nslots = 4u * sa.nregions;
    const uint32_t nw = gridDim.x * 4u;
    const uint32_t gw = blockIdx.x * 4u + (uint32_t)w;
    for (uint32_t base = gw; base < nslots; base += 64u * nw) {
        // lane l checks slot base + l nw
        const uint32_t my = base + (uint32_t)lane * nw;
        uint32_t c = 0;
        if (my < nslots) c = min(sa.bcount[my >> 2], (uint32_t)kSurRegion);
        const int myn = (int)c - 64 * (int)(my & 3u);  // rows of this lane's tile
        unsigned long long todo = __ballot(myn > 0);
        if (!todo) continue;
        int b = __ffsll(todo) - 1;
        todo &= todo - 1ull;
        int n_cur = __shfl(myn, b, 64);
        uint32_t q_cur = ((base + (uint32_t)b * nw) >> 2) * (uint32_t)kSurRegion + 64u * ((base + (uint32_t)b * nw) & 3u);
        float r[kSurQueueRow];
        sur_load_row(sa.queue, q_cur + (uint32_t)lane, lane < n_cur, r);
        for (;;) {
            // next tile of this batch: its rows are loaded while this one runs
            const bool more = todo != 0ull;
            int n_nxt = 0;
            uint32_t q_nxt = 0;
            float rn[kSurQueueRow];
            if (more) {
                const int bn = __ffsll(todo) - 1;
                todo &= todo - 1ull;
                n_nxt = __shfl(myn, bn, 64);
                const uint32_t sn = base + (uint32_t)bn * nw;
                q_nxt = (sn >> 2) * (uint32_t)kSurRegion + 64u * (sn & 3u);
            }
            sur_load_row(sa.queue, q_nxt + (uint32_t)lane, more && lane < n_nxt, rn);
            const bool live = lane < n_cur;
            float xn[16];
#pragma unroll
            for (int f = 0; f < 16; f++) xn[f] = (f < kMlpInUsed) ? (r[1 + f] - scl[f]) * scl[16 + f] : 0.f;
            float y[kMlpOutUsed];
            mlp_wave_infer(sw, xn, y, xs[w], lane);
            bool accept = false;
            if (live) {
                const int slot = __float_as_int(r[0]);
                // trial state from the queued trial stress (same arithmetic as j2_trial)
                J2Trial t;
                const float E = r[20], nu = r[21];
                t.E = E;
                t.nu = nu;
                t.sy = r[22];
                t.Hh = r[23];
                t.ep = r[13];
                t.lmd = E * nu / ((1.f + nu) * (1.f - 2.f * nu));
                t.mu = E / (2.f * (1.f + nu));
#pragma unroll
                for (int cc = 0; cc < 6; cc++) t.st[cc] = r[14 + cc];
                t.pm = (t.st[0] + t.st[1] + t.st[2]) * (1.f / 3.f);
                t.d0 = t.st[0] - t.pm;
                t.d1 = t.st[1] - t.pm;
                t.d2 = t.st[2] - t.pm;
                const float j2 = 0.5f * (t.d0 * t.d0 + t.d1 * t.d1 + t.d2 * t.d2) + t.st[3] * t.st[3] +
                                 t.st[4] * t.st[4] + t.st[5] * t.st[5];
                t.q = sqrtf(3.f * j2);
                float ss[6];
#pragma unroll
                for (int cc = 0; cc < 6; cc++) ss[cc] = t.st[cc] - (y[cc] * scl[48 + cc] + scl[32 + cc]);
                const float dg = y[6] * scl[54] + scl[38];
                if (j2_residual(t, ss, dg) <= sa.tol) {
                    accept = true;
                    const size_t ld = P.ld;
#pragma unroll
                    for (int cc = 0; cc < 6; cc++) P.sig[cc * ld + slot] = ss[cc];
                    if (P.eps_p) P.eps_p[slot] = t.ep + dg;
                }
            }
            const uint32_t ev = (uint32_t)__popcll(__ballot(live));
            const uint32_t rj = (uint32_t)__popcll(__ballot(live && !accept));
            if (lane == 0) {
                const uint32_t rg = q_cur / (uint32_t)kSurRegion;
                atomicAdd(&sa.rstat[2 * rg], ev);
                if (rj) atomicAdd(&sa.rstat[2 * rg + 1], rj);
            }
            if (!more) break;
            n_cur = n_nxt;
            q_cur = q_nxt;
#pragma unroll
            for (int cc = 0; cc < kSurQueueRow; cc++) r[cc] = rn[cc];
        }
    }
}

__global__ void k32_permute(PartF32 src, PartF32 dst, const uint32_t *__restrict__ perm, const int *__restrict__ pid_src,
                            int *__restrict__ pid_dst) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= src.n) return;
    const uint32_t o = perm[s];
    const size_t ls = src.ld, lt = dst.ld;
    for (int a = 0; a < 3; a++) {
        dst.r[a * lt + s] = src.r[a * ls + o];
        dst.v[a * lt + s] = src.v[a * ls + o];
        if (src.trac) dst.trac[a * lt + s] = src.trac[a * ls + o];
    }
    for (int q = 0; q < 9; q++) dst.H[q * lt + s] = src.H[q * ls + o];
    for (int q = 0; q < 6; q++) dst.sig[q * lt + s] = src.sig[q * ls + o];
    dst.m[s] = src.m[o];
    dst.V0[s] = src.V0[o];
    dst.V32[s] = src.V32[o];
    dst.V[s] = src.V[o];
    dst.J[s] = src.J[o];
    dst.key[s] = src.key[o];
    if (src.eps_p) dst.eps_p[s] = src.eps_p[o];
    if (src.matid) dst.matid[s] = src.matid[o];
    pid_dst[s] = pid_src[o];
}

__global__ void k32_refresh_jv(GridGeom g, PartF32 P) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.n || P.key[s] >= (uint32_t)g.nelem) return;
    float H[9];
#pragma unroll
    for (int q = 0; q < 9; q++) H[q] = P.H[q * P.ld + s];
    const double J = det_I_plus_H(H);
    P.J[s] = J;
    P.V[s] = J * (double)P.V0[s];
}

}  // namespace

void launch_f32_refresh_jv(hipStream_t s, const GridGeom &g, const PartF32 &P) {
    if (P.n > 0) hipLaunchKernelGGL(k32_refresh_jv, dim3(cdiv(P.n, 256)), dim3(256), 0, s, g, P);
}

static int cell_groups_grid(const GridGeom &g) {
    int ng = cdiv(g.nelem, 16);
    return ng < 8192 ? ng : 8192;
}

template <int LPC, bool PF>
static void p2g_cells_launch_pf(hipStream_t s, int ng, bool diag, bool trac, const GridGeom &g, const PartF32 &P,
                                const CellLists &L, const NodeTables &T, float *partials) {
    if (diag) {
        if (trac)
            hipLaunchKernelGGL((k32_p2g_cells<LPC, true, true, PF>), dim3(ng), dim3(kSeg), 0, s, g, P, L, T, partials);
        else
            hipLaunchKernelGGL((k32_p2g_cells<LPC, true, false, PF>), dim3(ng), dim3(kSeg), 0, s, g, P, L, T,
                               partials);
    } else {
        if (trac)
            hipLaunchKernelGGL((k32_p2g_cells<LPC, false, true, PF>), dim3(ng), dim3(kSeg), 0, s, g, P, L, T,
                               partials);
        else
            hipLaunchKernelGGL((k32_p2g_cells<LPC, false, false, PF>), dim3(ng), dim3(kSeg), 0, s, g, P, L, T,
                               partials);
    }
}

template <int LPC>
static void p2g_cells_launch(hipStream_t s, int ng, bool diag, bool trac, bool pf, const GridGeom &g,
                             const PartF32 &P, const CellLists &L, const NodeTables &T, float *partials) {
    if (pf)
        p2g_cells_launch_pf<LPC, true>(s, ng, diag, trac, g, P, L, T, partials);
    else
        p2g_cells_launch_pf<LPC, false>(s, ng, diag, trac, g, P, L, T, partials);
}

void launch_f32_p2g(hipStream_t s, const GridGeom &g, const PartF32 &P, const NodesF32 &N, const CellLists &L,
                    const BcDense &bc, const NodeTables &T, const TileGeom &tg, const StepConsts &c, float *partials,
                    int diag, hipEvent_t mid) {
    const int ng = c.cells_grid > 0 ? c.cells_grid : cell_groups_grid(g);
    const bool trac = P.trac != nullptr;
    const bool pf = c.cells_prefetch != 0;
    switch (c.lanes_per_cell) {
    case 8: p2g_cells_launch<8>(s, ng, diag, trac, pf, g, P, L, T, partials); break;
    case 4: p2g_cells_launch<4>(s, ng, diag, trac, pf, g, P, L, T, partials); break;
    case 2: p2g_cells_launch<2>(s, ng, diag, trac, pf, g, P, L, T, partials); break;
    default: p2g_cells_launch<1>(s, ng, diag, trac, pf, g, P, L, T, partials); break;
    }
    if (mid) (void)hipEventRecord(mid, s);
    if (diag)
        hipLaunchKernelGGL(k32_p2g_nodes<true>, dim3(tg.ntiles), dim3(256), 0, s, g, L, tg, (const float *)partials, N,
                           bc, c);
    else
        hipLaunchKernelGGL(k32_p2g_nodes<false>, dim3(tg.ntiles), dim3(256), 0, s, g, L, tg, (const float *)partials,
                           N, bc, c);
}

template <int LPC>
static void musl_cells_launch(hipStream_t s, int ng, const GridGeom &g, const PartF32 &P, const CellLists &L,
                              const NodeTables &T, const NodesF32 &N, const StepConsts &c, float *partials) {
    if (c.scheme == 1)
        hipLaunchKernelGGL((k32_musl_cells<LPC, true>), dim3(ng), dim3(kSeg), 0, s, g, P, L, T, N, c, partials);
    else
        hipLaunchKernelGGL((k32_musl_cells<LPC, false>), dim3(ng), dim3(kSeg), 0, s, g, P, L, T, N, c, partials);
}

void launch_f32_g2p_vel_musl(hipStream_t s, const GridGeom &g, const PartF32 &P, const NodesF32 &N,
                             const CellLists &L, const BcDense &bc, const NodeTables &T, const TileGeom &tg,
                             const StepConsts &c, float *partials, int diag, hipEvent_t mid) {
    const int ng = cell_groups_grid(g);
    switch (c.lanes_per_cell) {
    case 8: musl_cells_launch<8>(s, ng, g, P, L, T, N, c, partials); break;
    case 4: musl_cells_launch<4>(s, ng, g, P, L, T, N, c, partials); break;
    case 2: musl_cells_launch<2>(s, ng, g, P, L, T, N, c, partials); break;
    default: musl_cells_launch<1>(s, ng, g, P, L, T, N, c, partials); break;
    }
    if (mid) (void)hipEventRecord(mid, s);
    if (diag)
        hipLaunchKernelGGL(k32_musl_nodes<true>, dim3(tg.ntiles), dim3(256), 0, s, g, L, tg, (const float *)partials, N,
                           bc, c);
    else
        hipLaunchKernelGGL(k32_musl_nodes<false>, dim3(tg.ntiles), dim3(256), 0, s, g, L, tg, (const float *)partials,
                           N, bc, c);
}

void launch_f32_j2_surrogate(hipStream_t s, const PartF32 &P, const SurrogateArgs &sa) {
    if (P.n <= 0) return;
    // persistent blocks: as many as are resident at once (queried once)
    static int resident = 0;
    if (resident == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k32_j2_surrogate, 256, 0);
        resident = std::max(1, cus) * std::max(1, per_cu);
    }
    const int blocks = std::min((int)sa.nregions, resident);
    hipLaunchKernelGGL(k32_j2_surrogate, dim3(blocks), dim3(256), 0, s, P, sa);
}

void launch_f32_g2p_update(hipStream_t s, const GridGeom &g, const PartF32 &P, const NodesF32 &N,
                           const NodeTables &T, const MaterialTable &mt, const StepConsts &c, uint32_t *err,
                           const SurrogateArgs &sa, const BinFuse &bf) {
    if (P.n <= 0) return;
    const dim3 grid(cdiv(P.n, 256)), block(256);
    if (sa.mode == 1)
        hipLaunchKernelGGL(k32_g2p_update<1>, grid, block, 0, s, g, P, N, T, mt, c, err, sa, bf);
    else if (sa.mode == 2)
        hipLaunchKernelGGL(k32_g2p_update<2>, grid, block, 0, s, g, P, N, T, mt, c, err, sa, bf);
    else
        hipLaunchKernelGGL(k32_g2p_update<0>, grid, block, 0, s, g, P, N, T, mt, c, err, sa, bf);
}

__global__ void k_iota(uint32_t *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (uint32_t)i;
}

void launch_iota_u32(hipStream_t s, uint32_t *p, int n) {
    if (n > 0) hipLaunchKernelGGL(k_iota, dim3(cdiv(n, 256)), dim3(256), 0, s, p, n);
}

void launch_f32_permute(hipStream_t s, const PartF32 &src, const PartF32 &dst, const uint32_t *perm,
                        const int *pid_src, int *pid_dst) {
    if (src.n <= 0) return;
    hipLaunchKernelGGL(k32_permute, dim3(cdiv(src.n, 256)), dim3(256), 0, s, src, dst, perm, pid_src, pid_dst);
}

}  // namespace mpm

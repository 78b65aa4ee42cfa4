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
// group of 64 / L consecutive non-empty cells (L lanes per cell). Their
// particles are one contiguous range of the cell-sorted order, which is
// staged through LDS in chunks with coalesced loads; the lanes of a cell then
// fold its particles, in sorted order, into the 8 corner partials of that
// cell (registers). Where the previous cell of the group is the cell's left
// neighbour, the two corner partials of each node between them are added in
// the group, so that node pass reads one value for them. A node pass adds the
// partials around each node in ascending cell order with Neumaier-compensated
// FP32 sums and applies the grid update (velocity, acceleration, BCs) in the
// same pass.
// The MUSL remap uses the same structure, with the G2P velocity of every
// particle computed from the 8 node values of its cell held in registers.
#include "fast32_device.h"

#include <algorithm>
#include <type_traits>

namespace mpm {

namespace {

// Corner partials of the cells kernels. A cell's values are grouped by
// corner group g = y-side + 2 z-side and x-side `side` (0: the cell's left
// corners, 1: its right ones), NQ quantities each. part_entry orders them so
// that after ReduceScatter the lane of sub s holds whole (group, side) runs in
// consecutive registers: L <= 4 lanes per cell hold 4 / L groups with both
// sides (group s 4 / L + gl, run start (side 4 / L + gl) NQ), L = 8 holds side
// s & 1 of group s >> 1 (run start 0).
template <int LPC>
__host__ __device__ constexpr int groups_per_lane() { return LPC <= 4 ? 4 / LPC : 1; }
template <int LPC, int NQ>
__device__ __forceinline__ constexpr int part_entry(int g, int side, int f) {
    constexpr int GPL = groups_per_lane<LPC>();
    return LPC <= 4 ? LPC * ((side * GPL + g % GPL) * NQ + f) + g / GPL : 8 * f + 2 * g + side;
}

// Stored layout: slot side * 4 + g, then the cell, then the NQ quantities
// padded to NQP = 8 (P2G) or 4 (MUSL) floats, so that a cell's run of one
// slot is stored and loaded with 16-byte accesses.
template <int NQ>
__host__ __device__ constexpr int padded_nq() { return NQ <= 4 ? 4 : 8; }

// Stores a cell's corner partials after ReduceScatter, merged along x. A cell
// whose left neighbour is the previous cell of its group adds that cell's
// right-corner runs into its own left-corner ones (the lanes that hold them
// are L apart, or L - 1 for L = 8). Stored: the left-corner runs of every
// non-empty cell, and the right-corner runs only of a cell whose right
// neighbour did not take them (empty, past the row's end, or the first cell
// of the next group) -- the cases in which the node passes read them
// (node_rows4). With L.gmask == 0 no cell merges and every run is stored.
template <int LPC, int NQ>
__device__ __forceinline__ void xmerge_store(float *v, int sub, const CellInfo &si, const CellLists &L, int nex0,
                                             float *__restrict__ part, size_t ne) {
    constexpr int NQP = padded_nq<NQ>();
    constexpr int GPL = groups_per_lane<LPC>();
    const int lane = threadIdx.x;
    const bool xm = L.gsize > 1u;
    const int pc = __shfl_up(si.c, LPC, 64);    // previous cell of the group
    const int nc = __shfl_down(si.c, LPC, 64);  // next cell of the group
    const bool merge = xm && lane >= LPC && si.c > 0 && pc == si.c - 1 && si.ix > 0;
    const bool keep_right = !(xm && lane < kSeg - LPC && nc == si.c + 1 && si.ix + 1 < nex0);
    const bool on = si.ce > si.cs;
    if constexpr (LPC <= 4) {
#pragma unroll
        for (int gl = 0; gl < GPL; gl++)
#pragma unroll
            for (int f = 0; f < NQ; f++) {
                const float up = __shfl_up(v[(GPL + gl) * NQ + f], LPC, 64);
                if (merge) v[gl * NQ + f] += up;
            }
        if (on) {
#pragma unroll
            for (int side = 0; side < 2; side++) {
                if (side == 1 && !keep_right) break;
#pragma unroll
                for (int gl = 0; gl < GPL; gl++) {
                    const int g = sub * GPL + gl;
                    float4 *dst = (float4 *)(part + ((size_t)(side * 4 + g) * ne + si.c) * NQP);
                    const float *run = v + (side * GPL + gl) * NQ;
#pragma unroll
                    for (int h = 0; h < NQP / 4; h++)
                        dst[h] = make_float4(run[4 * h], 4 * h + 1 < NQ ? run[4 * h + 1] : 0.f,
                                             4 * h + 2 < NQ ? run[4 * h + 2] : 0.f,
                                             4 * h + 3 < NQ ? run[4 * h + 3] : 0.f);
                }
            }
        }
    } else {
        const int side = sub & 1, g = sub >> 1;
#pragma unroll
        for (int f = 0; f < NQ; f++) {
            const float up = __shfl_up(v[f], LPC - 1, 64);
            if (merge && side == 0) v[f] += up;
        }
        if (on && (side == 0 || keep_right)) {
            float4 *dst = (float4 *)(part + ((size_t)(side * 4 + g) * ne + si.c) * NQP);
#pragma unroll
            for (int h = 0; h < NQP / 4; h++)
                dst[h] = make_float4(v[4 * h], 4 * h + 1 < NQ ? v[4 * h + 1] : 0.f, 4 * h + 2 < NQ ? v[4 * h + 2] : 0.f,
                                     4 * h + 3 < NQ ? v[4 * h + 3] : 0.f);
        }
    }
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
        // quantities per corner: m, p, force (f = 0..6) and the volume (7)
        constexpr int NQ = DIAG ? 8 : 7;
        constexpr int NV = 8 * NQ;
        float v[NV];
#pragma unroll
        for (int k = 0; k < 2; k++)
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int ii = 0; ii < 2; ii++) {
                    const int q = ii + 2 * j + 4 * k;
#pragma unroll
                    for (int f = 0; f < 4; f++) v[part_entry<LPC, NQ>(j + 2 * k, ii, f)] = acc[q][f];
#pragma unroll
                    for (int a = 0; a < 3; a++) {
                        const float fx_ = ii ? FX[a][j + 2 * k] : -FX[a][j + 2 * k];
                        const float fy_ = j ? FY[a][ii + 2 * k] : -FY[a][ii + 2 * k];
                        const float fz_ = k ? FZ[a][ii + 2 * j] : -FZ[a][ii + 2 * j];
                        v[part_entry<LPC, NQ>(j + 2 * k, ii, 4 + a)] = (fx_ + fy_) + fz_ + (TRAC ? tr[q][a] : 0.f);
                    }
                    if (DIAG) v[part_entry<LPC, NQ>(j + 2 * k, ii, 7)] = vof[q];
                }
        ReduceScatter<1, LPC, NV>::run(v, sub);
        xmerge_store<LPC, NQ>(v, sub, si, L, g.nex[0], part, (size_t)g.nelem);
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


// The <= 4 cell rows around node (i, j, k) in ascending cell order, each with
// the cell left (i - 1) and right (i) of the node, and which of their partials
// to read: the right cell's left-corner entries (with the left cell's
// right-corner ones added when the cells kernel merged the pair: both
// non-empty and the right cell not the first of its group) and, when the pair
// was not merged, the left cell's right-corner entries. The occupancy and
// group-start words are loaded from clamped addresses, all at once.
struct NodeRows {
    int cl[4], cr[4];
    bool rl[4], rr[4];
};
__device__ __forceinline__ bool node_rows4(const GridGeom &g, const CellLists &L, int i, int j, int k, NodeRows &nr) {
    bool any = false;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int cj = j - 1 + (r & 1), ck = k - 1 + (r >> 1);
        const bool row = (cj >= 0) & (cj < g.nex[1]) & (ck >= 0) & (ck < g.nex[2]);
        const int base = row ? (cj + ck * g.nex[1]) * g.nex[0] : 0;
        const bool in_l = row & (i >= 1);  // a node has i <= nex
        const bool in_r = row & (i < g.nex[0]);
        const int cl = in_l ? base + i - 1 : 0, cr = in_r ? base + i : 0;
        const uint32_t wl = L.occ[cl >> 5], wr = L.occ[cr >> 5], gr = L.gs[cr >> 5];
        // bitwise, not short-circuit: a branch would let the compiler sink a
        // word's load into it and wait for each row in turn
        const bool on_l = in_l & (bool)((wl >> (cl & 31)) & 1u);
        const bool on_r = in_r & (bool)((wr >> (cr & 31)) & 1u);
        const bool merged = on_l & on_r & !((gr >> (cr & 31)) & 1u);
        nr.cl[r] = cl;
        nr.cr[r] = cr;
        nr.rl[r] = on_l & !merged;
        nr.rr[r] = on_r;
        any = any | on_l | on_r;
    }
    return any;
}

// Sum over the rows around a node of NQ partial quantities, in ascending
// cell order, with Neumaier sums (16-byte loads of the stored runs).
template <int NQ>
__device__ __forceinline__ void node_partial_sums(const NodeRows &nr, const float *__restrict__ part, size_t ne,
                                                  float *s, float *cp) {
    constexpr int NQP = padded_nq<NQ>();
    float4 val[4][2][NQP / 4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        // the node is in corner group (1 - dj) + 2 (1 - dk) of the cells of row r
        const int g = (1 - (r & 1)) + 2 * (1 - (r >> 1));
        // a run that is not added reads the buffer's first run (one line for
        // the whole wave), so that every load is issued without a branch
        const size_t ol = nr.rl[r] ? ((size_t)(4 + g) * ne + nr.cl[r]) * NQP : 0;
        const size_t orr = nr.rr[r] ? ((size_t)g * ne + nr.cr[r]) * NQP : 0;
        const float4 *pl = (const float4 *)(part + ol);
        const float4 *pr = (const float4 *)(part + orr);
#pragma unroll
        for (int h = 0; h < NQP / 4; h++) {
            val[r][0][h] = pl[h];
            val[r][1][h] = pr[h];
        }
    }
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int side = 0; side < 2; side++) {
            if (!(side ? nr.rr[r] : nr.rl[r])) continue;
#pragma unroll
            for (int f = 0; f < NQ; f++) {
                const float4 q = val[r][side][f >> 2];
                neu_add(s[f], cp[f], (f & 3) == 0 ? q.x : (f & 3) == 1 ? q.y : (f & 3) == 2 ? q.z : q.w);
            }
        }
}

template <bool DIAG>
__global__ void __launch_bounds__(256) k32_p2g_nodes(GridGeom g, CellLists L, TileGeom tg,
                                                      const float *__restrict__ part, NodesF32 N, BcDense bc,
                                                      StepConsts sc) {
    int n, i, j, k;
    if (!tile_node(g, L, tg, DIAG, n, i, j, k)) return;
    const size_t ne = (size_t)g.nelem;
    const int nodec = g.nodec;
    constexpr int NQ = DIAG ? 8 : 7;
    float s[NQ], cp[NQ];
#pragma unroll
    for (int f = 0; f < NQ; f++) {
        s[f] = 0.f;
        cp[f] = 0.f;
    }
    // Straight-line up to the stores, so that no load waits behind the sums:
    // the boundary mask is read with the occupancy words and the prescribed
    // velocities with the runs (see node_bc); a node next to no particle sums
    // nothing (its runs read the buffer's first one and are skipped).
    const uint32_t mk = bc.mask ? bc.mask[n] : 0u;
    NodeRows nr;
    const bool any = node_rows4(g, L, i, j, k, nr);
    double bv[3];
    node_bc(bc, n, nodec, mk, bv);
    node_partial_sums<NQ>(nr, part, ne, s, cp);
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
#pragma unroll
    for (int c = 0; c < 3; c++)
        if (mk & (1u << c)) {
            v[c] = (float)(bv[c] * sc.facl);
            a[c] = 0.f;
        }
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
// MDT: per-material time steps (each particle takes its material's dt); the
// PIC node velocity v + dt a is then formed per particle from v and a.
template <int LPC, bool FLIP, bool MDT>
__global__ void __launch_bounds__(kSeg) k32_musl_cells(GridGeom g, PartF32 P, CellLists L, NodeTables T, NodesF32 N,
                                                        StepConsts sc, float *__restrict__ part, MaterialTable mt) {
    constexpr int G = kSeg / LPC;
    // staged fields: 0-2 r, 3 m, [4-6 old v for FLIP], [matid for MDT]; results: G2P velocity per particle
    constexpr int F_MAT = 4 + (FLIP ? 3 : 0);
    constexpr int NF = F_MAT + (MDT ? 1 : 0);
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
        // (MDT PIC: v here, a in nal)
        constexpr bool TWO = MDT && !FLIP;
        float nvl[8][3], nal[TWO ? 8 : 1][3];
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
                            nvl[q][c] = FLIP ? N.na[c * nodec + nid] : (MDT ? N.nv : N.nvp)[c * nodec + nid];
                            if constexpr (TWO) nal[q][c] = N.na[c * nodec + nid];
                        }
                    }
        } else {
#pragma unroll
            for (int q = 0; q < 8; q++)
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    nvl[q][c] = 0.f;
                    if constexpr (TWO) nal[q][c] = 0.f;
                }
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
                    if (MDT && P.matid) lds_gather<SS>(P.matid, sl, stage, F_MAT, r);
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
                float vp[3] = {0.f, 0.f, 0.f}, ap[3] = {0.f, 0.f, 0.f};
#pragma unroll
                for (int k = 0; k < 2; k++)
#pragma unroll
                    for (int j = 0; j < 2; j++)
#pragma unroll
                        for (int ii = 0; ii < 2; ii++) {
                            const int q = ii + 2 * j + 4 * k;
                            Nw[q] = wx[ii] * (wy[j] * wz[k]);
#pragma unroll
                            for (int c = 0; c < 3; c++) {
                                vp[c] += Nw[q] * nvl[q][c];
                                if constexpr (TWO) ap[c] += Nw[q] * nal[TWO ? q : 0][c];
                            }
                        }
                float dtp = sc.dt;
                // (an engine uploaded without material ids has material 0 only)
                if constexpr (MDT) dtp = mt.m[P.matid ? __float_as_int(stage[st_idx<SS>(F_MAT, i)]) : 0].dt;
                if (FLIP) {
                    vp[0] = stage[st_idx<SS>(4, i)] + dtp * vp[0];
                    vp[1] = stage[st_idx<SS>(5, i)] + dtp * vp[1];
                    vp[2] = stage[st_idx<SS>(6, i)] + dtp * vp[2];
                } else if (TWO) {
#pragma unroll
                    for (int c = 0; c < 3; c++) vp[c] += dtp * ap[c];
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
        // corner (i, j, k) = x-side i of corner group j + 2 k; momentum component c
        float v[24];
#pragma unroll
        for (int q = 0; q < 8; q++)
#pragma unroll
            for (int c = 0; c < 3; c++) v[part_entry<LPC, 3>(q >> 1, q & 1, c)] = acc[q][c];
        ReduceScatter<1, LPC, 24>::run(v, sub);
        xmerge_store<LPC, 3>(v, sub, si, L, g.nex[0], part, (size_t)g.nelem);
    }  // groups
}

template <bool DIAG>
__global__ void __launch_bounds__(256) k32_musl_nodes(GridGeom g, CellLists L, TileGeom tg,
                                                       const float *__restrict__ part, NodesF32 N, BcDense bc,
                                                       StepConsts sc) {
    int n, i, j, k;
    if (!tile_node(g, L, tg, DIAG, n, i, j, k)) return;
    const size_t ne = (size_t)g.nelem;
    const int nodec = g.nodec;
    float s[3] = {0.f, 0.f, 0.f}, cp[3] = {0.f, 0.f, 0.f};
    // straight-line up to the stores, as in k32_p2g_nodes: the node mass and
    // the boundary mask travel with the occupancy words
    const float m = N.nm[n];
    const uint32_t mk = bc.mask ? bc.mask[n] : 0u;
    NodeRows nr;
    const bool any = node_rows4(g, L, i, j, k, nr);
    double bv[3];
    node_bc(bc, n, nodec, mk, bv);
    node_partial_sums<3>(nr, part, ne, s, cp);
    float v[3] = {0.f, 0.f, 0.f};
    float p[3] = {s[0] + cp[0], s[1] + cp[1], s[2] + cp[2]};
    if ((double)m > sc.mtol) {
#pragma unroll
        for (int c = 0; c < 3; c++) v[c] = safe_div(p[c], m, sc.m_small);
    }
#pragma unroll
    for (int c = 0; c < 3; c++)
        if (mk & (1u << c)) v[c] = (float)(bv[c] * sc.facl);
    if (!any) {
        if (DIAG)
            for (int c = 0; c < 3; c++) N.nvm[c * nodec + n] = 0.f;
        return;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) N.nvm[c * nodec + n] = v[c];
    if (DIAG) {
#pragma unroll
        for (int c = 0; c < 3; c++) N.np[c * nodec + n] = p[c];
    }
}


// Mode 3 of the update kernel: the network weights in MFMA-fragment order,
// the feature scaling, and one 64-row exchange tile per wave.
struct SurFusedSmem {
    MlpInferSmem sw;
    float scl[64];
    __attribute__((aligned(16))) uint16_t xs[4][kMlpWaveScratchBytes / 2];
};

// SUR: 0 plain; 1 plain with sample logging; 2 queues the network's
// candidates for k32_surrogate_pass; 3 runs the network here, on the rows of
// each wave's own lanes (no queue), with the same arithmetic as mode 2.
template <int SUR>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k32_g2p_update(GridGeom g, PartF32 P, NodesF32 N, NodeTables T,
                                                       MaterialTable mt, StepConsts sc, uint32_t *err,
                                                       SurrogateArgs sa, BinFuse bf) {
    const int slot0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (bf.cnt) {
        for (int i = slot0; i < bf.ntiles; i += gridDim.x * blockDim.x) bf.tiles[i] = 0u;
        if (slot0 == 0) {
            bf.ctr->n_active = 0u;
        }
    }
    __shared__ std::conditional_t<SUR == 3, SurFusedSmem, int> fsm;
    if constexpr (SUR == 3) {
        // staged while the particle streams load; first read after the trial state
        mlp_load_infer_smem(fsm.sw, sa.net);
        if (threadIdx.x < 64) fsm.scl[threadIdx.x] = sa.net_scaling[threadIdx.x];
    }
    // Wave-wide steps (the network, the run-aggregated cell counts) need every
    // lane: no lane leaves early; invalid lanes compute on slot 0 / cell 0
    // and store nothing.
    bool valid = slot0 < P.n;
    const int slot = valid ? slot0 : 0;
    const size_t ld = P.ld;
    // The particle's own streams do not depend on its cell key: they are
    // issued with the key, so their latency overlaps the key -> node gather chain.
    const int mat = P.matid ? P.matid[slot] : 0;
    uint32_t key = P.key[slot];
    float r[3], H[9];
#pragma unroll
    for (int a = 0; a < 3; a++) r[a] = P.r[a * ld + slot];
#pragma unroll
    for (int q = 0; q < 9; q++) H[q] = __builtin_nontemporal_load(&P.H[q * ld + slot]);  // read once
    const float V0p = P.V0[slot];
    if (key >= (uint32_t)g.nelem) valid = false;
    if (!valid) key = 0;
    const int nodec = g.nodec;
    const int nexy = g.nex[0] * g.nex[1];
    int rem, ix, iy;
    const int iz = fdivmod((int)key, nexy, T.inv_nexy, rem);
    iy = fdivmod(rem, g.nex[0], T.inv_nex, ix);
    const float fx = r[0], fy = r[1], fz = r[2];
    const float wx[2] = {1.f - fx, fx}, wy[2] = {1.f - fy, fy}, wz[2] = {1.f - fz, fz};
    const float inv0 = T.inv_dx[0], inv1 = T.inv_dx[1], inv2 = T.inv_dx[2];
    const int n0 = ix + g.nxc[0] * iy + g.nxc[0] * g.nxc[1] * iz;
    const MaterialDev md = mt.m[mat];  // copied before the node gather: its loads overlap the gather's
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
    const float dtp = sc.mat_dt ? md.dt : sc.dt;
    float D[9];
#pragma unroll
    for (int q = 0; q < 9; q++) D[q] = dtp * Dv[q];
    // H_new = (I + D)(I + H) - I = H + D + D H
    float Hn[9];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++)
            Hn[3 * a + b] = (H[3 * a + b] + D[3 * a + b]) +
                            (D[3 * a] * H[b] + D[3 * a + 1] * H[3 + b] + D[3 * a + 2] * H[6 + b]);
    const double J = det_I_plus_H(Hn);
    const double V = J * (double)V0p;
    // position: offset in cell units, carried into the cell index when it leaves [0, 1)
    int ci[3] = {ix, iy, iz};
    bool inside = true;
    const double b0 = sc.sh_geup ? *sc.sh_b0 : 0.0;  // particle shifting: x += -b0 geup
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float ra = r[a] + (dtp * dxp[a]) * T.inv_dx[a];
        if (sc.sh_geup) ra += (float)(-b0 * sc.sh_geup[a * ld + slot] * (double)T.inv_dx[a]);
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
    // only the rate forms (linear elastic increment, J2) read the old stress
    const bool rate_form = (md.model == 0 || md.model == 10);
    float s[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (rate_form) {
#pragma unroll
        for (int q = 0; q < 6; q++) s[q] = P.sig[q * ld + slot];
    }
    float ep = (md.model == 10 && P.eps_p) ? P.eps_p[slot] : 0.f;
    bool defer = false;  // the stress (and plastic strain) of this slot is written by the network pass
    if constexpr (SUR == 0) {
        stress32(md, Hn, D, J, s, &ep, 1.f / dtp);
    } else {
        const bool j2 = md.model == 10;
        const bool tait = md.model == 20;
        J2Trial t;
        // the network's candidates: J2 plastic trial states, or Tait particles
        bool plastic = false;
        float pex = 0.f, Bt = 0.f;  // Tait: exact pressure and its scale B
        if (j2) {
            j2_trial(md.prop, D, s, ep, t);
            plastic = valid && t.f > 0.f && sa.kind == kSurJ2;
        } else {
            stress32(md, Hn, D, J, s, &ep, 1.f / dtp);
            if (tait && sa.kind == kSurTait) {
                plastic = valid;
                pex = tait_pressure(md.prop, J, &Bt);
            }
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
                        float x[16], y[16];
                        if (sa.kind == kSurTait) {
                            tait_features(J, D, dtp, x);
                            y[0] = pex;
#pragma unroll
                            for (int c = 1; c < 16; c++) y[c] = 0.f;
                        } else {
                            j2_features(t, D, s_old, x);
#pragma unroll
                            for (int c = 0; c < 6; c++) y[c] = t.st[c] - s[c];
                            y[6] = ep - t.ep;
#pragma unroll
                            for (int c = 7; c < 16; c++) y[c] = 0.f;
                        }
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
        } else if constexpr (SUR == 3) {
            // the wave's candidates run through the network here, one 64-row
            // tile per wave (rows of non-candidate lanes are computed and
            // ignored); each output depends only on its own row, so an
            // accepted prediction equals the one the network pass of mode 2
            // makes for the same state
            __syncthreads();  // weights staged (every lane of the block gets here)
            const unsigned long long m = __ballot(plastic);
            bool accept = false;
            if (m) {
                float x[16];
                if (sa.kind == kSurTait)
                    tait_features(J, D, dtp, x);
                else
                    j2_features(t, D, s, x);
                float xn[16];
#pragma unroll
                for (int f = 0; f < 16; f++) xn[f] = (plastic && f < kMlpInUsed) ? (x[f] - fsm.scl[f]) * fsm.scl[16 + f] : 0.f;
                float y[kMlpOutUsed];
                mlp_wave_infer(fsm.sw, xn, y, fsm.xs[threadIdx.x >> 6], lane);
                if (plastic) {
                    if (sa.kind == kSurTait) {
                        // the exact pressure checks the network's; an accepted one replaces it
                        const float pnet = y[0] * fsm.scl[48] + fsm.scl[32];
                        if (fabsf(pnet - pex) < sa.tol * (fabsf(pex) + 1e-6f * Bt)) {
                            accept = true;
                            const float d = pex - pnet;
#pragma unroll
                            for (int cc = 0; cc < 3; cc++) s[cc] += d;
                        }
                    } else {
                        float ss[6];
#pragma unroll
                        for (int cc = 0; cc < 6; cc++) ss[cc] = t.st[cc] - (y[cc] * fsm.scl[48 + cc] + fsm.scl[32 + cc]);
                        const float dg = y[6] * fsm.scl[54] + fsm.scl[38];
                        float epn = t.ep + dg;
                        accept = j2_residual(t, ss, dg) <= sa.tol;
                        if (!accept) j2_return(t, ss, &epn);  // fallback: the exact radial return
#pragma unroll
                        for (int cc = 0; cc < 6; cc++) s[cc] = ss[cc];
                        ep = epn;
                    }
                }
                const uint32_t ev = (uint32_t)__popcll(m);
                const uint32_t rj = (uint32_t)__popcll(__ballot(plastic && !accept));
                if (lane == 0) {
                    atomicAdd(&sa.rstat[2 * blockIdx.x], ev);
                    if (rj) atomicAdd(&sa.rstat[2 * blockIdx.x + 1], rj);
                }
            }
            if (j2 && !plastic) j2_return(t, s, &ep);  // elastic trial state
        } else {
            // the candidates of this block are compacted, in slot order, into
            // the block's own queue region for the network pass that follows
            // this kernel. J2: the plastic corrector is left to that pass (the
            // network where its residual is within the tolerance, the exact
            // return otherwise), so the stress of a queued state is not
            // stored here. Tait: the exact stress is stored and its pressure
            // checks the network's.
            float x[16];
            if (plastic) {
                if (sa.kind == kSurTait)
                    tait_features(J, D, dtp, x);
                else
                    j2_features(t, D, s, x);
            }
            if (j2 && !plastic) j2_return(t, s, &ep);  // elastic trial state (or a J2 state the network skips)
            defer = j2 && plastic;
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
                    if (sa.kind == kSurTait) {
                        float4 *row = (float4 *)(sa.queue + (size_t)qi * kSurRowTait);
                        row[0] = make_float4(__int_as_float(slot), x[0], x[1], pex);
                        row[1] = make_float4(Bt, 0.f, 0.f, 0.f);
                    } else {
                        float4 *row = (float4 *)(sa.queue + (size_t)qi * kSurRowJ2);
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
    if (md.model == 10 && P.eps_p && !defer) P.eps_p[slot] = ep;
    // streamed out with non-temporal stores: the particle state is next read
    // a whole step later, and stores that stay dirty in L2 slow the next
    // kernels (P2G cells 77 -> 66 us on the 4M-particle cube)
#pragma unroll
    for (int q = 0; q < 9; q++) __builtin_nontemporal_store(Hn[q], &P.H[q * ld + slot]);
    if (!defer) {
#pragma unroll
        for (int q = 0; q < 6; q++) __builtin_nontemporal_store(s[q], &P.sig[q * ld + slot]);
    }
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
template <int RF>
__device__ __forceinline__ void sur_load_row(const float *queue, uint32_t qi, bool live, float r[RF]) {
    if (live) {
        const float4 *row = (const float4 *)(queue + (size_t)qi * RF);
#pragma unroll
        for (int c = 0; c < RF / 4; c++) {
            const float4 v = row[c];
            r[4 * c] = v.x;
            r[4 * c + 1] = v.y;
            r[4 * c + 2] = v.z;
            r[4 * c + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int c = 0; c < RF; c++) r[c] = 0.f;
    }
}

// Network pass over the states queued by the update kernel of a step. A tile
// slot is (region, k): rows [64 k, 64 k + 64) of queue region k / 4 (the
// queued states of one update block). The blocks are persistent (the
// network weights are staged in LDS once per block); each wave visits every
// nw-th tile slot, finds the non-empty ones of its next 64 slots with one
// load of the region counts per lane, and loads the rows of its next tile
// while the current tile runs through the network on the BF16 MFMA.
// J2: a prediction is accepted where the return-mapping residual is within
// the tolerance; otherwise the exact radial return of the queued trial state
// is computed here (the update kernel left these stresses to this pass).
// Tait: the network's pressure replaces the exact one where they agree
// within the tolerance; otherwise the exact stress the update kernel wrote
// stays. The evaluated / rejected tallies go to per-region counters: a
// region's tiles may be spread over several waves, so they are added
// atomically, but to distinct addresses.
template <int KIND>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k32_surrogate_pass(PartF32 P, SurrogateArgs sa) {
    constexpr int RF = KIND == kSurTait ? kSurRowTait : kSurRowJ2;
    constexpr int NIN = KIND == kSurTait ? 2 : kMlpInUsed;  // features in a row
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
        float r[RF];
        sur_load_row<RF>(sa.queue, q_cur + (uint32_t)lane, lane < n_cur, r);
        for (;;) {
            // next tile of this batch: its rows are loaded while this one runs
            const bool more = todo != 0ull;
            int n_nxt = 0;
            uint32_t q_nxt = 0;
            float rn[RF];
            if (more) {
                const int bn = __ffsll(todo) - 1;
                todo &= todo - 1ull;
                n_nxt = __shfl(myn, bn, 64);
                const uint32_t sn = base + (uint32_t)bn * nw;
                q_nxt = (sn >> 2) * (uint32_t)kSurRegion + 64u * (sn & 3u);
            }
            sur_load_row<RF>(sa.queue, q_nxt + (uint32_t)lane, more && lane < n_nxt, rn);
            const bool live = lane < n_cur;
            float xn[16];
#pragma unroll
            for (int f = 0; f < 16; f++)  // a row holds NIN features; the rest are zero
                xn[f] = (f < kMlpInUsed) ? ((f < NIN ? r[1 + f] : 0.f) - scl[f]) * scl[16 + f] : 0.f;
            float y[kMlpOutUsed];
            mlp_wave_infer(sw, xn, y, xs[w], lane);
            bool accept = false;
            if (live) {
                const int slot = __float_as_int(r[0]);
                const size_t ld = P.ld;
                if constexpr (KIND == kSurTait) {
                    // pressure closure: the exact Tait pressure checks the
                    // network's; an accepted one replaces it in the stored stress
                    const float pex = r[3], B = r[4];
                    const float pnet = y[0] * scl[48] + scl[32];
                    if (fabsf(pnet - pex) < sa.tol * (fabsf(pex) + 1e-6f * B)) {
                        accept = true;
                        const float d = pex - pnet;
#pragma unroll
                        for (int cc = 0; cc < 3; cc++) P.sig[cc * ld + slot] += d;
                    }
                } else {
                    // trial state from the queued trial stress (the update kernel's arithmetic)
                    J2Trial t;
                    t.E = r[20];
                    t.nu = r[21];
                    t.sy = r[22];
                    t.Hh = r[23];
                    t.ep = r[13];
#pragma unroll
                    for (int cc = 0; cc < 6; cc++) t.st[cc] = r[14 + cc];
                    j2_trial_derived(t);
                    float ss[6];
#pragma unroll
                    for (int cc = 0; cc < 6; cc++) ss[cc] = t.st[cc] - (y[cc] * scl[48 + cc] + scl[32 + cc]);
                    const float dg = y[6] * scl[54] + scl[38];
                    float epn = t.ep + dg;
                    accept = j2_residual(t, ss, dg) <= sa.tol;
                    if (!accept) j2_return(t, ss, &epn);  // fallback: the exact radial return
#pragma unroll
                    for (int cc = 0; cc < 6; cc++) P.sig[cc * ld + slot] = ss[cc];
                    if (P.eps_p) P.eps_p[slot] = epn;
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
            for (int cc = 0; cc < RF; cc++) r[cc] = rn[cc];
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
    for (int q = 0; q < 9; q++) {
        if (src.C) dst.C[q * lt + s] = src.C[q * ls + o];
        if (src.Dinv) dst.Dinv[q * lt + s] = src.Dinv[q * ls + o];
        if (src.Hb) dst.Hb[q * lt + s] = src.Hb[q * ls + o];
    }
    if (src.jb) dst.jb[s] = src.jb[o];
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

template <int LPC, bool MDT>
static void musl_cells_launch_mdt(hipStream_t s, int ng, const GridGeom &g, const PartF32 &P, const CellLists &L,
                                  const NodeTables &T, const NodesF32 &N, const StepConsts &c, float *partials,
                                  const MaterialTable &mt) {
    if (c.scheme == 1)
        hipLaunchKernelGGL((k32_musl_cells<LPC, true, MDT>), dim3(ng), dim3(kSeg), 0, s, g, P, L, T, N, c, partials,
                           mt);
    else
        hipLaunchKernelGGL((k32_musl_cells<LPC, false, MDT>), dim3(ng), dim3(kSeg), 0, s, g, P, L, T, N, c, partials,
                           mt);
}

template <int LPC>
static void musl_cells_launch(hipStream_t s, int ng, const GridGeom &g, const PartF32 &P, const CellLists &L,
                              const NodeTables &T, const NodesF32 &N, const StepConsts &c, float *partials,
                              const MaterialTable *mt) {
    if (c.mat_dt && mt)
        musl_cells_launch_mdt<LPC, true>(s, ng, g, P, L, T, N, c, partials, *mt);
    else
        musl_cells_launch_mdt<LPC, false>(s, ng, g, P, L, T, N, c, partials, MaterialTable{nullptr, 0});
}

void launch_f32_g2p_vel_musl(hipStream_t s, const GridGeom &g, const PartF32 &P, const NodesF32 &N,
                             const CellLists &L, const BcDense &bc, const NodeTables &T, const TileGeom &tg,
                             const StepConsts &c, float *partials, int diag, hipEvent_t mid, const MaterialTable *mt) {
    const int ng = cell_groups_grid(g);
    switch (c.lanes_per_cell) {
    case 8: musl_cells_launch<8>(s, ng, g, P, L, T, N, c, partials, mt); break;
    case 4: musl_cells_launch<4>(s, ng, g, P, L, T, N, c, partials, mt); break;
    case 2: musl_cells_launch<2>(s, ng, g, P, L, T, N, c, partials, mt); break;
    default: musl_cells_launch<1>(s, ng, g, P, L, T, N, c, partials, mt); break;
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
    static int resident[2] = {0, 0};
    const bool tait = sa.kind == kSurTait;
    if (resident[tait] == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (tait)
            (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k32_surrogate_pass<kSurTait>, 256, 0);
        else
            (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k32_surrogate_pass<kSurJ2>, 256, 0);
        resident[tait] = std::max(1, cus) * std::max(1, per_cu);
    }
    const int blocks = std::min((int)sa.nregions, resident[tait]);
    if (tait)
        hipLaunchKernelGGL(k32_surrogate_pass<kSurTait>, dim3(blocks), dim3(256), 0, s, P, sa);
    else
        hipLaunchKernelGGL(k32_surrogate_pass<kSurJ2>, dim3(blocks), dim3(256), 0, s, P, sa);
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
    else if (sa.mode == 3)
        hipLaunchKernelGGL(k32_g2p_update<3>, grid, block, 0, s, g, P, N, T, mt, c, err, sa, bf);
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

// FP32 explicit-MPM step kernels for stencils of 2 or 3 control points per
// axis: the open quadratic B-spline (idimc = 2, module/shape_function.cpp:
// 53-104) on any axis, the tent functions (idimc = 1) on the others, and the
// affine transfers TPIC and APIC (module/map_and_interpolate.h:57-176) with
// either. The linear PIC / FLIP step keeps its tuned kernels (fast32.hip).
//
// Same structure and determinism as fast32.hip: a one-wave workgroup owns 64
// consecutive non-empty cells (one lane per cell), stages their particles
// (one contiguous range of the cell-sorted order) through LDS with the LDS
// DMA, and folds each cell's particles in sorted order into per-cell corner
// partials, NX NY NZ corners per cell. The node passes add, for each control
// point, the partials of the cells whose stencil holds it, in ascending cell
// order with Neumaier-compensated FP32 sums, and apply the grid update. The
// update kernel gathers the NX NY NZ node velocities of each particle in slot
// order. No float atomics anywhere; runs are bit-reproducible.
//
// Positions are cell index + FP32 offset r in [0, 1) in cell units, so the
// weights are evaluated from r without cancellation: for the quadratic
// B-spline the three interior pieces are 0.5 (1 - r)^2, 0.75 - (r - 1/2)^2 and
// 0.5 r^2, and the clamped pieces of the first two and last two control
// points are the reference's table rewritten in r.
#include "fast32_device.h"

#include <type_traits>

#include <algorithm>

namespace mpm {

namespace {

// Weights of the control points c + o (o < W) of a particle at offset r of
// cell c on an axis of nex cells: w (N), d (dN/dr: times 1/dx gives dN/dx),
// x (control-point coordinate minus the particle's, in cell units; the
// control points are DefineCpPos's: the nodes for the tent functions, the
// Greville abscissae of the open B-spline -- domain ends for the first and
// last, cell centres in between, module/mesh.h:32-47).
template <int W>
struct AxisW {
    float w[W], d[W], x[W];
};

template <int W>
__device__ __forceinline__ void axis_weights(int c, float r, int nex, AxisW<W> &A) {
    if constexpr (W == 2) {
        A.w[0] = 1.f - r;
        A.w[1] = r;
        A.d[0] = -1.f;  // at r == 0 the left node takes the right-hand slope, as shape_function.cpp:45-50
        A.d[1] = 1.f;
        A.x[0] = -r;
        A.x[1] = 1.f - r;
    } else {
        const float s = 1.f - r, h = r - 0.5f;
        A.w[0] = 0.5f * s * s;
        A.d[0] = -s;
        A.x[0] = -0.5f - r;
        A.w[1] = 0.75f - h * h;
        A.d[1] = -2.f * h;
        A.x[1] = 0.5f - r;
        A.w[2] = 0.5f * r * r;
        A.d[2] = r;
        A.x[2] = 1.5f - r;
        // clamped pieces, with the reference's precedence: global index 0,
        // then 1, then the last, then the one before it (selects, not
        // branches: the cell, and so the piece, is fixed per lane). Cells 2 ..
        // nex - 3 have none: a lane's cell is fixed, so interior waves skip
        // the selects.
        const int last = nex + 1;
        if (c >= 2 && c <= nex - 3) return;
#pragma unroll
        for (int o = 0; o < 3; o++) {
            const int gi = c + o;
            const bool e0 = gi == 0;
            const bool e1 = !e0 && gi == 1;
            const bool eL = !e0 && !e1 && gi == last;
            const bool eL1 = !e0 && !e1 && !eL && gi == last - 1;
            float w = A.w[o], d = A.d[o], x = A.x[o];
            if (o == 1) {
                w = e1 ? r * (2.f - 1.5f * r) : w;
                d = e1 ? 2.f - 3.f * r : d;
                w = eL1 ? 0.5f + r * (1.f - 1.5f * r) : w;
                d = eL1 ? 1.f - 3.f * r : d;
            }
            w = e0 ? s * s : w;
            d = e0 ? -2.f * s : d;
            x = e0 ? -r : x;
            w = eL ? r * r : w;
            d = eL ? 2.f * r : d;
            x = eL ? (float)(nex - c) - r : x;
            A.w[o] = w;
            A.d[o] = d;
            A.x[o] = x;
        }
    }
}

// One control point o (0, 1 or 2, the same for a lane's whole particle loop)
// of the quadratic B-spline, with axis_weights<3>'s expressions: the weight,
// its r-derivative and the offset, by selects only (an indexed register array
// would be placed in scratch memory).
__device__ __forceinline__ void axis_weight_at(int c, float r, int nex, int o, float &w, float &d, float &x) {
    const float s = 1.f - r, h = r - 0.5f;
    const float w0 = 0.5f * s * s, w1 = 0.75f - h * h, w2 = 0.5f * r * r;
    const float d0 = -s, d1 = -2.f * h, d2 = r;
    const float x0 = -0.5f - r, x1 = 0.5f - r, x2 = 1.5f - r;
    w = o == 0 ? w0 : (o == 1 ? w1 : w2);
    d = o == 0 ? d0 : (o == 1 ? d1 : d2);
    x = o == 0 ? x0 : (o == 1 ? x1 : x2);
    if (c >= 2 && c <= nex - 3) return;  // no clamped piece in this cell
    const int last = nex + 1;
    const int gi = c + o;
    const bool e0 = gi == 0;
    const bool e1 = !e0 && gi == 1;
    const bool eL = !e0 && !e1 && gi == last;
    const bool eL1 = !e0 && !e1 && !eL && gi == last - 1;
    w = (o == 1 && e1) ? r * (2.f - 1.5f * r) : w;
    d = (o == 1 && e1) ? 2.f - 3.f * r : d;
    w = (o == 1 && eL1) ? 0.5f + r * (1.f - 1.5f * r) : w;
    d = (o == 1 && eL1) ? 1.f - 3.f * r : d;
    w = e0 ? s * s : w;
    d = e0 ? -2.f * s : d;
    x = e0 ? -r : x;
    w = eL ? r * r : w;
    d = eL ? 2.f * r : d;
    x = eL ? (float)(nex - c) - r : x;
}

// 3x3 inverse through FP64 (the APIC D matrix is about dx^2 / 4 I).
__device__ __forceinline__ void inv3_f(const float a[9], float o[9]) {
    double A[9];
#pragma unroll
    for (int q = 0; q < 9; q++) A[q] = (double)a[q];
    const double c0 = A[4] * A[8] - A[5] * A[7], c1 = A[5] * A[6] - A[3] * A[8], c2 = A[3] * A[7] - A[4] * A[6];
    const double id = 1.0 / (A[0] * c0 + A[1] * c1 + A[2] * c2);
    o[0] = (float)(c0 * id);
    o[1] = (float)((A[2] * A[7] - A[1] * A[8]) * id);
    o[2] = (float)((A[1] * A[5] - A[2] * A[4]) * id);
    o[3] = (float)(c1 * id);
    o[4] = (float)((A[0] * A[8] - A[2] * A[6]) * id);
    o[5] = (float)((A[2] * A[3] - A[0] * A[5]) * id);
    o[6] = (float)(c2 * id);
    o[7] = (float)((A[1] * A[6] - A[0] * A[7]) * id);
    o[8] = (float)((A[0] * A[4] - A[1] * A[3]) * id);
}

constexpr int kStageS = 65;  // stage block stride (64 + 1: conflict-free lane-per-cell reads)

// Affine transfer matrix of a particle in cell units of the offset: Mh[a][b]
// = M_ab dx_b, with M the TPIC velocity gradient or the APIC B D^-1, so that
// M (x_a - x_p) = sum_b Mh[a][b] x_b for the cell-unit offsets x_b.
template <int AFF, int NB = kBlk>
__device__ __forceinline__ void affine_matrix(const float *stage, int i, int fC, const float dx[3], float Mh[3][3]) {
    float M[9];
    if constexpr (AFF == 1) {
#pragma unroll
        for (int q = 0; q < 9; q++) M[q] = stage[st_idx<kStageS, NB>(fC + q, i)];
    } else {
        float B[9], Di[9];
#pragma unroll
        for (int q = 0; q < 9; q++) {
            B[q] = stage[st_idx<kStageS, NB>(fC + q, i)];
            Di[q] = stage[st_idx<kStageS, NB>(fC + 9 + q, i)];
        }
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) M[3 * a + b] = B[3 * a] * Di[b] + B[3 * a + 1] * Di[3 + b] + B[3 * a + 2] * Di[6 + b];
    }
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) Mh[a][b] = M[3 * a + b] * dx[b];
}

// ------------------------------------------------------------------ x-merge
// Cells that follow each other along x share NX - 1 of their NX corner
// columns. In a cells kernel the lanes of one z-layer hold the consecutive
// non-empty cells of a group (lane = layer * G + u). A cell is linked to its
// left neighbour when that neighbour is the previous cell of its group; its
// column values become S[NX - 1] = col[NX - 1] and S[ii] = col[ii] + S[ii + 1]
// of the left neighbour (ii < NX - 1), so S[0] holds everything the cells of
// its run up to it add to the control point of its column 0. A cell stores
// S[0], and S[ii >= 1] only when its right neighbour does not take them (it
// is empty, past the row's end, or starts the next group). A control point's
// row then reads S[0] of the cell of its column 0, S[d] of the cell d to the
// left where the cell right of that one did not take it (row_reads).
// The node passes learn the links from the occupancy bits and the binning's
// group-start bits of the same grouping; a group size of 1 links nothing.
struct XLinks {
    bool link, keep_right;
};
template <int G>
__device__ __forceinline__ XLinks x_links(const CellInfo &si, int u, int nex0, bool xm) {
    const int pc = __shfl_up(si.c, 1, 64);
    const int nc = __shfl_down(si.c, 1, 64);
    XLinks x;
    x.link = xm && u >= 1 && si.c > 0 && pc == si.c - 1 && si.ix > 0;
    x.keep_right = !(xm && u + 1 < G && nc == si.c + 1 && si.ix + 1 < nex0);
    return x;
}

// S[ii] += S[ii + 1] of the lane below (the left neighbour) where linked,
// for one row of NX corners of NQ quantities, the first NM of which are
// merged (called by all lanes)
template <int NX, int NQ, int NM = NQ>
__device__ __forceinline__ void x_merge_row(float (&S)[NX][NQ], bool link) {
#pragma unroll
    for (int ii = NX - 2; ii >= 0; ii--)
#pragma unroll
        for (int f = 0; f < NM; f++) {
            const float up = __shfl_up(S[ii + 1][f], 1, 64);
            if (link) S[ii][f] += up;
        }
}

// stores NQ quantities padded to NQP floats at slot q of cell c ([q][cell][NQP])
template <int NQ, int NQP>
__device__ __forceinline__ void store_run(float *__restrict__ part, size_t ne, int q, int c, const float (&v)[NQ]) {
    float4 *dst = (float4 *)(part + ((size_t)q * ne + (size_t)c) * NQP);
#pragma unroll
    for (int h = 0; h < NQP / 4; h++)
        dst[h] = make_float4(4 * h < NQ ? v[4 * h] : 0.f, 4 * h + 1 < NQ ? v[4 * h + 1] : 0.f,
                             4 * h + 2 < NQ ? v[4 * h + 2] : 0.f, 4 * h + 3 < NQ ? v[4 * h + 3] : 0.f);
}

// Which stored runs control point i of one cell row adds: rd[d] for the cell
// i - d (its column d), from the occupancy and group-start bits of the cells
// i - NX + 1 .. i of the row (base = the row's first cell). The words are
// loaded by row_words for every row of a control point before any is decoded
// (one round trip for all rows); a row outside the grid, or past the row's
// end, reads word 0 and decodes to no runs.
struct RowWords {
    uint32_t o0, o1, g0, g1;
    int clo, xlo, xhi;
    bool ok;
};
template <int NX>
__device__ __forceinline__ RowWords row_words(const uint32_t *occ, const uint32_t *gs, int nex, int base, int i,
                                              bool row_ok) {
    RowWords w;
    w.xlo = max(i - (NX - 1), 0);
    w.xhi = min(i, nex - 1);
    w.ok = row_ok && w.xlo <= w.xhi;
    w.clo = w.ok ? base + w.xlo : 0;
    const int chi = w.ok ? base + w.xhi : 0;
    w.o0 = occ[w.clo >> 5];
    w.o1 = occ[chi >> 5];
    w.g0 = gs[w.clo >> 5];
    w.g1 = gs[chi >> 5];
    return w;
}
template <int NX>
__device__ __forceinline__ bool row_reads(const RowWords &w, int base, int i, bool (&rd)[NX]) {
    auto bit = [&](uint32_t w0, uint32_t w1, int c) -> bool {
        return (((c >> 5) == (w.clo >> 5) ? w0 : w1) >> (c & 31)) & 1u;
    };
    bool any = false;
#pragma unroll
    for (int d = 0; d < NX; d++) {
        const int cx = i - d;
        const bool on = w.ok && cx >= w.xlo && cx <= w.xhi && bit(w.o0, w.o1, base + cx);
        any = any || on;
        // taken by the right neighbour: it holds particles and does not start a group
        const bool taken =
            d > 0 && cx + 1 <= w.xhi && bit(w.o0, w.o1, base + cx + 1) && !bit(w.g0, w.g1, base + cx + 1);
        rd[d] = on && !taken;
    }
    return any;
}

// ------------------------------------------------------------------ P2G
// AFF: 0 none, 1 TPIC, 2 APIC. Partials: one 32-byte run {m, px, py, pz,
// fx, fy, fz, volume} per (corner q, cell c) at (q * nelem + c) * 8, corner
// q = i + NX (j + NY k) in stencil order k, j, i as MakSf; columns 1 and 2
// of a row only where the x-merge below leaves them to the cell.
// LZ lanes per cell (1 or NZ): with LZ = NZ each lane folds the cell's
// particles into one z-layer of the corners (lane = layer * G + cell of the
// group, G = 64 / NZ cells per group), which cuts the accumulators of the
// 27-point stencil from 216 to 72 registers; every corner's sum is the same
// sequence of additions either way.
// LEAN: no traction and no diagnostic output (the node pass then adds 7
// quantities): the corner volumes and the traction term are left out of the
// particle loop and the traction fields out of the stage; the volume slot of
// the stored runs holds 0.
template <int NX, int NY, int NZ, int AFF, int LZ, bool LEAN>
__global__ void __launch_bounds__(kSeg) ks_p2g_cells(GridGeom g, PartF32 P, CellLists L, NodeTables T,
                                                      float *__restrict__ part) {
    static_assert(LZ == 1 || LZ == NZ, "a lane holds all z-layers of a cell or one");
    constexpr int KL = NZ / LZ;       // z-layers per lane
    constexpr int G = kSeg / LZ;      // cells per group
    constexpr int NN = NX * NY * NZ;
    constexpr int NL = NX * NY * KL;  // corners per lane
    // staged fields: 0-2 r, 3-5 v, 6 m, 7 V, 8-13 sigma, [14-16 traction], then C (9) and D^-1 (9)
    constexpr int F_C = LEAN ? 14 : 17;
    constexpr int NF = F_C + (AFF >= 1 ? 9 : 0) + (AFF == 2 ? 9 : 0);
    // chunk of PB 64-particle blocks: three for the layered stencils (a group
    // of 21 cells holds 168 particles at 8 per cell), which keeps the stage
    // small enough for three waves per SIMD
    constexpr int PB = (LZ > 1) ? 3 : kBlk;
    constexpr int PCH = 64 * PB;
    __shared__ float stage[NF * PB * kStageS];
    const uint32_t n_active = L.ctr->n_active;
    const int ngroups = (int)((n_active + G - 1) / G);
    const int t = threadIdx.x;
    const int k0 = (LZ == 1) ? 0 : t / G;  // first z-layer of this lane (LZ for an idle lane)
    const int u = (LZ == 1) ? t : t - k0 * G;
    const size_t ld = P.ld;
    const bool trac = !LEAN && P.trac != nullptr;
    const float inv0 = T.inv_dx[0], inv1 = T.inv_dx[1], inv2 = T.inv_dx[2];
    const float dxs[3] = {T.dx[0], T.dx[1], T.dx[2]};
    for (int grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
        CellInfo si;
        uint32_t S, E;
        wave_fence();
        {
            const uint32_t a0 = (uint32_t)grp * G;
            const uint32_t a = a0 + (uint32_t)u;
            const uint32_t a_end = min(a0 + (uint32_t)G, n_active);
            S = ((const __attribute__((address_space(4))) uint32_t *)L.astart)[a0];
            E = ((const __attribute__((address_space(4))) uint32_t *)L.astart)[a_end];
            si.c = -1;
            si.cs = si.ce = 0u;
            si.ix = si.iy = si.iz = 0;
            if (k0 < LZ && a < a_end) {
                si.c = (int)L.active[a];
                si.cs = L.astart[a];
                si.ce = L.astart[a + 1];
                const int nexy = g.nex[0] * g.nex[1];
                si.iz = si.c / nexy;
                const int r = si.c - si.iz * nexy;
                si.iy = r / g.nex[0];
                si.ix = r - si.iy * g.nex[0];
            }
        }
        float acc[NL][7];
        float vof[NL];
#pragma unroll
        for (int q = 0; q < NL; q++) {
#pragma unroll
            for (int f = 0; f < 7; f++) acc[q][f] = 0.f;
            vof[q] = 0.f;
        }
        for (uint32_t base = S; base < E; base += PCH) {
            const int nch = (int)min((uint32_t)PCH, E - base);
            uint32_t slot[PB];
            load_slots<PB>(L, base, nch, slot);
#pragma unroll
            for (int r = 0; r < PB; r++) {
                if (64 * r + t < nch) {
                    const uint32_t sl = slot[r];
#pragma unroll
                    for (int a = 0; a < 3; a++) {
                        lds_gather<kStageS, kNT, PB>(P.r + a * ld, sl, stage, a, r);
                        lds_gather<kStageS, kNT, PB>(P.v + a * ld, sl, stage, 3 + a, r);
                    }
                    lds_gather<kStageS, kNT, PB>(P.m, sl, stage, 6, r);
                    lds_gather<kStageS, kNT, PB>(P.V32, sl, stage, 7, r);
#pragma unroll
                    for (int q = 0; q < 6; q++) lds_gather<kStageS, kNT, PB>(P.sig + q * ld, sl, stage, 8 + q, r);
                    if (trac) {
#pragma unroll
                        for (int a = 0; a < 3; a++) lds_gather<kStageS, kNT, PB>(P.trac + a * ld, sl, stage, 14 + a, r);
                    }
                    if constexpr (AFF >= 1) {
#pragma unroll
                        for (int q = 0; q < 9; q++) lds_gather<kStageS, kNT, PB>(P.C + q * ld, sl, stage, F_C + q, r);
                    }
                    if constexpr (AFF == 2) {
#pragma unroll
                        for (int q = 0; q < 9; q++)
                            lds_gather<kStageS, kNT, PB>(P.Dinv + q * ld, sl, stage, F_C + 9 + q, r);
                    }
                }
            }
            wait_vm_all();
            wave_fence();
            const uint32_t lo = max(si.cs, base), hi = min(si.ce, base + (uint32_t)nch);
            for (uint32_t pos = lo; pos < hi; pos++) {
                const int i = (int)(pos - base);
                AxisW<NX> Ax;
                AxisW<NY> Ay;
                AxisW<KL> Az;  // this lane's z-layers
                axis_weights<NX>(si.ix, stage[st_idx<kStageS, PB>(0, i)], g.nex[0], Ax);
                axis_weights<NY>(si.iy, stage[st_idx<kStageS, PB>(1, i)], g.nex[1], Ay);
                if constexpr (LZ == 1) {
                    axis_weights<NZ>(si.iz, stage[st_idx<kStageS, PB>(2, i)], g.nex[2], Az);
                } else {
                    static_assert(NZ == 3 && KL == 1, "one quadratic z-layer per lane");
                    axis_weight_at(si.iz, stage[st_idx<kStageS, PB>(2, i)], g.nex[2], k0, Az.w[0], Az.d[0], Az.x[0]);
                }
                const float m = stage[st_idx<kStageS, PB>(6, i)];
                const float mv[3] = {m * stage[st_idx<kStageS, PB>(3, i)], m * stage[st_idx<kStageS, PB>(4, i)],
                                     m * stage[st_idx<kStageS, PB>(5, i)]};
                const float Vp = stage[st_idx<kStageS, PB>(7, i)];
                const float nV = -Vp;
                float sg[6];
#pragma unroll
                for (int q = 0; q < 6; q++) sg[q] = nV * stage[st_idx<kStageS, PB>(8 + q, i)];
                // -V sigma_ab / dx_b (Voigt xx,yy,zz,yz,xz,xy): f_a = sum_b (.)_ab dN/dr_b
                const float X[3] = {sg[0] * inv0, sg[5] * inv0, sg[4] * inv0};
                const float Y[3] = {sg[5] * inv1, sg[1] * inv1, sg[3] * inv1};
                const float Z[3] = {sg[4] * inv2, sg[3] * inv2, sg[2] * inv2};
                float tr[3] = {0.f, 0.f, 0.f};
                if (trac) {
#pragma unroll
                    for (int a = 0; a < 3; a++) tr[a] = stage[st_idx<kStageS, PB>(14 + a, i)];
                }
                // momentum per corner: m v + m M (x_a - x_p), split by axis
                float px[3][NX], py[3][NY], pz[3][KL];
                if constexpr (AFF >= 1) {
                    float Mh[3][3];
                    affine_matrix<AFF, PB>(stage, i, F_C, dxs, Mh);
#pragma unroll
                    for (int a = 0; a < 3; a++) {
#pragma unroll
                        for (int o = 0; o < NX; o++) px[a][o] = mv[a] + (m * Mh[a][0]) * Ax.x[o];
#pragma unroll
                        for (int o = 0; o < NY; o++) py[a][o] = (m * Mh[a][1]) * Ay.x[o];
#pragma unroll
                        for (int o = 0; o < KL; o++) pz[a][o] = (m * Mh[a][2]) * Az.x[o];
                    }
                }
                // Tensor-product form: with the y and z factors of a corner row
                // folded once per row, a corner costs one FMA per quantity for
                // mass, momentum and volume and two for each force component:
                // f_a = dN/dr_x (X_a wy wz) + N_x (Y_a dy wz + Z_a wy dz + t_a wy wz)
#pragma unroll
                for (int k = 0; k < KL; k++)
#pragma unroll
                    for (int j = 0; j < NY; j++) {
                        const float wyz = Ay.w[j] * Az.w[k];
                        const float dyz = Ay.d[j] * Az.w[k];
                        const float wydz = Ay.w[j] * Az.d[k];
                        float Fx[3], Fw[3], Mv[3];
#pragma unroll
                        for (int a = 0; a < 3; a++) {
                            Fx[a] = X[a] * wyz;
                            // explicit FMAs: the lean form is the full one with tr = 0
                            const float zw = Z[a] * wydz;
                            Fw[a] = __builtin_fmaf(Y[a], dyz, LEAN ? zw : __builtin_fmaf(tr[a], wyz, zw));
                            Mv[a] = mv[a] * wyz;
                        }
                        const float Mw = m * wyz, Vw = Vp * wyz;
#pragma unroll
                        for (int ii = 0; ii < NX; ii++) {
                            const int q = ii + NX * (j + NY * k);
                            const float wx = Ax.w[ii];
                            acc[q][0] += wx * Mw;
#pragma unroll
                            for (int a = 0; a < 3; a++) {
                                if constexpr (AFF >= 1)
                                    acc[q][1 + a] += (wx * wyz) * ((px[a][ii] + py[a][j]) + pz[a][k]);
                                else
                                    acc[q][1 + a] += wx * Mv[a];
                                acc[q][4 + a] += Ax.d[ii] * Fx[a];
                                acc[q][4 + a] += wx * Fw[a];
                            }
                            if constexpr (!LEAN) vof[q] += wx * Vw;
                        }
                    }
            }
            wave_fence();
        }
        // corner rows merged along x, then stored as 32-byte runs (m, p, f, volume)
        const XLinks xl = x_links<G>(si, u, g.nex[0], L.gsize > 1u);
        const bool on = si.ce > si.cs;
        // the cell count is re-read here (an empty asm makes it opaque) so
        // that the compiler does not keep the store addresses live in
        // registers across the whole kernel
        int nel = g.nelem;
        asm volatile("" : "+s"(nel));
        const size_t ne = (size_t)nel;
#pragma unroll
        for (int kl = 0; kl < KL; kl++)
#pragma unroll
            for (int j = 0; j < NY; j++) {
                float S[NX][8];
#pragma unroll
                for (int ii = 0; ii < NX; ii++) {
                    const int ql = ii + NX * (j + NY * kl);
#pragma unroll
                    for (int f = 0; f < 7; f++) S[ii][f] = acc[ql][f];
                    S[ii][7] = LEAN ? 0.f : vof[ql];
                }
                x_merge_row<NX, 8, LEAN ? 7 : 8>(S, xl.link);
                if (on) {
#pragma unroll
                    for (int ii = 0; ii < NX; ii++)
                        if (ii == 0 || xl.keep_right)
                            store_run<8, 8>(part, ne, ii + NX * (j + NY * (k0 + kl)), si.c, S[ii]);
                }
            }
    }
}

// Cells whose stencil holds control point (i, j, k), in ascending cell
// order: cell (i - NX + 1 + di, ...) for di < NX, the point being its corner
// (NX - 1 - di, ...). Visits f(cell, corner) for the cells inside the grid
// that hold particles.
template <int NX, int NY, int NZ, typename F>
__device__ __forceinline__ bool for_node_cells(const GridGeom &g, const CellLists &L, int i, int j, int k, F f) {
    bool any = false;
#pragma unroll
    for (int dk = 0; dk < NZ; dk++) {
        const int ck = k - (NZ - 1) + dk;
        if (ck < 0 || ck >= g.nex[2]) continue;
#pragma unroll
        for (int dj = 0; dj < NY; dj++) {
            const int cj = j - (NY - 1) + dj;
            if (cj < 0 || cj >= g.nex[1]) continue;
#pragma unroll
            for (int di = 0; di < NX; di++) {
                const int ci = i - (NX - 1) + di;
                if (ci < 0 || ci >= g.nex[0]) continue;
                const int c = ci + cj * g.nex[0] + ck * g.nex[0] * g.nex[1];
                if (!cell_nonempty(L, c)) continue;
                any = true;
                f(c, (NX - 1 - di) + NX * ((NY - 1 - dj) + NY * (NZ - 1 - dk)));
            }
        }
    }
    return any;
}

// Sum over the cell rows around control point (i, j, k), in ascending row
// order, of the stored runs (NQ of NQP floats) the x-merge leaves for it,
// with Neumaier sums; false when no cell around holds particles.
template <int NX, int NY, int NZ, int NQ, int NQP>
__device__ __forceinline__ bool stencil_node_sums(const GridGeom &g, const uint32_t *occ, const uint32_t *gs, int i,
                                                  int j, int k, const float *__restrict__ part, size_t ne, float *s,
                                                  float *cp) {
    // every row's occupancy and group-start words first (one round trip)
    RowWords w[NZ][NY];
    int base[NZ][NY];
#pragma unroll
    for (int dk = 0; dk < NZ; dk++) {
        const int ck = k - (NZ - 1) + dk;
#pragma unroll
        for (int dj = 0; dj < NY; dj++) {
            const int cj = j - (NY - 1) + dj;
            const bool ok = ck >= 0 && ck < g.nex[2] && cj >= 0 && cj < g.nex[1];
            base[dk][dj] = ok ? (cj + ck * g.nex[1]) * g.nex[0] : 0;
            w[dk][dj] = row_words<NX>(occ, gs, g.nex[0], base[dk][dj], i, ok);
        }
    }
    bool any = false;
    // then the runs, one z-layer of rows per round trip; a run that is not
    // added reads the buffer's first run (one line for the whole wave)
#pragma unroll
    for (int dk = 0; dk < NZ; dk++) {
        bool rd[NY][NX];
        float4 v[NY][NX][NQP / 4];
#pragma unroll
        for (int dj = 0; dj < NY; dj++) {
            any = row_reads<NX>(w[dk][dj], base[dk][dj], i, rd[dj]) || any;
            const int qyz = NX * ((NY - 1 - dj) + NY * (NZ - 1 - dk));
#pragma unroll
            for (int d = 0; d < NX; d++) {
                const size_t off = rd[dj][d] ? ((size_t)(d + qyz) * ne + (size_t)(base[dk][dj] + i - d)) * NQP : 0;
                const float4 *src = (const float4 *)(part + off);
#pragma unroll
                for (int h = 0; h < NQP / 4; h++) v[dj][d][h] = src[h];
            }
        }
        // ascending cell order within each row: column NX - 1 of the leftmost cell first
#pragma unroll
        for (int dj = 0; dj < NY; dj++)
#pragma unroll
            for (int d = NX - 1; d >= 0; d--) {
                if (!rd[dj][d]) continue;
#pragma unroll
                for (int f = 0; f < NQ; f++) {
                    const float4 q = v[dj][d][f >> 2];
                    neu_add(s[f], cp[f], (f & 3) == 0 ? q.x : (f & 3) == 1 ? q.y : (f & 3) == 2 ? q.z : q.w);
                }
            }
    }
    return any;
}

template <int NX, int NY, int NZ, bool DIAG>
__global__ void __launch_bounds__(256) ks_p2g_nodes(GridGeom g, CellLists L, TileGeom tg, const float *__restrict__ part,
                                                     NodesF32 N, BcDense bc, StepConsts sc) {
    constexpr int NN = NX * NY * NZ;
    int n, i, j, k;
    if (!tile_node(g, L, tg, DIAG, n, i, j, k)) return;
    const size_t ne = (size_t)g.nelem;
    constexpr int NQ = DIAG ? 8 : 7;
    float s[NQ], cp[NQ];
#pragma unroll
    for (int f = 0; f < NQ; f++) s[f] = cp[f] = 0.f;
    const bool any = stencil_node_sums<NX, NY, NZ, NQ, 8>(g, L.occ, L.gs, i, j, k, part, ne, s, cp);
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
                N.nvp[c * nodec + n] = 0.f;
            }
        }
        return;
    }
    const float m = s[0] + cp[0];
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
    const uint8_t mk = bc.mask ? bc.mask[n] : 0;
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
// SCH: 0 PIC, 1 FLIP, 2 TPIC, 3 APIC. G2P velocity from the cell's node
// values (PIC family: v + dt a, FLIP: the particle's own velocity plus dt a),
// the TPIC velocity gradient C_ab = sum dN/dx_b v_a (VarGradG2P) or the APIC
// B_ab = sum N v_a (x_a - x_p)_b and D = sum N (x_a - x_p)(x_a - x_p)^T
// (VarBmatG2P, DmatG2P, InvDmatG2P), then the MUSL momentum partials
// N m (v + M (x_a - x_p)) of the corners.
template <int NX, int NY, int NZ, int SCH>
__global__ void __launch_bounds__(kSeg) ks_musl_cells(GridGeom g, PartF32 P, CellLists L, NodeTables T, NodesF32 N,
                                                       StepConsts sc, float *__restrict__ part) {
    constexpr int NN = NX * NY * NZ;
    constexpr bool FLIP = SCH == 1;
    constexpr int AFF = SCH >= 2 ? SCH - 1 : 0;
    // staged: 0-2 r, 3 m, [4-6 old v for FLIP]; results: v (3), then C or B (9), then D^-1 (9)
    constexpr int NF = 4 + (FLIP ? 3 : 0);
    constexpr int NO = 3 + (AFF >= 1 ? 9 : 0) + (AFF == 2 ? 9 : 0);
    __shared__ float stage[NF * kBlk * kStageS];
    __shared__ float vout[NO * kBlk * kStageS];
    const uint32_t n_active = L.ctr->n_active;
    const int ngroups = (int)((n_active + kSeg - 1) / kSeg);
    const int t = threadIdx.x;
    const size_t ld = P.ld;
    const int nodec = g.nodec;
    const float inv[3] = {T.inv_dx[0], T.inv_dx[1], T.inv_dx[2]};
    const float dxs[3] = {T.dx[0], T.dx[1], T.dx[2]};
    for (int grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
        CellInfo si;
        uint32_t S, E;
        wave_fence();
        group_setup<1>(g, L, n_active, grp, si, S, E);
        const bool ne = si.ce > si.cs;
        float nvl[NN][3];
        if (ne) {
            const int n0 = si.ix + g.nxc[0] * si.iy + g.nxc[0] * g.nxc[1] * si.iz;
#pragma unroll
            for (int k = 0; k < NZ; k++)
#pragma unroll
                for (int j = 0; j < NY; j++)
#pragma unroll
                    for (int i = 0; i < NX; i++) {
                        const int q = i + NX * (j + NY * k);
                        const int nid = n0 + i + j * g.nxc[0] + k * g.nxc[0] * g.nxc[1];
#pragma unroll
                        for (int c = 0; c < 3; c++) nvl[q][c] = FLIP ? N.na[c * nodec + nid] : N.nvp[c * nodec + nid];
                    }
        } else {
#pragma unroll
            for (int q = 0; q < NN; q++)
#pragma unroll
                for (int c = 0; c < 3; c++) nvl[q][c] = 0.f;
        }
        float acc[NN][3];
#pragma unroll
        for (int q = 0; q < NN; q++)
#pragma unroll
            for (int c = 0; c < 3; c++) acc[q][c] = 0.f;

        for (uint32_t base = S; base < E; base += kChunk) {
            const int nch = (int)min((uint32_t)kChunk, E - base);
            uint32_t slot[kBlk];
            load_slots(L, base, nch, slot);
#pragma unroll
            for (int r = 0; r < kBlk; r++) {
                if (64 * r + t < nch) {
                    const uint32_t sl = slot[r];
#pragma unroll
                    for (int a = 0; a < 3; a++) {
                        lds_gather<kStageS>(P.r + a * ld, sl, stage, a, r);
                        if (FLIP) lds_gather<kStageS>(P.v + a * ld, sl, stage, 4 + a, r);
                    }
                    lds_gather<kStageS>(P.m, sl, stage, 3, r);
                }
            }
            wait_vm_all();
            wave_fence();
            const uint32_t lo = max(si.cs, base), hi = min(si.ce, base + (uint32_t)nch);
            for (uint32_t pos = lo; pos < hi; pos++) {
                const int i = (int)(pos - base);
                AxisW<NX> Ax;
                AxisW<NY> Ay;
                AxisW<NZ> Az;
                axis_weights<NX>(si.ix, stage[st_idx<kStageS>(0, i)], g.nex[0], Ax);
                axis_weights<NY>(si.iy, stage[st_idx<kStageS>(1, i)], g.nex[1], Ay);
                axis_weights<NZ>(si.iz, stage[st_idx<kStageS>(2, i)], g.nex[2], Az);
                const float m = stage[st_idx<kStageS>(3, i)];
                float vp[3] = {0.f, 0.f, 0.f};
                float Cm[9];
#pragma unroll
                for (int q = 0; q < 9; q++) Cm[q] = 0.f;
#pragma unroll
                for (int k = 0; k < NZ; k++)
#pragma unroll
                    for (int j = 0; j < NY; j++)
#pragma unroll
                        for (int ii = 0; ii < NX; ii++) {
                            const int q = ii + NX * (j + NY * k);
                            const float wyz = Ay.w[j] * Az.w[k];
                            const float Nw = Ax.w[ii] * wyz;
#pragma unroll
                            for (int c = 0; c < 3; c++) vp[c] += Nw * nvl[q][c];
                            if constexpr (AFF == 1) {
                                // C_ab += dN/dx_b v_a
                                const float gr[3] = {Ax.d[ii] * wyz * inv[0], Ax.w[ii] * (Ay.d[j] * Az.w[k]) * inv[1],
                                                     Ax.w[ii] * (Ay.w[j] * Az.d[k]) * inv[2]};
#pragma unroll
                                for (int a = 0; a < 3; a++)
#pragma unroll
                                    for (int b = 0; b < 3; b++) Cm[3 * a + b] += gr[b] * nvl[q][a];
                            } else if constexpr (AFF == 2) {
                                // B_ab += N v_a (x_a - x_p)_b
                                const float dd[3] = {Ax.x[ii] * dxs[0], Ay.x[j] * dxs[1], Az.x[k] * dxs[2]};
#pragma unroll
                                for (int a = 0; a < 3; a++) {
                                    const float nv = Nw * nvl[q][a];
#pragma unroll
                                    for (int b = 0; b < 3; b++) Cm[3 * a + b] += nv * dd[b];
                                }
                            }
                        }
                if (FLIP) {
#pragma unroll
                    for (int c = 0; c < 3; c++) vp[c] = stage[st_idx<kStageS>(4 + c, i)] + sc.dt * vp[c];
                }
                float Mh[3][3];
                float Di[9];
                if constexpr (AFF == 2) {
                    // D = sum N (x_a - x_p)(x_a - x_p)^T factorises over the axes:
                    // D_aa = dx_a^2 S2_a prod S0, D_ab = dx_a dx_b S1_a S1_b S0_c
                    float s0[3] = {0.f, 0.f, 0.f}, s1[3] = {0.f, 0.f, 0.f}, s2[3] = {0.f, 0.f, 0.f};
#pragma unroll
                    for (int u = 0; u < NX; u++) {
                        s0[0] += Ax.w[u];
                        s1[0] += Ax.w[u] * Ax.x[u];
                        s2[0] += Ax.w[u] * Ax.x[u] * Ax.x[u];
                    }
#pragma unroll
                    for (int u = 0; u < NY; u++) {
                        s0[1] += Ay.w[u];
                        s1[1] += Ay.w[u] * Ay.x[u];
                        s2[1] += Ay.w[u] * Ay.x[u] * Ay.x[u];
                    }
#pragma unroll
                    for (int u = 0; u < NZ; u++) {
                        s0[2] += Az.w[u];
                        s1[2] += Az.w[u] * Az.x[u];
                        s2[2] += Az.w[u] * Az.x[u] * Az.x[u];
                    }
                    float D[9];
#pragma unroll
                    for (int a = 0; a < 3; a++)
#pragma unroll
                        for (int b = 0; b < 3; b++) {
                            const int c3 = 3 - a - b;  // the third axis when a != b
                            D[3 * a + b] = (a == b) ? dxs[a] * dxs[a] * s2[a] * s0[(a + 1) % 3] * s0[(a + 2) % 3]
                                                    : dxs[a] * dxs[b] * s1[a] * s1[b] * s0[c3];
                        }
                    inv3_f(D, Di);
#pragma unroll
                    for (int a = 0; a < 3; a++)
#pragma unroll
                        for (int b = 0; b < 3; b++)
                            Mh[a][b] = (Cm[3 * a] * Di[b] + Cm[3 * a + 1] * Di[3 + b] + Cm[3 * a + 2] * Di[6 + b]) *
                                       dxs[b];
                } else if constexpr (AFF == 1) {
#pragma unroll
                    for (int a = 0; a < 3; a++)
#pragma unroll
                        for (int b = 0; b < 3; b++) Mh[a][b] = Cm[3 * a + b] * dxs[b];
                }
                vout[st_idx<kStageS>(0, i)] = vp[0];
                vout[st_idx<kStageS>(1, i)] = vp[1];
                vout[st_idx<kStageS>(2, i)] = vp[2];
                if constexpr (AFF >= 1) {
#pragma unroll
                    for (int q = 0; q < 9; q++) vout[st_idx<kStageS>(3 + q, i)] = Cm[q];
                }
                if constexpr (AFF == 2) {
#pragma unroll
                    for (int q = 0; q < 9; q++) vout[st_idx<kStageS>(12 + q, i)] = Di[q];
                }
                const float mv[3] = {m * vp[0], m * vp[1], m * vp[2]};
                float px[3][NX], py[3][NY], pz[3][NZ];
                if constexpr (AFF >= 1) {
#pragma unroll
                    for (int a = 0; a < 3; a++) {
#pragma unroll
                        for (int u = 0; u < NX; u++) px[a][u] = mv[a] + (m * Mh[a][0]) * Ax.x[u];
#pragma unroll
                        for (int u = 0; u < NY; u++) py[a][u] = (m * Mh[a][1]) * Ay.x[u];
#pragma unroll
                        for (int u = 0; u < NZ; u++) pz[a][u] = (m * Mh[a][2]) * Az.x[u];
                    }
                }
#pragma unroll
                for (int k = 0; k < NZ; k++)
#pragma unroll
                    for (int j = 0; j < NY; j++)
#pragma unroll
                        for (int ii = 0; ii < NX; ii++) {
                            const int q = ii + NX * (j + NY * k);
                            const float Nw = Ax.w[ii] * (Ay.w[j] * Az.w[k]);
#pragma unroll
                            for (int a = 0; a < 3; a++) {
                                const float va = (AFF >= 1) ? (px[a][ii] + py[a][j]) + pz[a][k] : mv[a];
                                acc[q][a] += Nw * va;
                            }
                        }
            }
            wave_fence();
#pragma unroll
            for (int r = 0; r < kBlk; r++) {
                const int i = 64 * r + t;
                if (i < nch) {
                    const uint32_t sl = slot[r];
#pragma unroll
                    for (int c = 0; c < 3; c++) __builtin_nontemporal_store(vout[st_idx<kStageS>(c, i)], &P.v[c * ld + sl]);
                    if constexpr (AFF >= 1) {
#pragma unroll
                        for (int q = 0; q < 9; q++) P.C[q * ld + sl] = vout[st_idx<kStageS>(3 + q, i)];
                    }
                    if constexpr (AFF == 2) {
#pragma unroll
                        for (int q = 0; q < 9; q++) P.Dinv[q * ld + sl] = vout[st_idx<kStageS>(12 + q, i)];
                    }
                }
            }
            wave_fence();
        }
        // corner rows merged along x (the MUSL grouping: one lane per cell), 16-byte runs
        const XLinks xl = x_links<kSeg>(si, t, g.nex[0], L.gsize2 > 1u);
        const size_t nel = (size_t)g.nelem;
#pragma unroll
        for (int k = 0; k < NZ; k++)
#pragma unroll
            for (int j = 0; j < NY; j++) {
                float S[NX][3];
#pragma unroll
                for (int ii = 0; ii < NX; ii++)
#pragma unroll
                    for (int c = 0; c < 3; c++) S[ii][c] = acc[ii + NX * (j + NY * k)][c];
                x_merge_row<NX, 3>(S, xl.link);
                if (ne) {
#pragma unroll
                    for (int ii = 0; ii < NX; ii++)
                        if (ii == 0 || xl.keep_right) store_run<3, 4>(part, nel, ii + NX * (j + NY * k), si.c, S[ii]);
                }
            }
    }
}

template <int NX, int NY, int NZ, bool DIAG>
__global__ void __launch_bounds__(256) ks_musl_nodes(GridGeom g, CellLists L, TileGeom tg,
                                                      const float *__restrict__ part, NodesF32 N, BcDense bc,
                                                      StepConsts sc) {
    int n, i, j, k;
    if (!tile_node(g, L, tg, DIAG, n, i, j, k)) return;
    const size_t ne = (size_t)g.nelem;
    float s[3] = {0.f, 0.f, 0.f}, cp[3] = {0.f, 0.f, 0.f};
    const bool any = stencil_node_sums<NX, NY, NZ, 3, 4>(g, L.occ, L.gs2, i, j, k, part, ne, s, cp);
    const int nodec = g.nodec;
    if (!any) {
        if (DIAG)
            for (int c = 0; c < 3; c++) N.nvm[c * nodec + n] = 0.f;
        return;
    }
    const float m = N.nm[n];
    float v[3] = {0.f, 0.f, 0.f};
    const float p[3] = {s[0] + cp[0], s[1] + cp[1], s[2] + cp[2]};
    if ((double)m > sc.mtol) {
#pragma unroll
        for (int c = 0; c < 3; c++) v[c] = safe_div(p[c], m, sc.m_small);
    }
    const uint8_t mk = bc.mask ? bc.mask[n] : 0;
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
// dF - I = dt sum v_a (x) grad N_a over the stencil, H, J (FP64), x, V,
// stress and the new cell key, in slot order; fused cell counting for the
// next binning as in the tent-function update kernel.
// FB (F-bar projection): F and the increment come from ks_fbar_grad; the
// node average of Jbar det dF is interpolated here (Jbar = sum N nJbar,
// fbar_projection.cpp:53-64), the increment is scaled by (Jbar / (Jbar_old
// det dF))^(1/3) (:66-71), Fbar advances by it (:73-86), and the stress
// follows Fbar, Jbar and the scaled increment (update_point_coord_and_stress
// with def_grad_bar, var_trans_solid_explicit.cpp:298-302).
template <int NX, int NY, int NZ, bool FB>
__global__ void __launch_bounds__(256) ks_g2p_update(GridGeom g, PartF32 P, NodesF32 N, NodeTables T,
                                                      MaterialTable mt, StepConsts sc, uint32_t *err, BinFuse bf) {
    // stencils of 18 or more points gather through a per-wave LDS tile of node values
    constexpr bool kTileGather = NX * NY * NZ >= 18;
    constexpr int kTileX = 32;
    __shared__ std::conditional_t<kTileGather, float4[4][kTileX * NY * NZ], int> ntile;
    const int slot0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (bf.cnt) {
        for (int i = slot0; i < bf.ntiles; i += gridDim.x * blockDim.x) bf.tiles[i] = 0u;
        if (slot0 == 0) bf.ctr->n_active = 0u;
    }
    // every lane takes part in the wave-wide counting; invalid lanes compute
    // on slot 0 / cell 0 and store nothing
    bool valid = slot0 < P.n;
    const int slot = valid ? slot0 : 0;
    const size_t ld = P.ld;
    // The particle's own streams do not depend on its cell key: they are
    // issued with the key (the material index first), so their latency
    // overlaps the key -> node gather chain.
    const int mat = P.matid ? P.matid[slot] : 0;
    uint32_t key = P.key[slot];
    float r[3], H[9];
#pragma unroll
    for (int a = 0; a < 3; a++) r[a] = P.r[a * ld + slot];
#pragma unroll
    for (int q = 0; q < 9; q++) {
        if constexpr (FB)
            H[q] = P.H[q * ld + slot];  // F - I, advanced by ks_fbar_grad
        else
            H[q] = __builtin_nontemporal_load(&P.H[q * ld + slot]);
    }
    const float V0p = P.V0[slot];
    if (key >= (uint32_t)g.nelem) valid = false;
    if (!valid) key = 0;
    const int nodec = g.nodec;
    const int nexy = g.nex[0] * g.nex[1];
    int rem, ix, iy;
    const int iz = fdivmod((int)key, nexy, T.inv_nexy, rem);
    iy = fdivmod(rem, g.nex[0], T.inv_nex, ix);
    // The material is copied before the node gather, so that its loads
    // overlap the gather's, except for the 27-point stencil: there its
    // registers would cost a wave per SIMD.
    constexpr bool kLateMat = NX * NY * NZ == 27;
    MaterialDev md;
    if constexpr (!kLateMat) md = mt.m[mat];
    AxisW<NX> Ax;
    AxisW<NY> Ay;
    AxisW<NZ> Az;
    axis_weights<NX>(ix, r[0], g.nex[0], Ax);
    axis_weights<NY>(iy, r[1], g.nex[1], Ay);
    axis_weights<NZ>(iz, r[2], g.nex[2], Az);
    const float inv0 = T.inv_dx[0], inv1 = T.inv_dx[1], inv2 = T.inv_dx[2];
    const int n0 = ix + g.nxc[0] * iy + g.nxc[0] * g.nxc[1] * iz;
    float Dv[9];
#pragma unroll
    for (int q = 0; q < 9; q++) Dv[q] = 0.f;
    float dxp[3] = {0.f, 0.f, 0.f};
    float je = 0.f;  // FB: Jbar - 1 = sum N (nJbar - 1)
    // tensor-product form: along each node row (fixed j, k) the x sums
    // sum_i v N_x and sum_i v dN_x/dr first, then the row's y and z factors;
    // nodev(i, j, k) gives the node's MUSL velocity and (F-bar) nJbar - 1
    auto accumulate = [&](auto nodev) {
#pragma unroll
        for (int k = 0; k < NZ; k++)
#pragma unroll
            for (int j = 0; j < NY; j++) {
                float rw[3] = {0.f, 0.f, 0.f}, rd[3] = {0.f, 0.f, 0.f}, rj = 0.f;
#pragma unroll
                for (int i = 0; i < NX; i++) {
                    const float4 v4 = nodev(i, j, k);
                    const float v[3] = {v4.x, v4.y, v4.z};
#pragma unroll
                    for (int a = 0; a < 3; a++) {
                        rw[a] += Ax.w[i] * v[a];
                        if constexpr (!FB) rd[a] += Ax.d[i] * v[a];
                    }
                    if constexpr (FB) rj += Ax.w[i] * v4.w;
                }
                const float wyz = Ay.w[j] * Az.w[k];
#pragma unroll
                for (int a = 0; a < 3; a++) dxp[a] += wyz * rw[a];
                if constexpr (FB) {
                    je += wyz * rj;
                } else {
                    const float dyz = Ay.d[j] * Az.w[k], wydz = Ay.w[j] * Az.d[k];
#pragma unroll
                    for (int a = 0; a < 3; a++) {
                        Dv[3 * a] += wyz * rd[a];
                        Dv[3 * a + 1] += dyz * rw[a];
                        Dv[3 * a + 2] += wydz * rw[a];
                    }
                }
            }
    };
    if constexpr (kTileGather) {
        // The wave stages the node values of a run of cells in one cell row
        // (the lanes of the lowest pending lane's row whose stencils fit in
        // kTileX control points from the run's first cell) in its LDS tile
        // with a few coalesced loads per lane; those lanes then read their
        // NX NY NZ nodes from the tile. Repeated until every particle of the
        // wave is done: once after a relayout, when the wave's particles sit in
        // a few consecutive cells.
        float4 *tl = ntile[threadIdx.x >> 6];
        const int lane = threadIdx.x & 63;
        const int row = valid ? iy + g.nex[1] * iz : -1;
        unsigned long long pending = __ballot(valid);
        while (pending) {
            const int leader = __ffsll((unsigned long long)pending) - 1;
            const int rowL = __shfl(row, leader, 64);
            const bool cand = ((pending >> lane) & 1ull) && row == rowL;
            int xl = cand ? ix : 0x7fffffff, xh = cand ? ix : -1;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                xl = min(xl, __shfl_xor(xl, o, 64));
                xh = max(xh, __shfl_xor(xh, o, 64));
            }
            const int x0 = xl;
            const int xn = min(xh, x0 + kTileX - NX) - x0 + NX;  // control points staged along x
            const bool mine = cand && ix - x0 <= kTileX - NX;
            const int iyL = __shfl(iy, leader, 64), izL = __shfl(iz, leader, 64);
            const int cnt = xn * NY * NZ;
            for (int e = lane; e < cnt; e += 64) {
                const int xx = e % xn, yz = e / xn;
                const int nid = (x0 + xx) + g.nxc[0] * (iyL + yz % NY) + g.nxc[0] * g.nxc[1] * (izL + yz / NY);
                tl[e] = make_float4(N.nvm[nid], N.nvm[nodec + nid], N.nvm[2 * nodec + nid], FB ? N.nje[nid] : 0.f);
            }
            wave_fence();  // the tile is written before any lane reads it (one wave: LDS in order)
            if (mine) {
                const int xo = ix - x0;
                accumulate([&](int i, int j, int k) { return tl[(k * NY + j) * xn + xo + i]; });
            }
            pending &= ~__ballot(mine);
            wave_fence();  // every read of this tile precedes the next staging
        }
    } else {
        accumulate([&](int i, int j, int k) {
            const int nid = n0 + i + j * g.nxc[0] + k * g.nxc[0] * g.nxc[1];
            return make_float4(N.nvm[nid], N.nvm[nodec + nid], N.nvm[2 * nodec + nid], FB ? N.nje[nid] : 0.f);
        });
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        Dv[3 * a] *= inv0;
        Dv[3 * a + 1] *= inv1;
        Dv[3 * a + 2] *= inv2;
    }
    float D[9], Hn[9];
    if constexpr (FB) {
#pragma unroll
        for (int q = 0; q < 9; q++) {
            D[q] = P.Dd[q * ld + slot];
            Hn[q] = H[q];
        }
    } else {
#pragma unroll
        for (int q = 0; q < 9; q++) D[q] = sc.dt * Dv[q];
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++)
                Hn[3 * a + b] = (H[3 * a + b] + D[3 * a + b]) +
                                (D[3 * a] * H[b] + D[3 * a + 1] * H[3 + b] + D[3 * a + 2] * H[6 + b]);
    }
    const double J = det_I_plus_H(Hn);
    // F-bar: the scaled increment Ds and Hbar = Fbar - I it advances (the
    // stress then follows these two and Jbar)
    float Hs[9];
    double Jl = J;
    if constexpr (FB) {
        const double jm = (double)P.jm[slot];  // Jbar_old det dF - 1
        const double x = ((double)je - jm) / (1.0 + jm);
        const double c = cbrt(1.0 + x);
        const double cm1 = c - 1.0;
        float Hb[9];
#pragma unroll
        for (int q = 0; q < 9; q++) {
            D[q] = (float)(c * (double)D[q] + ((q == 0 || q == 4 || q == 8) ? cm1 : 0.0));
            Hb[q] = P.Hb[q * ld + slot];
        }
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++)
                Hs[3 * a + b] = (Hb[3 * a + b] + D[3 * a + b]) +
                                (D[3 * a] * Hb[b] + D[3 * a + 1] * Hb[3 + b] + D[3 * a + 2] * Hb[6 + b]);
        Jl = 1.0 + (double)je;
    } else {
#pragma unroll
        for (int q = 0; q < 9; q++) Hs[q] = Hn[q];
    }
    const double V = J * (double)V0p;
    if constexpr (kLateMat) md = mt.m[mat];
    int ci[3] = {ix, iy, iz};
    bool inside = true;
    const double b0 = sc.sh_geup ? *sc.sh_b0 : 0.0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float ra = r[a] + (sc.dt * dxp[a]) * T.inv_dx[a];
        if (sc.sh_geup) ra += (float)(-b0 * sc.sh_geup[a * ld + slot] * (double)T.inv_dx[a]);
        const float fl = floorf(ra);
        ra -= fl;
        ci[a] += (int)fl;
        if (ra >= 1.f) {
            ra = 0.f;
            ci[a] += 1;
        }
        r[a] = ra;
        inside = inside && ci[a] >= 0 && ci[a] < g.nex[a];
    }
    const bool rate_form = (md.model == 0 || md.model == 10);
    float s[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (rate_form) {
#pragma unroll
        for (int q = 0; q < 6; q++) s[q] = P.sig[q * ld + slot];
    }
    float ep = (md.model == 10 && P.eps_p) ? P.eps_p[slot] : 0.f;
    stress32(md, Hs, D, J, Jl, s, &ep, 1.f / sc.dt);
    const uint32_t nk = inside ? (uint32_t)(ci[0] + ci[1] * g.nex[0] + ci[2] * nexy) : (uint32_t)g.nelem;
    if (bf.cnt) {
        if (bf.delta) {
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
    if constexpr (FB) {
#pragma unroll
        for (int q = 0; q < 9; q++) P.Hb[q * ld + slot] = Hs[q];
        P.jb[slot] = je;
    } else {
#pragma unroll
        for (int q = 0; q < 9; q++) __builtin_nontemporal_store(Hn[q], &P.H[q * ld + slot]);
    }
#pragma unroll
    for (int q = 0; q < 6; q++) __builtin_nontemporal_store(s[q], &P.sig[q * ld + slot]);
#pragma unroll
    for (int a = 0; a < 3; a++) __builtin_nontemporal_store(r[a], &P.r[a * ld + slot]);
    __builtin_nontemporal_store((float)V, &P.V32[slot]);
    if (!inside) {
        P.J[slot] = J;
        P.V[slot] = V;
    }
    if (nk != key) P.key[slot] = nk;
    uint32_t e = 0;
    if (!inside) e |= kErrOutOfGrid;
    if (!(J > 0.0) || !(Jl > 0.0)) e |= kErrJacobian;
    if (!isfinite(r[0] + r[1] + r[2] + s[0] + s[1] + s[2])) e |= kErrNonFinite;
    if (e) atomicOr(err, e);
}

// ------------------------------------------------------ F-bar projection
// (fbar_projection.cpp:18-94 with the increment of
// var_trans_solid_explicit.cpp:244-296.) One thread per slot: dF - I, the
// new F (stored as H) and Jbar_old det dF - 1 for the node average.
template <int NX, int NY, int NZ>
__global__ void __launch_bounds__(256) ks_fbar_grad(GridGeom g, PartF32 P, NodesF32 N, NodeTables T, StepConsts sc) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= P.n) return;
    const uint32_t key = P.key[slot];
    if (key >= (uint32_t)g.nelem) return;
    const size_t ld = P.ld;
    const int nodec = g.nodec;
    const int nexy = g.nex[0] * g.nex[1];
    int rem, ix, iy;
    const int iz = fdivmod((int)key, nexy, T.inv_nexy, rem);
    iy = fdivmod(rem, g.nex[0], T.inv_nex, ix);
    AxisW<NX> Ax;
    AxisW<NY> Ay;
    AxisW<NZ> Az;
    axis_weights<NX>(ix, P.r[slot], g.nex[0], Ax);
    axis_weights<NY>(iy, P.r[ld + slot], g.nex[1], Ay);
    axis_weights<NZ>(iz, P.r[2 * ld + slot], g.nex[2], Az);
    const float inv0 = T.inv_dx[0], inv1 = T.inv_dx[1], inv2 = T.inv_dx[2];
    const int n0 = ix + g.nxc[0] * iy + g.nxc[0] * g.nxc[1] * iz;
    float Dv[9];
#pragma unroll
    for (int q = 0; q < 9; q++) Dv[q] = 0.f;
#pragma unroll
    for (int k = 0; k < NZ; k++)
#pragma unroll
        for (int j = 0; j < NY; j++)
#pragma unroll
            for (int i = 0; i < NX; i++) {
                const int nid = n0 + i + j * g.nxc[0] + k * g.nxc[0] * g.nxc[1];
                const float v0 = N.nvm[nid], v1 = N.nvm[nodec + nid], v2 = N.nvm[2 * nodec + nid];
                const float wyz = Ay.w[j] * Az.w[k];
                const float gx = (Ax.d[i] * inv0) * wyz;
                const float gy = (Ay.d[j] * inv1) * (Ax.w[i] * Az.w[k]);
                const float gz = (Az.d[k] * inv2) * (Ax.w[i] * Ay.w[j]);
                Dv[0] += v0 * gx;
                Dv[1] += v0 * gy;
                Dv[2] += v0 * gz;
                Dv[3] += v1 * gx;
                Dv[4] += v1 * gy;
                Dv[5] += v1 * gz;
                Dv[6] += v2 * gx;
                Dv[7] += v2 * gy;
                Dv[8] += v2 * gz;
            }
    float D[9], H[9];
#pragma unroll
    for (int q = 0; q < 9; q++) {
        D[q] = sc.dt * Dv[q];
        H[q] = P.H[q * ld + slot];
    }
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++)
            P.H[(3 * a + b) * ld + slot] = (H[3 * a + b] + D[3 * a + b]) +
                                           (D[3 * a] * H[b] + D[3 * a + 1] * H[3 + b] + D[3 * a + 2] * H[6 + b]);
#pragma unroll
    for (int q = 0; q < 9; q++) P.Dd[q * ld + slot] = D[q];
    const double dm = det_I_plus_H(D) - 1.0, jb = (double)P.jb[slot];
    P.jm[slot] = (float)(jb + dm + jb * dm);
}

// Cell partials of sum N (Jbar det dF - 1) V over each corner (the cell's
// particles in sorted order), one lane per cell, particles staged through
// LDS as in the P2G cells kernel.
template <int NX, int NY, int NZ>
__global__ void __launch_bounds__(kSeg) ks_fbar_cells(GridGeom g, PartF32 P, CellLists L, float *__restrict__ part) {
    constexpr int NN = NX * NY * NZ;
    __shared__ float stage[5 * kBlk * kStageS];  // r (3), Jbar det dF - 1, V
    const uint32_t n_active = L.ctr->n_active;
    const int ngroups = (int)((n_active + kSeg - 1) / kSeg);
    const int t = threadIdx.x;
    const size_t ld = P.ld;
    for (int grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
        CellInfo si;
        uint32_t S, E;
        wave_fence();
        group_setup<1>(g, L, n_active, grp, si, S, E);
        float acc[NN];
#pragma unroll
        for (int q = 0; q < NN; q++) acc[q] = 0.f;
        for (uint32_t base = S; base < E; base += kChunk) {
            const int nch = (int)min((uint32_t)kChunk, E - base);
            uint32_t slot[kBlk];
            load_slots(L, base, nch, slot);
#pragma unroll
            for (int r = 0; r < kBlk; r++) {
                if (64 * r + t < nch) {
                    const uint32_t sl = slot[r];
#pragma unroll
                    for (int a = 0; a < 3; a++) lds_gather<kStageS>(P.r + a * ld, sl, stage, a, r);
                    lds_gather<kStageS>(P.jm, sl, stage, 3, r);
                    lds_gather<kStageS>(P.V32, sl, stage, 4, r);
                }
            }
            wait_vm_all();
            wave_fence();
            const uint32_t lo = max(si.cs, base), hi = min(si.ce, base + (uint32_t)nch);
            for (uint32_t pos = lo; pos < hi; pos++) {
                const int i = (int)(pos - base);
                AxisW<NX> Ax;
                AxisW<NY> Ay;
                AxisW<NZ> Az;
                axis_weights<NX>(si.ix, stage[st_idx<kStageS>(0, i)], g.nex[0], Ax);
                axis_weights<NY>(si.iy, stage[st_idx<kStageS>(1, i)], g.nex[1], Ay);
                axis_weights<NZ>(si.iz, stage[st_idx<kStageS>(2, i)], g.nex[2], Az);
                const float jv = stage[st_idx<kStageS>(3, i)] * stage[st_idx<kStageS>(4, i)];
#pragma unroll
                for (int k = 0; k < NZ; k++)
#pragma unroll
                    for (int j = 0; j < NY; j++) {
                        const float wyz = Ay.w[j] * Az.w[k];
#pragma unroll
                        for (int ii = 0; ii < NX; ii++) acc[ii + NX * (j + NY * k)] += (Ax.w[ii] * wyz) * jv;
                    }
            }
            wave_fence();
        }
        if (si.ce > si.cs) {
            const size_t ne = (size_t)g.nelem;
#pragma unroll
            for (int q = 0; q < NN; q++) part[(size_t)q * ne + si.c] = acc[q];
        }
    }
}

// Node average nJbar - 1 = sum N (Jbar det dF - 1) V / nvof where nvof > mtol,
// and -1 (nJbar = 0) elsewhere (fbar_projection.cpp:44-51). The sum over a
// node's cells is Neumaier-compensated in ascending cell order.
template <int NX, int NY, int NZ>
__global__ void __launch_bounds__(256) ks_fbar_nodes(GridGeom g, CellLists L, TileGeom tg, const float *__restrict__ part,
                                                      NodesF32 N, StepConsts sc) {
    int n, i, j, k;
    if (!tile_node(g, L, tg, false, n, i, j, k)) return;
    const size_t ne = (size_t)g.nelem;
    float sm = 0.f, cp = 0.f;
    const bool any = for_node_cells<NX, NY, NZ>(g, L, i, j, k, [&](int c, int q) {
        neu_add(sm, cp, part[(size_t)q * ne + c]);
    });
    if (!any) return;
    const float nv = N.nvof[n];
    N.nje[n] = ((double)nv > sc.mtol) ? (float)((double)(sm + cp) / (double)nv) : -1.f;
}

// ------------------------------------------------------ particle shifting
// (module/material_point.cpp:174-227; global sums in blocks of kShiftBlock
// consecutive entries, then over the blocks in order, as the FP64 engine)
__global__ void ks_shift_nei(int nodec, const float *__restrict__ nvof, ShiftF64 S) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const int i0 = b * kShiftBlock;
    if (i0 >= nodec) return;
    double acc = 0.0;
    for (int i = i0; i < nodec && i < i0 + kShiftBlock; i++) {
        const double e = (double)nvof[i] - S.nvol[i];
        const double ne = e > 0.0 ? e : 0.0;
        S.nei[i] = ne;
        acc += ne * ne;
    }
    S.part_eu[b] = acc;
}

template <int NX, int NY, int NZ>
__global__ void __launch_bounds__(256) ks_shift_geup(GridGeom g, PartF32 P, NodeTables T, ShiftF64 S) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= P.n) return;
    const size_t ld = P.ld;
    const uint32_t key = P.key[slot];
    double gr[3] = {0.0, 0.0, 0.0};
    if (key < (uint32_t)g.nelem) {
        const int nexy = g.nex[0] * g.nex[1];
        int rem, ix;
        const int iz = fdivmod((int)key, nexy, T.inv_nexy, rem);
        const int iy = fdivmod(rem, g.nex[0], T.inv_nex, ix);
        AxisW<NX> Ax;
        AxisW<NY> Ay;
        AxisW<NZ> Az;
        axis_weights<NX>(ix, P.r[slot], g.nex[0], Ax);
        axis_weights<NY>(iy, P.r[ld + slot], g.nex[1], Ay);
        axis_weights<NZ>(iz, P.r[2 * ld + slot], g.nex[2], Az);
        const int n0 = ix + g.nxc[0] * iy + g.nxc[0] * g.nxc[1] * iz;
#pragma unroll
        for (int k = 0; k < NZ; k++)
#pragma unroll
            for (int j = 0; j < NY; j++)
#pragma unroll
                for (int i = 0; i < NX; i++) {
                    const double ne = S.nei[n0 + i + j * g.nxc[0] + k * g.nxc[0] * g.nxc[1]];
                    gr[0] += (double)(Ax.d[i] * T.inv_dx[0] * (Ay.w[j] * Az.w[k])) * ne;
                    gr[1] += (double)(Ay.d[j] * T.inv_dx[1] * (Ax.w[i] * Az.w[k])) * ne;
                    gr[2] += (double)(Az.d[k] * T.inv_dx[2] * (Ax.w[i] * Ay.w[j])) * ne;
                }
        const double s2 = 2.0 * (double)P.V32[slot];
#pragma unroll
        for (int c = 0; c < 3; c++) gr[c] *= s2;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) S.geup[c * ld + slot] = gr[c];
}

__global__ void ks_shift_dot(int n, size_t ld, ShiftF64 S) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const int i0 = b * kShiftBlock;
    if (i0 >= n) return;
    double acc = 0.0;
    for (int i = i0; i < n && i < i0 + kShiftBlock; i++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double v = S.geup[c * ld + i];
            acc += v * v;
        }
    S.part_dot[b] = acc;
}

__global__ void ks_shift_b0(int nbe, int nbd, double mtol, ShiftF64 S) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double eu = 0.0, dot = 0.0;
    for (int b = 0; b < nbe; b++) eu += S.part_eu[b];
    for (int b = 0; b < nbd; b++) dot += S.part_dot[b];
    *S.b0 = (dot < mtol) ? 0.0 : (eu / dot);
}

int cell_groups(const GridGeom &g, int cells_per_group = 64) {
    const int ng = cdiv(g.nelem, cells_per_group);
    return ng < 8192 ? ng : 8192;
}

template <int NX, int NY, int NZ>
void p2g_launch(hipStream_t s, const GridGeom &g, const PartF32 &P, const NodesF32 &N, const CellLists &L,
                const BcDense &bc, const NodeTables &T, const TileGeom &tg, const StepConsts &c, float *partials,
                int diag, hipEvent_t mid) {
    // the quadratic z-axis takes one z-layer per lane (3 lanes per cell)
    constexpr int LZ = (NZ == 3 && NX * NY * NZ >= 18) ? NZ : 1;
    const int ng = c.cells_grid > 0 ? c.cells_grid : cell_groups(g, kSeg / LZ);
    const bool lean = !diag && P.trac == nullptr;
    auto cells = [&](auto aff, auto lean_c) {
        hipLaunchKernelGGL((ks_p2g_cells<NX, NY, NZ, decltype(aff)::value, LZ, decltype(lean_c)::value>), dim3(ng),
                           dim3(kSeg), 0, s, g, P, L, T, partials);
    };
    using A0 = std::integral_constant<int, 0>;
    using A1 = std::integral_constant<int, 1>;
    using A2 = std::integral_constant<int, 2>;
    if (lean) {
        if (c.scheme == 2) cells(A1{}, std::true_type{});
        else if (c.scheme == 3) cells(A2{}, std::true_type{});
        else cells(A0{}, std::true_type{});
    } else {
        if (c.scheme == 2) cells(A1{}, std::false_type{});
        else if (c.scheme == 3) cells(A2{}, std::false_type{});
        else cells(A0{}, std::false_type{});
    }
    if (mid) (void)hipEventRecord(mid, s);
    if (diag)
        hipLaunchKernelGGL((ks_p2g_nodes<NX, NY, NZ, true>), dim3(tg.ntiles), dim3(256), 0, s, g, L, tg,
                           (const float *)partials, N, bc, c);
    else
        hipLaunchKernelGGL((ks_p2g_nodes<NX, NY, NZ, false>), dim3(tg.ntiles), dim3(256), 0, s, g, L, tg,
                           (const float *)partials, N, bc, c);
}

template <int NX, int NY, int NZ>
void musl_launch(hipStream_t s, const GridGeom &g, const PartF32 &P, const NodesF32 &N, const CellLists &L,
                 const BcDense &bc, const NodeTables &T, const TileGeom &tg, const StepConsts &c, float *partials,
                 int diag, hipEvent_t mid) {
    const int ng = c.cells_grid > 0 ? c.cells_grid : cell_groups(g);
    switch (c.scheme) {
    case 1:
        hipLaunchKernelGGL((ks_musl_cells<NX, NY, NZ, 1>), dim3(ng), dim3(kSeg), 0, s, g, P, L, T, N, c, partials);
        break;
    case 2:
        hipLaunchKernelGGL((ks_musl_cells<NX, NY, NZ, 2>), dim3(ng), dim3(kSeg), 0, s, g, P, L, T, N, c, partials);
        break;
    case 3:
        hipLaunchKernelGGL((ks_musl_cells<NX, NY, NZ, 3>), dim3(ng), dim3(kSeg), 0, s, g, P, L, T, N, c, partials);
        break;
    default:
        hipLaunchKernelGGL((ks_musl_cells<NX, NY, NZ, 0>), dim3(ng), dim3(kSeg), 0, s, g, P, L, T, N, c, partials);
        break;
    }
    if (mid) (void)hipEventRecord(mid, s);
    if (diag)
        hipLaunchKernelGGL((ks_musl_nodes<NX, NY, NZ, true>), dim3(tg.ntiles), dim3(256), 0, s, g, L, tg,
                           (const float *)partials, N, bc, c);
    else
        hipLaunchKernelGGL((ks_musl_nodes<NX, NY, NZ, false>), dim3(tg.ntiles), dim3(256), 0, s, g, L, tg,
                           (const float *)partials, N, bc, c);
}

// stencil widths (points per axis) of the supported idimc combinations:
// 0 = (1,1,1), 1 = (2,2,2), 2 = (2,1,2); -1 otherwise
int stencil_kind(const GridGeom &g) {
    const int a = g.idimc[0], b = g.idimc[1], c = g.idimc[2];
    if (a == 1 && b == 1 && c == 1) return 0;
    if (a == 2 && b == 2 && c == 2) return 1;
    if (a == 2 && b == 1 && c == 2) return 2;
    return -1;
}

}  // namespace

bool s32_supported(const GridGeom &g) { return stencil_kind(g) >= 0; }

int s32_p2g_cells_per_group(const GridGeom &g) {
    // p2g_launch: the quadratic z-axis of a stencil of 18 or more points takes one z-layer per lane
    const int nn = (g.idimc[0] + 1) * (g.idimc[1] + 1) * (g.idimc[2] + 1);
    return (g.idimc[2] == 2 && nn >= 18) ? kSeg / 3 : kSeg;
}

int s32_partials_per_cell(const GridGeom &g) {
    const int nn = (g.idimc[0] + 1) * (g.idimc[1] + 1) * (g.idimc[2] + 1);
    return 8 * nn;  // 7 P2G quantities + volume per corner (MUSL uses 3)
}

void launch_s32_p2g(hipStream_t s, const GridGeom &g, const PartF32 &P, const NodesF32 &N, const CellLists &L,
                    const BcDense &bc, const NodeTables &T, const TileGeom &tg, const StepConsts &c, float *partials,
                    int diag, hipEvent_t mid) {
    switch (stencil_kind(g)) {
    case 0: p2g_launch<2, 2, 2>(s, g, P, N, L, bc, T, tg, c, partials, diag, mid); break;
    case 1: p2g_launch<3, 3, 3>(s, g, P, N, L, bc, T, tg, c, partials, diag, mid); break;
    case 2: p2g_launch<3, 2, 3>(s, g, P, N, L, bc, T, tg, c, partials, diag, mid); break;
    default: break;
    }
}

void launch_s32_g2p_vel_musl(hipStream_t s, const GridGeom &g, const PartF32 &P, const NodesF32 &N,
                             const CellLists &L, const BcDense &bc, const NodeTables &T, const TileGeom &tg,
                             const StepConsts &c, float *partials, int diag, hipEvent_t mid) {
    switch (stencil_kind(g)) {
    case 0: musl_launch<2, 2, 2>(s, g, P, N, L, bc, T, tg, c, partials, diag, mid); break;
    case 1: musl_launch<3, 3, 3>(s, g, P, N, L, bc, T, tg, c, partials, diag, mid); break;
    case 2: musl_launch<3, 2, 3>(s, g, P, N, L, bc, T, tg, c, partials, diag, mid); break;
    default: break;
    }
}

void launch_s32_shift(hipStream_t s, const GridGeom &g, const PartF32 &P, const NodesF32 &N, const NodeTables &T,
                      const ShiftF64 &S, double mtol) {
    if (P.n <= 0) return;
    const int nbe = cdiv(g.nodec, kShiftBlock), nbd = cdiv(P.n, kShiftBlock);
    hipLaunchKernelGGL(ks_shift_nei, dim3(cdiv(nbe, 64)), dim3(64), 0, s, g.nodec, (const float *)N.nvof, S);
    const dim3 grid(cdiv(P.n, 256)), block(256);
    switch (stencil_kind(g)) {
    case 0: hipLaunchKernelGGL((ks_shift_geup<2, 2, 2>), grid, block, 0, s, g, P, T, S); break;
    case 1: hipLaunchKernelGGL((ks_shift_geup<3, 3, 3>), grid, block, 0, s, g, P, T, S); break;
    case 2: hipLaunchKernelGGL((ks_shift_geup<3, 2, 3>), grid, block, 0, s, g, P, T, S); break;
    default: break;
    }
    hipLaunchKernelGGL(ks_shift_dot, dim3(cdiv(nbd, 64)), dim3(64), 0, s, P.n, P.ld, S);
    hipLaunchKernelGGL(ks_shift_b0, dim3(1), dim3(64), 0, s, nbe, nbd, mtol, S);
}

namespace {

template <int NX, int NY, int NZ>
void fbar_launch(hipStream_t s, const GridGeom &g, const PartF32 &P, const NodesF32 &N, const CellLists &L,
                 const NodeTables &T, const TileGeom &tg, const StepConsts &c, float *partials) {
    hipLaunchKernelGGL((ks_fbar_grad<NX, NY, NZ>), dim3(cdiv(P.n, 256)), dim3(256), 0, s, g, P, N, T, c);
    const int ng = c.cells_grid > 0 ? c.cells_grid : cell_groups(g);
    hipLaunchKernelGGL((ks_fbar_cells<NX, NY, NZ>), dim3(ng), dim3(kSeg), 0, s, g, P, L, partials);
    hipLaunchKernelGGL((ks_fbar_nodes<NX, NY, NZ>), dim3(tg.ntiles), dim3(256), 0, s, g, L, tg, (const float *)partials,
                       N, c);
}

template <int NX, int NY, int NZ>
void update_launch(hipStream_t s, const GridGeom &g, const PartF32 &P, const NodesF32 &N, const NodeTables &T,
                   const MaterialTable &mt, const StepConsts &c, uint32_t *err, const BinFuse &bf) {
    const dim3 grid(cdiv(P.n, 256)), block(256);
    if (c.fbar)
        hipLaunchKernelGGL((ks_g2p_update<NX, NY, NZ, true>), grid, block, 0, s, g, P, N, T, mt, c, err, bf);
    else
        hipLaunchKernelGGL((ks_g2p_update<NX, NY, NZ, false>), grid, block, 0, s, g, P, N, T, mt, c, err, bf);
}

}  // namespace

void launch_s32_fbar(hipStream_t s, const GridGeom &g, const PartF32 &P, const NodesF32 &N, const CellLists &L,
                     const NodeTables &T, const TileGeom &tg, const StepConsts &c, float *partials) {
    if (P.n <= 0) return;
    switch (stencil_kind(g)) {
    case 0: fbar_launch<2, 2, 2>(s, g, P, N, L, T, tg, c, partials); break;
    case 1: fbar_launch<3, 3, 3>(s, g, P, N, L, T, tg, c, partials); break;
    case 2: fbar_launch<3, 2, 3>(s, g, P, N, L, T, tg, c, partials); break;
    default: break;
    }
}

void launch_s32_g2p_update(hipStream_t s, const GridGeom &g, const PartF32 &P, const NodesF32 &N,
                           const NodeTables &T, const MaterialTable &mt, const StepConsts &c, uint32_t *err,
                           const BinFuse &bf) {
    if (P.n <= 0) return;
    switch (stencil_kind(g)) {
    case 0: update_launch<2, 2, 2>(s, g, P, N, T, mt, c, err, bf); break;
    case 1: update_launch<3, 3, 3>(s, g, P, N, T, mt, c, err, bf); break;
    case 2: update_launch<3, 2, 3>(s, g, P, N, T, mt, c, err, bf); break;
    default: break;
    }
}

}  // namespace mpm

// Device helpers shared by the FP32 step kernels (fast32.hip: linear
// tent functions, tuned; stencil32.hip: any stencil of 2 or 3 control points
// per axis, with the affine transfers): chunked LDS staging of the
// cell-sorted particle fields, compensated sums, node-tile iteration, the
// FP32 stress update and the binning counters the update kernel keeps.
#pragma once

#include "mpm_common.h"
#include "mpm_kernels.h"

namespace mpm {

namespace {

constexpr int kSeg = 64;      // cells per row segment = threads per workgroup
constexpr int kChunk = 256;   // particles staged per LDS pass

// q = a / n, rem = a % n for a < 2^24 via the float reciprocal, corrected to exact
__device__ __forceinline__ int fdivmod(int a, int n, float inv_n, int &rem) {
    const int q = (int)((float)a * inv_n);
    const int r = a - q * n;
    // selects, not branches: keeps the caller's loads in one basic block
    const bool lo = r < 0, hi = r >= n;
    rem = lo ? r + n : (hi ? r - n : r);
    return lo ? q - 1 : (hi ? q + 1 : q);
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
// B: 64-particle blocks per chunk (kBlk unless a kernel stages shorter chunks)
template <int S, int B = kBlk>
__device__ __forceinline__ int st_idx(int f, int i) { return (f * B + (i >> 6)) * S + (i & 63); }

// cache policy of the LDS DMA: 0, or kNT for a stream read once per step
// (non-temporal, the gfx9 SLC / gfx940+ NT bit), which keeps L2 for the data
// the next kernels re-read
constexpr int kNT = 2;
template <int S, int CPOL = 0, int B = kBlk>
__device__ __forceinline__ void lds_gather(const void *field_base, uint32_t slot, float *stage, int f, int r) {
    __builtin_amdgcn_global_load_lds((const uint32_t *)field_base + slot, stage + st_idx<S, B>(f, 64 * r), 4, 0, CPOL);
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
template <int B = kBlk>
__device__ __forceinline__ void load_slots(const CellLists &L, uint32_t base, int nch, uint32_t slot[B]) {
#pragma unroll
    for (int r = 0; r < B; r++) slot[r] = (64 * r + (int)threadIdx.x < nch) ? L.perm[base + 64 * r + threadIdx.x] : 0u;
}

// the same for every lane, positions clamped into [base, end): exactly kBlk
// load instructions, whatever the chunk length
__device__ __forceinline__ void load_slots_all(const CellLists &L, uint32_t base, uint32_t end, uint32_t slot[kBlk]) {
#pragma unroll
    for (int r = 0; r < kBlk; r++) slot[r] = L.perm[min(base + 64u * r + threadIdx.x, end - 1u)];
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
    // Blocks take the tiles with x slowest. Workgroups go round the 8 XCDs in
    // block order; with x fastest, a particle region one tile wide in x puts
    // every busy block on ntx / gcd(ntx, 8) of them (two for ntx = 4).
    const int nyz = tg.nty * tg.ntz;
    const int tx = (int)blockIdx.x / nyz;
    const int ryz = (int)blockIdx.x - tx * nyz;
    const int ty = ryz % tg.nty;
    const int tz = ryz / tg.nty;
    const int tile = tx + tg.ntx * (ty + tg.nty * tz);
    if (!all_tiles && L.tiles[tile] == 0u) return false;
    const int l = threadIdx.x;
    i = tx * kNodeTileX + (l % kNodeTileX);
    j = ty * kNodeTileY + (l / kNodeTileX) % kNodeTileY;
    k = tz * kNodeTileZ + l / (kNodeTileX * kNodeTileY);
    if (i >= g.nxc[0] || j >= g.nxc[1] || k >= g.nxc[2]) return false;
    n = i + g.nxc[0] * (j + g.nxc[1] * k);
    return true;
}

// Prescribed velocities of the masked components of node n (bv[c] for mask
// bit c; an unmasked component reads the table's first value, one line for
// the whole wave), loaded without a branch so that they travel with the node
// pass's run loads instead of after its sums.
__device__ __forceinline__ void node_bc(const BcDense &bc, int n, int nodec, uint32_t mk, double (&bv)[3]) {
#pragma unroll
    for (int c = 0; c < 3; c++) bv[c] = bc.mask ? bc.val[((mk >> c) & 1u) ? (size_t)c * nodec + n : 0] : 0.0;
}

__device__ __forceinline__ bool cell_nonempty(const CellLists &L, int c) { return (L.occ[c >> 5] >> (c & 31)) & 1u; }
static_assert(kSeg == kGroupLanes, "the binning's group-start bits assume one wave per group");

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

// Elastic constants, pressure, deviator, von Mises stress and yield function
// of a trial state from t.E, t.nu, t.sy, t.Hh, t.ep and t.st (the update
// kernel and the surrogate's network pass share this arithmetic, so that the
// exact return of a rejected prediction is the update kernel's own).
__device__ __forceinline__ void j2_trial_derived(J2Trial &t) {
    const float E = t.E, nu = t.nu;
    t.lmd = E * nu / ((1.f + nu) * (1.f - 2.f * nu));
    t.mu = E / (2.f * (1.f + nu));
    t.pm = (t.st[0] + t.st[1] + t.st[2]) * (1.f / 3.f);
    t.d0 = t.st[0] - t.pm;
    t.d1 = t.st[1] - t.pm;
    t.d2 = t.st[2] - t.pm;
    const float j2 = 0.5f * (t.d0 * t.d0 + t.d1 * t.d1 + t.d2 * t.d2) + t.st[3] * t.st[3] + t.st[4] * t.st[4] +
                     t.st[5] * t.st[5];
    t.q = sqrtf(3.f * j2);
    t.f = t.q - (t.sy + t.Hh * t.ep);
}

__device__ __forceinline__ void j2_trial(const double *p, const float D[9], const float s[6], float ep, J2Trial &t) {
    const float E = (float)p[0], nu = (float)p[1];
    t.E = E;
    t.nu = nu;
    t.sy = (float)p[2];
    t.Hh = (float)p[3];
    t.ep = ep;
    const float lmd = E * nu / ((1.f + nu) * (1.f - 2.f * nu));
    const float mu = E / (2.f * (1.f + nu));
    float de[6] = {D[0], D[4], D[8], D[5] + D[7], D[6] + D[2], D[1] + D[3]};
    float tr = de[0] + de[1] + de[2];
    t.st[0] = s[0] + lmd * tr + 2.f * mu * de[0];
    t.st[1] = s[1] + lmd * tr + 2.f * mu * de[1];
    t.st[2] = s[2] + lmd * tr + 2.f * mu * de[2];
    t.st[3] = s[3] + mu * de[3];
    t.st[4] = s[4] + mu * de[4];
    t.st[5] = s[5] + mu * de[5];
    j2_trial_derived(t);
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

// Tait pressure p = B ((rho / rho0)^gamma - 1), B = rho0 c^2 / gamma, rho =
// rho0 / J, as B expm1(-gamma log1p(J - 1)): no cancellation near J = 1.
// prop = {rho0, c, gamma, mu}; *Bout receives B when not null.
__device__ __forceinline__ float tait_pressure(const double *p, double J, float *Bout) {
    const float rho0 = (float)p[0], c = (float)p[1], gam = (float)p[2];
    const float B = rho0 * c * c / gam;
    if (Bout) *Bout = B;
    return B * expm1f(-gam * log1pf((float)(J - 1.0)));
}

// Jl is the determinant in the Neo-Hookean log term: J, or Jbar under the
// F-bar projection (which passes H = Fbar - I and the scaled increment; the
// other determinants stay det F, as ConstitutiveModel with det_def_grad_bar).
__device__ __forceinline__ void stress32(const MaterialDev &md, const float H[9], const float D[9], double J,
                                         double Jl, float s[6], float *eps_p, float inv_dt) {
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
        float lnj = log1pf((float)(Jl - 1.0));
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
        const float mu = (float)p[3];
        const float pr = tait_pressure(p, J, nullptr);
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

__device__ __forceinline__ void stress32(const MaterialDev &md, const float H[9], const float D[9], double J,
                                         float s[6], float *eps_p, float inv_dt) {
    stress32(md, H, D, J, J, s, eps_p, inv_dt);
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

// 13 surrogate inputs of the Tait pressure closure: J - 1 and the volumetric
// strain rate tr(dF - I) / dt (the other inputs zero)
__device__ __forceinline__ void tait_features(double J, const float D[9], float dt, float x[16]) {
    x[0] = (float)(J - 1.0);
    x[1] = (D[0] + D[4] + D[8]) / dt;
#pragma unroll
    for (int c = 2; c < 16; c++) x[c] = 0.f;
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

}  // namespace

}  // namespace mpm

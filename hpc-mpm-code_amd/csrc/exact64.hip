// FP64 kernels that follow the reference's arithmetic operation by operation.
//
// Compiled with -ffp-contract=off: no fused multiply-adds, so every product
// and sum is rounded as in the reference's x86-64 build. Node sums are
// gathered per node in the reference's order (cells ascending, particles in
// linked-list order), which makes the P2G and MUSL sums identical to the
// sequential scatter loops of var_trans_solid_explicit.cpp:35-70 and :109-127.
// Particle-side passes are one thread per particle and visit the stencil in
// MakSf's k, j, i order. IEEE division and sqrt are correctly rounded on the
// device as on the host; only libm functions (log in Neo-Hookean, cbrt/log in
// Mooney-Rivlin) may differ in the last bit.
#include "mpm_common.h"
#include "mpm_kernels.h"

namespace mpm {

namespace {

// One axis of MakSf's linear branch (shape_function.cpp:40-51) for control
// point index ixyn.
__device__ __forceinline__ void sf_linear(double xyg, double xmin, double dxi, int ixyn, double &s, double &d) {
    double xy = xmin + dxi * (double)ixyn;
    double dxnpi = xyg - xy;
    double xi = fabs(dxnpi) / dxi;
    s = 1.0 - xi;
    d = (dxnpi < 0.0) ? (1.0 / dxi) : (-1.0 / dxi);
}

// Quadratic B-spline branch (shape_function.cpp:53-104), single rank.
__device__ __forceinline__ void sf_bspline(double xyg, double xmin, double dxi, int ixyn, int nw, double &s,
                                           double &d) {
    int ncxy = ixyn;
    double xy = xmin + dxi * (double)(ixyn - 2);
    double dxnpi = xyg - xy;
    double dxh = dxnpi / dxi;
    double s1, s2, s3, d1, d2;
    if (ncxy == 0) {
        dxh -= 2.0;
        s1 = 1.0; s2 = -2.0; s3 = 1.0; d1 = 2.0; d2 = -2.0;
    } else if (ncxy == 1) {
        dxh -= 1.0;
        if (dxh <= 1.0) {
            s1 = -1.5; s2 = 2.0; s3 = 0.0; d1 = -3.0; d2 = 2.0;
        } else {
            s1 = 0.5; s2 = -2.0; s3 = 2.0; d1 = 1.0; d2 = -2.0;
        }
    } else if (ncxy == nw - 1) {
        s1 = 1.0; s2 = 0.0; s3 = 0.0; d1 = 2.0; d2 = 0.0;
    } else if (ncxy == nw - 2) {
        if (dxh <= 1.0) {
            s1 = 0.5; s2 = 0.0; s3 = 0.0; d1 = 1.0; d2 = 0.0;
        } else {
            s1 = -1.5; s2 = 4.0; s3 = -2.0; d1 = -3.0; d2 = 4.0;
        }
    } else {
        if (dxh <= 1.0) {
            s1 = 0.5; s2 = 0.0; s3 = 0.0; d1 = 1.0; d2 = 0.0;
        } else if (dxh <= 2.0) {
            s1 = -1.0; s2 = 3.0; s3 = -1.5; d1 = -2.0; d2 = 3.0;
        } else {
            s1 = 0.5; s2 = -3.0; s3 = 4.5; d1 = 1.0; d2 = -3.0;
        }
    }
    s = (s1 * dxh + s2) * dxh + s3;
    d = (d1 * dxh + d2) / dxi;
}

__device__ __forceinline__ void sf_axis(const GridGeom &g, int ii, double xyg, int ixyn, double &s, double &d) {
    if (g.idimc[ii] == 1)
        sf_linear(xyg, g.xmin[ii], g.dx[ii], ixyn, s, d);
    else
        sf_bspline(xyg, g.xmin[ii], g.dx[ii], ixyn, g.nxc[ii], s, d);
}

// Stencil entry of MakSf for control point (ix, iy, iz) and position xyg
// (tensor product of shape_function.cpp:107-110).
__device__ __forceinline__ void sf_node(const GridGeom &g, const double xyg[3], int ix, int iy, int iz, double &sfi,
                                        double ds[3]) {
    double s0, s1, s2, d0, d1, d2;
    sf_axis(g, 0, xyg[0], ix, s0, d0);
    sf_axis(g, 1, xyg[1], iy, s1, d1);
    sf_axis(g, 2, xyg[2], iz, s2, d2);
    sfi = s0 * s1 * s2;
    ds[0] = d0 * s1 * s2;
    ds[1] = s0 * d1 * s2;
    ds[2] = s0 * s1 * d2;
}

__device__ __forceinline__ void cell_ijk(const GridGeom &g, int m, int &ixe, int &iye, int &ize) {
    int nex = g.nex[0], ney = g.nex[1];
    ize = m / (nex * ney);
    iye = (m - ize * (nex * ney)) / nex;
    ixe = m - ize * (nex * ney) - iye * nex;
}

__device__ __forceinline__ void apply_vel_bc(const BcDense &bc, int n, int nodec, double facl, double &v0, double &v1,
                                             double &v2) {
    uint8_t mk = bc.mask ? bc.mask[n] : 0;
    if (mk) {
        if (mk & 1) v0 = bc.val[n] * facl;
        if (mk & 2) v1 = bc.val[nodec + n] * facl;
        if (mk & 4) v2 = bc.val[2 * nodec + n] * facl;
    }
}

// Cells adjacent to control point (i, j, k), ascending cell index; returns count.
__device__ __forceinline__ int node_cells(const GridGeom &g, int i, int j, int k, int *cells) {
    int cnt = 0;
    const int r0 = g.idimc[0], r1 = g.idimc[1], r2 = g.idimc[2];
    for (int ck = k - r2; ck <= k; ck++) {
        if (ck < 0 || ck >= g.nex[2]) continue;
        for (int cj = j - r1; cj <= j; cj++) {
            if (cj < 0 || cj >= g.nex[1]) continue;
            for (int ci = i - r0; ci <= i; ci++) {
                if (ci < 0 || ci >= g.nex[0]) continue;
                cells[cnt++] = ci + cj * g.nex[0] + ck * g.nex[0] * g.nex[1];
            }
        }
    }
    return cnt;
}

// Control-point coordinate on one axis: DefineCpPos (module/mesh.h:32-47),
// one rank.
__device__ __forceinline__ double node_coord(const GridGeom &g, int a, int idx) {
    if (g.idimc[a] == 1) return g.xmin[a] + g.dx[a] * (double)idx;
    if (idx == 0) return g.xmin[a];
    if (idx == g.nxc[a] - 1) return g.xmax[a];
    return g.xmin[a] + g.dx[a] * ((double)idx - 0.5);
}

// Velocity particle pid carries to control point (i, j, k) under TPIC / APIC
// (map_and_interpolate.h:62-78 and :107-131): v_p + M (x_a - x_p), M the
// velocity gradient (TPIC) or B D^-1 (APIC, summed over k first as there).
// The node coordinate is xyc[nid][j]; the reference's xyc[j][nid] at :69 and
// :123 reads out of bounds.
__device__ __forceinline__ void affine_vel(const GridGeom &g, const PartF64 &P, uint32_t pid, int i, int j, int k,
                                           const double xyg[3], int scheme, double va[3]) {
    const size_t ld = P.ld;
    double d[3] = {node_coord(g, 0, i) - xyg[0], node_coord(g, 1, j) - xyg[1], node_coord(g, 2, k) - xyg[2]};
    double M[9];
    for (int q = 0; q < 9; q++) M[q] = P.C[q * ld + pid];
    if (scheme == 3) {
        double Di[9], BD[9];
        for (int q = 0; q < 9; q++) Di[q] = P.Dinv[q * ld + pid];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) {
                double acc = 0.0;
                for (int q = 0; q < 3; q++) acc += M[3 * r + q] * Di[3 * q + c];
                BD[3 * r + c] = acc;
            }
        for (int q = 0; q < 9; q++) M[q] = BD[q];
    }
    for (int r = 0; r < 3; r++) {
        double t = 0.0;
        for (int c = 0; c < 3; c++) t += M[3 * r + c] * d[c];
        va[r] = va[r] + t;
    }
}

// Particle2NodeS (var_trans_solid_explicit.cpp:20-96) for one node.
__global__ void __launch_bounds__(128) k64_p2g(GridGeom g, PartF64 P, NodesF64 N, CellLists L, BcDense bc, int scheme, double fx, double fy,
                        double fz, double mtol, double facl) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nodec) return;
    const int nodec = g.nodec;
    const size_t ld = P.ld;
    int i = n % g.nxc[0];
    int j = (n / g.nxc[0]) % g.nxc[1];
    int k = n / (g.nxc[0] * g.nxc[1]);
    int cells[27];
    int nc = node_cells(g, i, j, k, cells);
    double nm = 0.0, nvof = 0.0, p0 = 0.0, p1 = 0.0, p2 = 0.0, f0 = 0.0, f1 = 0.0, f2 = 0.0;
    for (int q = 0; q < nc; q++) {
        int c = cells[q];
        uint32_t s0 = L.cstart[c], s1 = L.cend[c];
        for (uint32_t s = s0; s < s1; s++) {
            uint32_t pid = L.perm[s];
            double xyg[3] = {P.x[pid], P.x[ld + pid], P.x[2 * ld + pid]};
            double sfi, ds[3];
            sf_node(g, xyg, i, j, k, sfi, ds);
            double mp = P.m[pid], vp = P.V[pid];
            double s_0 = P.sig[pid], s_1 = P.sig[ld + pid], s_2 = P.sig[2 * ld + pid];
            double s_3 = P.sig[3 * ld + pid], s_4 = P.sig[4 * ld + pid], s_5 = P.sig[5 * ld + pid];
            double t0 = 0.0, t1 = 0.0, t2 = 0.0;
            if (P.trac) {
                t0 = P.trac[pid];
                t1 = P.trac[ld + pid];
                t2 = P.trac[2 * ld + pid];
            }
            nm += sfi * mp;
            nvof += sfi * vp;
            double va[3] = {P.v[pid], P.v[ld + pid], P.v[2 * ld + pid]};
            if (scheme >= 2) affine_vel(g, P, pid, i, j, k, xyg, scheme, va);
            p0 += sfi * mp * va[0];
            p1 += sfi * mp * va[1];
            p2 += sfi * mp * va[2];
            double fi0 = -vp * (s_0 * ds[0] + s_5 * ds[1] + s_4 * ds[2]);
            double fi1 = -vp * (s_5 * ds[0] + s_1 * ds[1] + s_3 * ds[2]);
            double fi2 = -vp * (s_4 * ds[0] + s_3 * ds[1] + s_2 * ds[2]);
            double fe0 = sfi * (mp * fx + t0);
            double fe1 = sfi * (mp * fy + t1);
            double fe2 = sfi * (mp * fz + t2);
            f0 += fi0 + fe0;
            f1 += fi1 + fe1;
            f2 += fi2 + fe2;
        }
    }
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    if (nm > mtol) {
        v0 = p0 / nm;
        v1 = p1 / nm;
        v2 = p2 / nm;
    }
    if (nvof < mtol) nvof = 0.0;
    apply_vel_bc(bc, n, nodec, facl, v0, v1, v2);
    N.nm[n] = nm;
    N.nvof[n] = nvof;
    N.np[n] = p0;
    N.np[nodec + n] = p1;
    N.np[2 * nodec + n] = p2;
    N.nf[n] = f0;
    N.nf[nodec + n] = f1;
    N.nf[2 * nodec + n] = f2;
    N.nv[n] = v0;
    N.nv[nodec + n] = v1;
    N.nv[2 * nodec + n] = v2;
}

// ------------------------------------------------------------------
// The same P2G and MUSL sums in two passes. The per-node gathers above
// evaluate each particle's weights once for every node of its stencil and
// read its fields from scattered addresses; here one thread per binned
// particle (sorted position s) forms the values a node adds for it -- the
// same expressions as above, so the same doubles -- and stores them at
// [s][corner][quantity]; a node then adds them in the same order (cells
// ascending, particles in list order) from contiguous memory. The sums are
// bit-identical to the single-pass gathers and to the reference's scatter.
constexpr int kP2GQ = 8;   // nm, nvof, p (3), f (3)
constexpr int kMuslQ = 3;  // p (3)

__device__ __forceinline__ int stencil_corners(const GridGeom &g) {
    return (g.idimc[0] + 1) * (g.idimc[1] + 1) * (g.idimc[2] + 1);
}

template <bool MUSL>
__global__ void __launch_bounds__(128) k64_contrib(GridGeom g, PartF64 P, CellLists L, int scheme, double fx, double fy, double fz,
                            double *__restrict__ out) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t total = L.astart[L.ctr->n_active];
    if (s >= total) return;
    constexpr int Q = MUSL ? kMuslQ : kP2GQ;
    const uint32_t pid = L.perm[s];
    const size_t ld = P.ld;
    const int nen = stencil_corners(g);
    int ixe, iye, ize;
    cell_ijk(g, (int)P.key[pid], ixe, iye, ize);
    double xyg[3] = {P.x[pid], P.x[ld + pid], P.x[2 * ld + pid]};
    const double mp = P.m[pid];
    double vp = 0.0, s_0 = 0.0, s_1 = 0.0, s_2 = 0.0, s_3 = 0.0, s_4 = 0.0, s_5 = 0.0, t0 = 0.0, t1 = 0.0, t2 = 0.0;
    if (!MUSL) {
        vp = P.V[pid];
        s_0 = P.sig[pid];
        s_1 = P.sig[ld + pid];
        s_2 = P.sig[2 * ld + pid];
        s_3 = P.sig[3 * ld + pid];
        s_4 = P.sig[4 * ld + pid];
        s_5 = P.sig[5 * ld + pid];
        if (P.trac) {
            t0 = P.trac[pid];
            t1 = P.trac[ld + pid];
            t2 = P.trac[2 * ld + pid];
        }
    }
    double *o = out + (size_t)s * nen * Q;
    int q = 0;
    for (int dk = 0; dk <= g.idimc[2]; dk++)
        for (int dj = 0; dj <= g.idimc[1]; dj++)
            for (int di = 0; di <= g.idimc[0]; di++, q++) {
                const int i = ixe + di, j = iye + dj, k = ize + dk;
                double sfi, ds[3];
                sf_node(g, xyg, i, j, k, sfi, ds);
                double va[3] = {P.v[pid], P.v[ld + pid], P.v[2 * ld + pid]};
                if (scheme >= 2) affine_vel(g, P, pid, i, j, k, xyg, scheme, va);
                double *c = o + q * Q;
                if (MUSL) {
                    c[0] = sfi * mp * va[0];
                    c[1] = sfi * mp * va[1];
                    c[2] = sfi * mp * va[2];
                } else {
                    c[0] = sfi * mp;
                    c[1] = sfi * vp;
                    c[2] = sfi * mp * va[0];
                    c[3] = sfi * mp * va[1];
                    c[4] = sfi * mp * va[2];
                    const double fi0 = -vp * (s_0 * ds[0] + s_5 * ds[1] + s_4 * ds[2]);
                    const double fi1 = -vp * (s_5 * ds[0] + s_1 * ds[1] + s_3 * ds[2]);
                    const double fi2 = -vp * (s_4 * ds[0] + s_3 * ds[1] + s_2 * ds[2]);
                    const double fe0 = sfi * (mp * fx + t0);
                    const double fe1 = sfi * (mp * fy + t1);
                    const double fe2 = sfi * (mp * fz + t2);
                    c[5] = fi0 + fe0;
                    c[6] = fi1 + fe1;
                    c[7] = fi2 + fe2;
                }
            }
}

// Adds the stored values of node n in the reference's order: cells around the
// node ascending (k, j, i), each cell's particles in list order.
template <int Q, typename F>
__device__ __forceinline__ void gather_contrib(const GridGeom &g, const CellLists &L, const double *__restrict__ in,
                                               int i, int j, int k, F add) {
    const int r0 = g.idimc[0], r1 = g.idimc[1], r2 = g.idimc[2];
    const int nen = (r0 + 1) * (r1 + 1) * (r2 + 1);
    for (int ck = k - r2; ck <= k; ck++) {
        if (ck < 0 || ck >= g.nex[2]) continue;
        for (int cj = j - r1; cj <= j; cj++) {
            if (cj < 0 || cj >= g.nex[1]) continue;
            for (int ci = i - r0; ci <= i; ci++) {
                if (ci < 0 || ci >= g.nex[0]) continue;
                const int c = ci + cj * g.nex[0] + ck * g.nex[0] * g.nex[1];
                const int corner = (i - ci) + (r0 + 1) * ((j - cj) + (r1 + 1) * (k - ck));
                const uint32_t s0 = L.cstart[c], s1 = L.cend[c];
                // four particles' values loaded together, then added in order
                uint32_t s = s0;
                for (; s + 4 <= s1; s += 4) {
                    double v[4][Q];
#pragma unroll
                    for (int u = 0; u < 4; u++)
#pragma unroll
                        for (int q = 0; q < Q; q++) v[u][q] = in[((size_t)(s + u) * nen + corner) * Q + q];
#pragma unroll
                    for (int u = 0; u < 4; u++) add(v[u]);
                }
                for (; s < s1; s++) add(in + ((size_t)s * nen + corner) * Q);
            }
        }
    }
}

__global__ void __launch_bounds__(128) k64_p2g_gather(GridGeom g, NodesF64 N, CellLists L, BcDense bc, double mtol, double facl,
                               const double *__restrict__ in) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nodec) return;
    const int nodec = g.nodec;
    const int i = n % g.nxc[0], j = (n / g.nxc[0]) % g.nxc[1], k = n / (g.nxc[0] * g.nxc[1]);
    double nm = 0.0, nvof = 0.0, p0 = 0.0, p1 = 0.0, p2 = 0.0, f0 = 0.0, f1 = 0.0, f2 = 0.0;
    gather_contrib<kP2GQ>(g, L, in, i, j, k, [&](const double *c) {
        nm += c[0];
        nvof += c[1];
        p0 += c[2];
        p1 += c[3];
        p2 += c[4];
        f0 += c[5];
        f1 += c[6];
        f2 += c[7];
    });
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    if (nm > mtol) {
        v0 = p0 / nm;
        v1 = p1 / nm;
        v2 = p2 / nm;
    }
    if (nvof < mtol) nvof = 0.0;
    apply_vel_bc(bc, n, nodec, facl, v0, v1, v2);
    N.nm[n] = nm;
    N.nvof[n] = nvof;
    N.np[n] = p0;
    N.np[nodec + n] = p1;
    N.np[2 * nodec + n] = p2;
    N.nf[n] = f0;
    N.nf[nodec + n] = f1;
    N.nf[2 * nodec + n] = f2;
    N.nv[n] = v0;
    N.nv[nodec + n] = v1;
    N.nv[2 * nodec + n] = v2;
}

__global__ void __launch_bounds__(128) k64_musl_gather(GridGeom g, NodesF64 N, CellLists L, BcDense bc, double mtol, double facl,
                                const double *__restrict__ in) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nodec) return;
    const int nodec = g.nodec;
    const int i = n % g.nxc[0], j = (n / g.nxc[0]) % g.nxc[1], k = n / (g.nxc[0] * g.nxc[1]);
    double p0 = 0.0, p1 = 0.0, p2 = 0.0;
    gather_contrib<kMuslQ>(g, L, in, i, j, k, [&](const double *c) {
        p0 += c[0];
        p1 += c[1];
        p2 += c[2];
    });
    const double nm = N.nm[n];
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    if (nm > mtol) {
        v0 = p0 / nm;
        v1 = p1 / nm;
        v2 = p2 / nm;
    }
    apply_vel_bc(bc, n, nodec, facl, v0, v1, v2);
    N.np[n] = p0;
    N.np[nodec + n] = p1;
    N.np[2 * nodec + n] = p2;
    N.nv[n] = v0;
    N.nv[nodec + n] = v1;
    N.nv[2 * nodec + n] = v2;
}

// SolveSolid (solve_solid_explicit.cpp:8-31)
__global__ void k64_solve_grid(GridGeom g, NodesF64 N, BcDense bc, double mtol) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nodec) return;
    const int nodec = g.nodec;
    double nm = N.nm[n];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    if (nm > mtol) {
        a0 = N.nf[n] / nm;
        a1 = N.nf[nodec + n] / nm;
        a2 = N.nf[2 * nodec + n] / nm;
    }
    uint8_t mk = bc.mask ? bc.mask[n] : 0;
    if (mk & 1) a0 = 0.0;
    if (mk & 2) a1 = 0.0;
    if (mk & 4) a2 = 0.0;
    N.na[n] = a0;
    N.na[nodec + n] = a1;
    N.na[2 * nodec + n] = a2;
}

// var_trans_solid_explicit.cpp:194-200
__global__ void k64_vel_plus_dt_acc(int n3, double *nv, const double *na, double dt) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < n3) nv[n] += dt * na[n];
}

// var_trans_solid_explicit.cpp:202-240 (PIC / FLIP velocity G2P)
__global__ void __launch_bounds__(128) k64_g2p_vel(GridGeom g, PartF64 P, NodesF64 N, int scheme, double dt) {
    int pid = blockIdx.x * blockDim.x + threadIdx.x;
    if (pid >= P.n) return;
    uint32_t m = P.key[pid];
    if (m >= (uint32_t)g.nelem) return;
    const size_t ld = P.ld;
    const int nodec = g.nodec;
    int ixe, iye, ize;
    cell_ijk(g, (int)m, ixe, iye, ize);
    double xyg[3] = {P.x[pid], P.x[ld + pid], P.x[2 * ld + pid]};
    double v0, v1, v2;
    if (scheme != 1) {
        v0 = 0.0;
        v1 = 0.0;
        v2 = 0.0;
    } else {
        v0 = P.v[pid];
        v1 = P.v[ld + pid];
        v2 = P.v[2 * ld + pid];
    }
    // TPIC velocity gradient / APIC B and D, zeroed at :203-209
    double Cm[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, AD[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k <= g.idimc[2]; k++)
        for (int j = 0; j <= g.idimc[1]; j++)
            for (int i = 0; i <= g.idimc[0]; i++) {
                int ix = ixe + i, iy = iye + j, iz = ize + k;
                int nid = ix + g.nxc[0] * iy + g.nxc[0] * g.nxc[1] * iz;
                double sfi, ds[3];
                sf_node(g, xyg, ix, iy, iz, sfi, ds);
                const double nv[3] = {N.nv[nid], N.nv[nodec + nid], N.nv[2 * nodec + nid]};
                if (scheme != 1) {
                    v0 += sfi * nv[0];
                    v1 += sfi * nv[1];
                    v2 += sfi * nv[2];
                } else {
                    v0 += sfi * dt * N.na[nid];
                    v1 += sfi * dt * N.na[nodec + nid];
                    v2 += sfi * dt * N.na[2 * nodec + nid];
                }
                if (scheme == 2) {
                    // TPIC::VarGradG2P (map_and_interpolate.h:81-99)
                    for (int r = 0; r < 3; r++)
                        for (int c = 0; c < 3; c++) Cm[3 * r + c] += ds[c] * nv[r];
                } else if (scheme == 3) {
                    // APIC::VarBmatG2P :134-148, DmatG2P :151-164
                    double d[3] = {node_coord(g, 0, ix) - xyg[0], node_coord(g, 1, iy) - xyg[1],
                                   node_coord(g, 2, iz) - xyg[2]};
                    for (int r = 0; r < 3; r++)
                        for (int c = 0; c < 3; c++) Cm[3 * r + c] += sfi * nv[r] * d[c];
                    for (int r = 0; r < 3; r++)
                        for (int c = 0; c < 3; c++) AD[3 * r + c] += sfi * d[r] * d[c];
                }
            }
    P.v[pid] = v0;
    P.v[ld + pid] = v1;
    P.v[2 * ld + pid] = v2;
    if (scheme >= 2)
        for (int q = 0; q < 9; q++) P.C[q * ld + pid] = Cm[q];
    if (scheme == 3) {
        // APIC::InvDmatG2P :167-175 with InvMat3 (cal_mat.h:48-66)
        double det = AD[0] * (AD[4] * AD[8] - AD[5] * AD[7]) + AD[1] * (AD[5] * AD[6] - AD[3] * AD[8]) +
                     AD[2] * (AD[3] * AD[7] - AD[4] * AD[6]);
        det = 1.0 / det;
        double o[9];
        o[0] = (AD[4] * AD[8] - AD[5] * AD[7]) * det;
        o[1] = (AD[2] * AD[7] - AD[1] * AD[8]) * det;
        o[2] = (AD[1] * AD[5] - AD[2] * AD[4]) * det;
        o[3] = (AD[5] * AD[6] - AD[3] * AD[8]) * det;
        o[4] = (AD[0] * AD[8] - AD[2] * AD[6]) * det;
        o[5] = (AD[2] * AD[3] - AD[0] * AD[5]) * det;
        o[6] = (AD[3] * AD[7] - AD[4] * AD[6]) * det;
        o[7] = (AD[1] * AD[6] - AD[0] * AD[7]) * det;
        o[8] = (AD[0] * AD[4] - AD[1] * AD[3]) * det;
        for (int q = 0; q < 9; q++) P.Dinv[q * ld + pid] = o[q];
    }
}

// MUSL (var_trans_solid_explicit.cpp:101-148) for one node
__global__ void __launch_bounds__(128) k64_musl(GridGeom g, PartF64 P, NodesF64 N, CellLists L, BcDense bc, int scheme, double mtol,
                         double facl) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nodec) return;
    const int nodec = g.nodec;
    const size_t ld = P.ld;
    int i = n % g.nxc[0];
    int j = (n / g.nxc[0]) % g.nxc[1];
    int k = n / (g.nxc[0] * g.nxc[1]);
    int cells[27];
    int nc = node_cells(g, i, j, k, cells);
    double p0 = 0.0, p1 = 0.0, p2 = 0.0;
    for (int q = 0; q < nc; q++) {
        int c = cells[q];
        uint32_t s0 = L.cstart[c], s1 = L.cend[c];
        for (uint32_t s = s0; s < s1; s++) {
            uint32_t pid = L.perm[s];
            double xyg[3] = {P.x[pid], P.x[ld + pid], P.x[2 * ld + pid]};
            double sfi, ds[3];
            sf_node(g, xyg, i, j, k, sfi, ds);
            double mp = P.m[pid];
            double va[3] = {P.v[pid], P.v[ld + pid], P.v[2 * ld + pid]};
            if (scheme >= 2) affine_vel(g, P, pid, i, j, k, xyg, scheme, va);
            p0 += sfi * mp * va[0];
            p1 += sfi * mp * va[1];
            p2 += sfi * mp * va[2];
        }
    }
    double nm = N.nm[n];
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    if (nm > mtol) {
        v0 = p0 / nm;
        v1 = p1 / nm;
        v2 = p2 / nm;
    }
    apply_vel_bc(bc, n, nodec, facl, v0, v1, v2);
    N.np[n] = p0;
    N.np[nodec + n] = p1;
    N.np[2 * nodec + n] = p2;
    N.nv[n] = v0;
    N.nv[nodec + n] = v1;
    N.nv[2 * nodec + n] = v2;
}

__device__ __forceinline__ double det3(const double a[9]) {
    return a[0] * (a[4] * a[8] - a[5] * a[7]) + a[1] * (a[5] * a[6] - a[3] * a[8]) +
           a[2] * (a[3] * a[7] - a[4] * a[6]);
}

__device__ __forceinline__ void store_voigt(double *sig, size_t ld, int pid, const double s[9]) {
    sig[pid] = s[0];
    sig[ld + pid] = s[4];
    sig[2 * ld + pid] = s[8];
    sig[3 * ld + pid] = s[5];
    sig[4 * ld + pid] = s[2];
    sig[5 * ld + pid] = s[1];
}

__device__ __forceinline__ void bmat(const double F[9], double be[9]) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            be[3 * i + j] = F[3 * i] * F[3 * j] + F[3 * i + 1] * F[3 * j + 1] + F[3 * i + 2] * F[3 * j + 2];
}

// ConstitutiveModel::UpdateStress (constitutive_model.cpp:287-305), FP64,
// same expression order as the reference. J2 has no reference counterpart.
__device__ void update_stress64(const MaterialDev &md, double *sig, size_t ld, int pid, const double F[9],
                                const double dF[9], double jac, double jac_bar, double *eps_p, double dt) {
    const double *p = md.prop;
    switch (md.model) {
    case -1:
        for (int q = 0; q < 6; q++) sig[q * ld + pid] = 0.0;
        break;
    case 0: {  // MatLinElast :16-60
        double EE = p[0], nu = p[1];
        double C0 = (EE * (1.0 - nu)) / ((1.0 - 2.0 * nu) * (1.0 + nu));
        double C1 = (EE * nu) / ((1.0 - 2.0 * nu) * (1.0 + nu));
        double C2 = EE / (2.0 * (1.0 + nu));
        double e0 = dF[0] - 1.0, e1 = dF[4] - 1.0, e2 = dF[8] - 1.0;
        double e3 = dF[5] + dF[7], e4 = dF[6] + dF[2], e5 = dF[1] + dF[3];
        double d0 = C0 * e0 + C1 * e1 + C1 * e2;
        double d1 = C1 * e0 + C0 * e1 + C1 * e2;
        double d2 = C1 * e0 + C1 * e1 + C0 * e2;
        sig[pid] += d0;
        sig[ld + pid] += d1;
        sig[2 * ld + pid] += d2;
        sig[3 * ld + pid] += C2 * e3;
        sig[4 * ld + pid] += C2 * e4;
        sig[5 * ld + pid] += C2 * e5;
        break;
    }
    case 1: {  // MatNeoHookean :62-120
        double yng = p[0], nu = p[1];
        double lmd = nu * yng / ((1.0 + nu) * (1.0 - 2.0 * nu));
        double mu = yng / (2.0 * (1.0 + nu));
        double lnjac = log(jac_bar);
        double be[9], s[9];
        bmat(F, be);
        for (int q = 0; q < 9; q++) {
            double eye = (q == 0 || q == 4 || q == 8) ? 1.0 : 0.0;
            s[q] = (mu * (be[q] - eye) + lmd * lnjac * eye) / jac;
        }
        store_voigt(sig, ld, pid, s);
        break;
    }
    case 2: {  // MatStVk :122-185
        double yng = p[0], nu = p[1];
        double lmd = nu * yng / ((1.0 + nu) * (1.0 - 2.0 * nu));
        double mu = yng / (2.0 * (1.0 + nu));
        double be[9], be2[9], s[9];
        bmat(F, be);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                be2[3 * i + j] = be[3 * i] * be[j] + be[3 * i + 1] * be[3 + j] + be[3 * i + 2] * be[6 + j];
        double tr_be = 0.0;
        for (int i = 0; i < 3; i++) tr_be += be[4 * i];
        double tr_gls = 0.5 * (tr_be - 3.0);
        for (int q = 0; q < 9; q++) s[q] = (mu * be2[q] + (lmd * tr_gls - mu) * be[q]) / jac;
        store_voigt(sig, ld, pid, s);
        break;
    }
    case 3: {  // MatMooneyRivlin :187-285 (pow(I1,2) written as I1*I1)
        double c1 = p[0], c2 = p[1], bulk = p[2];
        double Fi[9], be[9], be2[9], s[9];
        double cj = cbrt(jac);
        for (int q = 0; q < 9; q++) Fi[q] = 1.0 / cj * F[q];
        bmat(Fi, be);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                be2[3 * i + j] = be[3 * i] * be[j] + be[3 * i + 1] * be[3 + j] + be[3 * i + 2] * be[6 + j];
        double I1 = 0.0, trb2 = 0.0;
        for (int i = 0; i < 3; i++) I1 += be[4 * i];
        for (int i = 0; i < 3; i++) trb2 += be2[4 * i];
        double I2 = 0.5 * (I1 * I1 - trb2);
        double hp = bulk * log(jac) / jac;
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++) s[3 * i + j] = 2.0 * ((c1 + c2 * I1) * be[3 * i + j] - c2 * be2[3 * i + j]) / jac;
            s[4 * i] += hp - 2.0 / (3.0 * jac) * (c1 * I1 + 2.0 * c2 * I2);
        }
        store_voigt(sig, ld, pid, s);
        break;
    }
    case 20: {  // Tait equation of state + Newtonian viscosity (no reference counterpart)
        const double rho0 = p[0], c = p[1], gam = p[2], mu = p[3];
        const double B = rho0 * c * c / gam;
        const double pr = B * expm1(-gam * log1p(jac - 1.0));
        const double idt = 1.0 / dt;
        double l[9];
        for (int q = 0; q < 9; q++) l[q] = (dF[q] - ((q == 0 || q == 4 || q == 8) ? 1.0 : 0.0)) * idt;
        const double tr = (l[0] + l[4] + l[8]) / 3.0;
        sig[pid] = -pr + 2.0 * mu * (l[0] - tr);
        sig[ld + pid] = -pr + 2.0 * mu * (l[4] - tr);
        sig[2 * ld + pid] = -pr + 2.0 * mu * (l[8] - tr);
        sig[3 * ld + pid] = mu * (l[5] + l[7]);
        sig[4 * ld + pid] = mu * (l[6] + l[2]);
        sig[5 * ld + pid] = mu * (l[1] + l[3]);
        break;
    }
    case 10: {  // J2 radial return, small-increment form (Simo & Hughes 1998, Box 3.2)
        double E = p[0], nu = p[1], sy = p[2], Hh = p[3];
        double lmd = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
        double mu = E / (2.0 * (1.0 + nu));
        double de[6] = {dF[0] - 1.0, dF[4] - 1.0, dF[8] - 1.0, dF[5] + dF[7], dF[6] + dF[2], dF[1] + dF[3]};
        double tr = de[0] + de[1] + de[2];
        double st[6];
        st[0] = sig[pid] + lmd * tr + 2.0 * mu * de[0];
        st[1] = sig[ld + pid] + lmd * tr + 2.0 * mu * de[1];
        st[2] = sig[2 * ld + pid] + lmd * tr + 2.0 * mu * de[2];
        st[3] = sig[3 * ld + pid] + mu * de[3];
        st[4] = sig[4 * ld + pid] + mu * de[4];
        st[5] = sig[5 * ld + pid] + mu * de[5];
        double pm = (st[0] + st[1] + st[2]) / 3.0;
        double d0 = st[0] - pm, d1 = st[1] - pm, d2 = st[2] - pm;
        double j2 = 0.5 * (d0 * d0 + d1 * d1 + d2 * d2) + st[3] * st[3] + st[4] * st[4] + st[5] * st[5];
        double q = sqrt(3.0 * j2);
        double ep = eps_p ? eps_p[pid] : 0.0;
        double f = q - (sy + Hh * ep);
        if (f <= 0.0) {
            for (int c = 0; c < 6; c++) sig[c * ld + pid] = st[c];
        } else {
            double dg = f / (3.0 * mu + Hh);
            double scale = 1.0 - 3.0 * mu * dg / q;
            sig[pid] = pm + d0 * scale;
            sig[ld + pid] = pm + d1 * scale;
            sig[2 * ld + pid] = pm + d2 * scale;
            sig[3 * ld + pid] = st[3] * scale;
            sig[4 * ld + pid] = st[4] * scale;
            sig[5 * ld + pid] = st[5] * scale;
            if (eps_p) eps_p[pid] = ep + dg;
        }
        break;
    }
    default: break;
    }
}

// Deformation gradient, coordinates, volume and stress
// (var_trans_solid_explicit.cpp:244-296 and :150-185), one thread per particle.
__global__ void __launch_bounds__(128) k64_g2p_update(GridGeom g, PartF64 P, NodesF64 N, MaterialTable mt, double dt, uint32_t *err,
                               const double *__restrict__ geup, const double *__restrict__ b0p, int fbar) {
    int pid = blockIdx.x * blockDim.x + threadIdx.x;
    if (pid >= P.n) return;
    uint32_t m = P.key[pid];
    if (m >= (uint32_t)g.nelem) return;
    const size_t ld = P.ld;
    const int nodec = g.nodec;
    int ixe, iye, ize;
    cell_ijk(g, (int)m, ixe, iye, ize);
    double xyg[3] = {P.x[pid], P.x[ld + pid], P.x[2 * ld + pid]};
    double dd[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double dx0 = xyg[0], dx1 = xyg[1], dx2 = xyg[2];
    for (int k = 0; k <= g.idimc[2]; k++)
        for (int j = 0; j <= g.idimc[1]; j++)
            for (int i = 0; i <= g.idimc[0]; i++) {
                int ix = ixe + i, iy = iye + j, iz = ize + k;
                int nid = ix + g.nxc[0] * iy + g.nxc[0] * g.nxc[1] * iz;
                double sfi, ds[3];
                sf_node(g, xyg, ix, iy, iz, sfi, ds);
                double vu = N.nv[nid], vv = N.nv[nodec + nid], vw = N.nv[2 * nodec + nid];
                dd[0] += ds[0] * vu * dt;
                dd[1] += ds[1] * vu * dt;
                dd[2] += ds[2] * vu * dt;
                dd[3] += ds[0] * vv * dt;
                dd[4] += ds[1] * vv * dt;
                dd[5] += ds[2] * vv * dt;
                dd[6] += ds[0] * vw * dt;
                dd[7] += ds[1] * vw * dt;
                dd[8] += ds[2] * vw * dt;
                dx0 += sfi * dt * vu;
                dx1 += sfi * dt * vv;
                dx2 += sfi * dt * vw;
            }
    dd[0] += 1.0;
    dd[4] += 1.0;
    dd[8] += 1.0;
    double Fo[9], F[9];
    for (int q = 0; q < 9; q++) Fo[q] = P.F[q * ld + pid];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double acc = 0.0;
            for (int kk = 0; kk < 3; kk++) acc += dd[3 * i + kk] * Fo[3 * kk + j];
            F[3 * i + j] = acc;
        }
    for (int q = 0; q < 9; q++) {
        P.F[q * ld + pid] = F[q];
        P.ddg[q * ld + pid] = dd[q];
    }
    double J = det3(F);
    P.J[pid] = J;
    if (fbar) {  // F-bar: positions, volume and stress follow the projection (k64_fbar_finish)
        P.detdd[pid] = det3(dd);
        return;
    }
    if (geup) {  // shifting correction after the position update (var_trans_solid_explicit.cpp:176)
        const double b0 = *b0p;
        dx0 += -b0 * geup[pid];
        dx1 += -b0 * geup[ld + pid];
        dx2 += -b0 * geup[2 * ld + pid];
    }
    P.x[pid] = dx0;
    P.x[ld + pid] = dx1;
    P.x[2 * ld + pid] = dx2;
    P.V[pid] = J * P.V0[pid];
    int mat = P.matid ? P.matid[pid] : 0;
    update_stress64(mt.m[mat], P.sig, ld, pid, F, dd, J, J, P.eps_p, dt);
    if (!(J > 0.0)) atomicOr(err, kErrJacobian);
    if (!isfinite(dx0) || !isfinite(dx1) || !isfinite(dx2)) atomicOr(err, kErrNonFinite);
}

// cal_def_grad_bar (fbar_projection.cpp:18-48) for one node: sum of
// N (Jbar_p det dF_p) V_p in the reference's order, divided by nvof.
__global__ void __launch_bounds__(128) k64_fbar_nodes(GridGeom g, PartF64 P, NodesF64 N, CellLists L, double mtol, double *njbar) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nodec) return;
    const size_t ld = P.ld;
    int i = n % g.nxc[0];
    int j = (n / g.nxc[0]) % g.nxc[1];
    int k = n / (g.nxc[0] * g.nxc[1]);
    int cells[27];
    int nc = node_cells(g, i, j, k, cells);
    double acc = 0.0;
    for (int q = 0; q < nc; q++) {
        int c = cells[q];
        for (uint32_t s = L.cstart[c]; s < L.cend[c]; s++) {
            uint32_t pid = L.perm[s];
            double xyg[3] = {P.x[pid], P.x[ld + pid], P.x[2 * ld + pid]};
            double sfi, ds[3];
            sf_node(g, xyg, i, j, k, sfi, ds);
            double jp = P.Jbar[pid] * P.detdd[pid];
            acc += sfi * jp * P.V[pid];
        }
    }
    double nv = N.nvof[n];
    njbar[n] = (nv > mtol) ? acc / nv : 0.0;
}

// cal_def_grad_bar (fbar_projection.cpp:50-91) and then
// update_point_coord_and_stress (var_trans_solid_explicit.cpp:150-185) with
// Fbar, Jbar and the scaled increment, one thread per particle.
__global__ void __launch_bounds__(128) k64_fbar_finish(GridGeom g, PartF64 P, NodesF64 N, MaterialTable mt, double dt, uint32_t *err,
                                const double *__restrict__ njbar, const double *__restrict__ geup,
                                const double *__restrict__ b0p) {
    int pid = blockIdx.x * blockDim.x + threadIdx.x;
    if (pid >= P.n) return;
    uint32_t m = P.key[pid];
    if (m >= (uint32_t)g.nelem) return;
    const size_t ld = P.ld;
    const int nodec = g.nodec;
    int ixe, iye, ize;
    cell_ijk(g, (int)m, ixe, iye, ize);
    double xyg[3] = {P.x[pid], P.x[ld + pid], P.x[2 * ld + pid]};
    double jbar = 0.0;
    double dx0 = xyg[0], dx1 = xyg[1], dx2 = xyg[2];
    for (int k = 0; k <= g.idimc[2]; k++)
        for (int j = 0; j <= g.idimc[1]; j++)
            for (int i = 0; i <= g.idimc[0]; i++) {
                int ix = ixe + i, iy = iye + j, iz = ize + k;
                int nid = ix + g.nxc[0] * iy + g.nxc[0] * g.nxc[1] * iz;
                double sfi, ds[3];
                sf_node(g, xyg, ix, iy, iz, sfi, ds);
                jbar += sfi * njbar[nid];
                dx0 += sfi * dt * N.nv[nid];
                dx1 += sfi * dt * N.nv[nodec + nid];
                dx2 += sfi * dt * N.nv[2 * nodec + nid];
            }
    const double Jp = P.Jbar[pid] * P.detdd[pid];
    double dd[9], Fb[9], Fn[9];
    for (int q = 0; q < 9; q++) {
        dd[q] = P.ddg[q * ld + pid];
        dd[q] *= cbrt((jbar / Jp));
        Fb[q] = P.Fbar[q * ld + pid];
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double acc = 0.0;
            for (int kk = 0; kk < 3; kk++) acc += dd[3 * i + kk] * Fb[3 * kk + j];
            Fn[3 * i + j] = acc;
        }
    for (int q = 0; q < 9; q++) {
        P.Fbar[q * ld + pid] = Fn[q];
        P.ddg[q * ld + pid] = dd[q];
    }
    P.Jbar[pid] = jbar;
    if (geup) {
        const double b0 = *b0p;
        dx0 += -b0 * geup[pid];
        dx1 += -b0 * geup[ld + pid];
        dx2 += -b0 * geup[2 * ld + pid];
    }
    P.x[pid] = dx0;
    P.x[ld + pid] = dx1;
    P.x[2 * ld + pid] = dx2;
    const double J = P.J[pid];
    P.V[pid] = J * P.V0[pid];
    int mat = P.matid ? P.matid[pid] : 0;
    update_stress64(mt.m[mat], P.sig, ld, pid, Fn, dd, J, jbar, P.eps_p, dt);
    if (!(J > 0.0) || !(jbar > 0.0)) atomicOr(err, kErrJacobian);
    if (!isfinite(dx0) || !isfinite(dx1) || !isfinite(dx2)) atomicOr(err, kErrNonFinite);
}

__global__ void __launch_bounds__(128) k64_maksf(GridGeom g, int n, const int32_t *cell, const double *xyg, int32_t *nc, double *sf,
                          double *dsf) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int nen = (g.idimc[0] + 1) * (g.idimc[1] + 1) * (g.idimc[2] + 1);
    int ixe, iye, ize;
    cell_ijk(g, cell[p], ixe, iye, ize);
    double x[3] = {xyg[3 * p], xyg[3 * p + 1], xyg[3 * p + 2]};
    int id = 0;
    for (int k = 0; k <= g.idimc[2]; k++)
        for (int j = 0; j <= g.idimc[1]; j++)
            for (int i = 0; i <= g.idimc[0]; i++) {
                int ix = ixe + i, iy = iye + j, iz = ize + k;
                size_t o = (size_t)p * nen + id;
                nc[o] = ix + g.nxc[0] * iy + g.nxc[0] * g.nxc[1] * iz;
                double s, ds[3];
                sf_node(g, x, ix, iy, iz, s, ds);
                sf[o] = s;
                dsf[3 * o] = ds[0];
                dsf[3 * o + 1] = ds[1];
                dsf[3 * o + 2] = ds[2];
                id++;
            }
}

// ------------------------------------------------------------------------
// Implicit-path transfers: the P2G and G2P halves of the reference's
// implicit solid step (its PETSc solve is the caller's). PIC and FLIP.

// SolidMaterialPoint::Particle2Node (solid/implicit/var_trans_solid_implicit.cpp:20-80)
// for one node: mass, volume, momentum (VelP2G) and m a (AccelP2G,
// map_and_interpolate.cpp:12-49) in the reference's order; velocity and
// acceleration where the mass exceeds mtol; velocity BCs set, acceleration
// BCs zeroed.
__global__ void __launch_bounds__(256) k64_implicit_p2g(GridGeom g, PartF64 P, NodesF64 N, CellLists L, BcDense bc, double mtol,
                                 double facl) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nodec) return;
    const int nodec = g.nodec;
    const size_t ld = P.ld;
    int i = n % g.nxc[0];
    int j = (n / g.nxc[0]) % g.nxc[1];
    int k = n / (g.nxc[0] * g.nxc[1]);
    int cells[27];
    int nc = node_cells(g, i, j, k, cells);
    double nm = 0.0, nvof = 0.0, p0 = 0.0, p1 = 0.0, p2 = 0.0, f0 = 0.0, f1 = 0.0, f2 = 0.0;
    for (int q = 0; q < nc; q++) {
        int c = cells[q];
        uint32_t s0 = L.cstart[c], s1 = L.cend[c];
        for (uint32_t s = s0; s < s1; s++) {
            uint32_t pid = L.perm[s];
            double xyg[3] = {P.x[pid], P.x[ld + pid], P.x[2 * ld + pid]};
            double sfi, ds[3];
            sf_node(g, xyg, i, j, k, sfi, ds);
            double mp = P.m[pid], vp = P.V[pid];
            nm += sfi * mp;
            nvof += sfi * vp;
            p0 += sfi * mp * P.v[pid];
            p1 += sfi * mp * P.v[ld + pid];
            p2 += sfi * mp * P.v[2 * ld + pid];
            f0 += sfi * mp * P.a[pid];
            f1 += sfi * mp * P.a[ld + pid];
            f2 += sfi * mp * P.a[2 * ld + pid];
        }
    }
    double v0 = 0.0, v1 = 0.0, v2 = 0.0, a0 = 0.0, a1 = 0.0, a2 = 0.0;
    if (nm > mtol) {
        v0 = p0 / nm;
        v1 = p1 / nm;
        v2 = p2 / nm;
        a0 = f0 / nm;
        a1 = f1 / nm;
        a2 = f2 / nm;
    }
    apply_vel_bc(bc, n, nodec, facl, v0, v1, v2);
    uint8_t mk = bc.mask ? bc.mask[n] : 0;
    if (mk & 1) a0 = 0.0;
    if (mk & 2) a1 = 0.0;
    if (mk & 4) a2 = 0.0;
    N.nm[n] = nm;
    N.nvof[n] = nvof;
    N.np[n] = p0;
    N.np[nodec + n] = p1;
    N.np[2 * nodec + n] = p2;
    N.nf[n] = f0;
    N.nf[nodec + n] = f1;
    N.nf[2 * nodec + n] = f2;
    N.nv[n] = v0;
    N.nv[nodec + n] = v1;
    N.nv[2 * nodec + n] = v2;
    N.na[n] = a0;
    N.na[nodec + n] = a1;
    N.na[2 * nodec + n] = a2;
}

// PredictNewmarkBetaVelAndAccel + CommitNodalKinematics
// (nonlinear_time_integration.cpp:59-77), every component of every node.
__global__ void k64_newmark_nodes(int n3, const double *__restrict__ ndispl, double *nv, double *na, NewmarkF64 nb) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n3) return;
    const double u = ndispl[n], v = nv[n], a = na[n];
    const double vk = nb.c[0] * u - nb.c[1] * v - nb.c[2] * a;
    const double ak = nb.c[3] * u - nb.c[4] * v - nb.c[5] * a;
    nv[n] = vk;
    na[n] = ak;
}

// Node2Particle (var_trans_solid_implicit.cpp:106-167) after the node update,
// one thread per particle: dF = I + sum dN (x) ndispl (ComputeDeltaDefGrad,
// material_point.cpp:127-151, af = 1, the identity added last), F = dF F, J,
// V = J V0, the stress update; for non-rigid materials displ = sum N ndispl,
// accel = sum N naccel and (PIC) vel = sum N nvel; then
// CommitParticleKinematics (nonlinear_time_integration.cpp:79-96).
__global__ void __launch_bounds__(256) k64_implicit_g2p(GridGeom g, PartF64 P, NodesF64 N, MaterialTable mt, const double *__restrict__ ndispl,
                                 double dt, double gamma_nb, int flip, uint32_t *err) {
    int pid = blockIdx.x * blockDim.x + threadIdx.x;
    if (pid >= P.n) return;
    uint32_t m = P.key[pid];
    if (m >= (uint32_t)g.nelem) return;
    const size_t ld = P.ld;
    const int nodec = g.nodec;
    const double af = 1.0;
    int ixe, iye, ize;
    cell_ijk(g, (int)m, ixe, iye, ize);
    double xyg[3] = {P.x[pid], P.x[ld + pid], P.x[2 * ld + pid]};
    double dd[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k <= g.idimc[2]; k++)
        for (int j = 0; j <= g.idimc[1]; j++)
            for (int i = 0; i <= g.idimc[0]; i++) {
                int nid = (ixe + i) + g.nxc[0] * (iye + j) + g.nxc[0] * g.nxc[1] * (ize + k);
                double sfi, ds[3];
                sf_node(g, xyg, ixe + i, iye + j, ize + k, sfi, ds);
                double uu = ndispl[nid], uv = ndispl[nodec + nid], uw = ndispl[2 * nodec + nid];
                dd[0] += ds[0] * af * uu;
                dd[1] += ds[1] * af * uu;
                dd[2] += ds[2] * af * uu;
                dd[3] += ds[0] * af * uv;
                dd[4] += ds[1] * af * uv;
                dd[5] += ds[2] * af * uv;
                dd[6] += ds[0] * af * uw;
                dd[7] += ds[1] * af * uw;
                dd[8] += ds[2] * af * uw;
            }
    dd[0] += 1.0;
    dd[4] += 1.0;
    dd[8] += 1.0;
    double Fo[9], F[9];
    for (int q = 0; q < 9; q++) Fo[q] = P.F[q * ld + pid];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double acc = 0.0;
            for (int kk = 0; kk < 3; kk++) acc += dd[3 * i + kk] * Fo[3 * kk + j];
            F[3 * i + j] = acc;
        }
    for (int q = 0; q < 9; q++) {
        P.F[q * ld + pid] = F[q];
        P.ddg[q * ld + pid] = dd[q];
    }
    const double J = det3(F);
    P.J[pid] = J;
    P.V[pid] = J * P.V0[pid];
    const int mat = P.matid ? P.matid[pid] : 0;
    const MaterialDev &md = mt.m[mat];
    update_stress64(md, P.sig, ld, pid, F, dd, J, J, P.eps_p, dt);
    double du[3] = {0.0, 0.0, 0.0}, ap[3] = {0.0, 0.0, 0.0};
    double vp[3];
    for (int c = 0; c < 3; c++) vp[c] = flip ? P.v[c * ld + pid] : 0.0;
    if (md.model != -1) {
        for (int k = 0; k <= g.idimc[2]; k++)
            for (int j = 0; j <= g.idimc[1]; j++)
                for (int i = 0; i <= g.idimc[0]; i++) {
                    int nid = (ixe + i) + g.nxc[0] * (iye + j) + g.nxc[0] * g.nxc[1] * (ize + k);
                    double sfi, ds[3];
                    sf_node(g, xyg, ixe + i, iye + j, ize + k, sfi, ds);
                    for (int c = 0; c < 3; c++) du[c] += sfi * ndispl[c * nodec + nid];
                    for (int c = 0; c < 3; c++) ap[c] += sfi * N.na[c * nodec + nid];
                    if (!flip)
                        for (int c = 0; c < 3; c++) vp[c] += sfi * N.nv[c * nodec + nid];
                }
    }
    for (int c = 0; c < 3; c++) {
        if (flip) vp[c] += dt * ((1.0 - gamma_nb) * P.a[c * ld + pid] + gamma_nb * ap[c]);
        P.v[c * ld + pid] = vp[c];
        P.a[c * ld + pid] = ap[c];
        P.x[c * ld + pid] = xyg[c] + du[c];
    }
    if (!(J > 0.0)) atomicOr(err, kErrJacobian);
    if (!isfinite(xyg[0] + du[0]) || !isfinite(xyg[1] + du[1]) || !isfinite(xyg[2] + du[2]))
        atomicOr(err, kErrNonFinite);
}

}  // namespace

size_t f64_contrib_doubles(const GridGeom &g, size_t ld) {
    return ld * (size_t)((g.idimc[0] + 1) * (g.idimc[1] + 1) * (g.idimc[2] + 1)) * kP2GQ;
}

void launch_f64_p2g(hipStream_t s, const GridGeom &g, const PartF64 &P, const NodesF64 &N, const CellLists &L,
                    const BcDense &bc, int scheme, double fx, double fy, double fz, double mtol, double facl,
                    double *contrib) {
    if (contrib) {
        if (P.n > 0)
            hipLaunchKernelGGL(k64_contrib<false>, dim3(cdiv(P.n, 128)), dim3(128), 0, s, g, P, L, scheme, fx, fy, fz,
                               contrib);
        hipLaunchKernelGGL(k64_p2g_gather, dim3(cdiv(g.nodec, 128)), dim3(128), 0, s, g, N, L, bc, mtol, facl,
                           (const double *)contrib);
        return;
    }
    hipLaunchKernelGGL(k64_p2g, dim3(cdiv(g.nodec, 128)), dim3(128), 0, s, g, P, N, L, bc, scheme, fx, fy, fz, mtol,
                       facl);
}

void launch_f64_solve_grid(hipStream_t s, const GridGeom &g, const NodesF64 &N, const BcDense &bc, double mtol) {
    hipLaunchKernelGGL(k64_solve_grid, dim3(cdiv(g.nodec, 256)), dim3(256), 0, s, g, N, bc, mtol);
}

// MakNodalVol (module/mesh.cpp:160-202) for one node: cells around it in
// ascending order, points of each cell in (z, y, x) order, as the reference's
// scatter adds them.
__global__ void __launch_bounds__(128) k64_nodal_vol(GridGeom g, int npx, int npy, int npz, GaussOffsets dec, double volp,
                              double *__restrict__ nvol) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nodec) return;
    const int i = n % g.nxc[0], j = (n / g.nxc[0]) % g.nxc[1], k = n / (g.nxc[0] * g.nxc[1]);
    int cells[27];
    const int nc = node_cells(g, i, j, k, cells);
    double acc = 0.0;
    for (int q = 0; q < nc; q++) {
        int ixe, iye, ize;
        cell_ijk(g, cells[q], ixe, iye, ize);
        double xye[3], xyp[3];
        xye[0] = g.xmin[0] + g.dx[0] * ((double)ixe + 0.5);
        xye[1] = g.xmin[1] + g.dx[1] * ((double)iye + 0.5);
        xye[2] = g.xmin[2] + g.dx[2] * ((double)ize + 0.5);
        for (int iz = 0; iz < npz; iz++) {
            xyp[2] = xye[2] + dec.d[iz][2];
            for (int iy = 0; iy < npy; iy++) {
                xyp[1] = xye[1] + dec.d[iy][1];
                for (int ix = 0; ix < npx; ix++) {
                    xyp[0] = xye[0] + dec.d[ix][0];
                    double sfi, ds[3];
                    sf_node(g, xyp, i, j, k, sfi, ds);
                    acc += sfi * volp;
                }
            }
        }
    }
    nvol[n] = acc;
}

// excess nodal volume nei = max(0, nvof - nvol) and the block sums of nei^2
__global__ void k64_shift_nei(int nodec, const double *__restrict__ nvof, ShiftF64 S) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const int i0 = b * kShiftBlock;
    if (i0 >= nodec) return;
    double s = 0.0;
    for (int i = i0; i < nodec && i < i0 + kShiftBlock; i++) {
        const double e = -S.nvol[i] + nvof[i];
        const double ne = e > 0.0 ? e : 0.0;
        S.nei[i] = ne;
        s += ne * ne;
    }
    S.part_eu[b] = s;
}

// gradient of nei at the particle (old position), times 2 V (old volume)
__global__ void __launch_bounds__(128) k64_shift_geup(GridGeom g, PartF64 P, ShiftF64 S) {
    const int pid = blockIdx.x * blockDim.x + threadIdx.x;
    if (pid >= P.n) return;
    const size_t ld = P.ld;
    const uint32_t m = P.key[pid];
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    if (m < (uint32_t)g.nelem) {
        int ixe, iye, ize;
        cell_ijk(g, (int)m, ixe, iye, ize);
        const double xyg[3] = {P.x[pid], P.x[ld + pid], P.x[2 * ld + pid]};
        for (int k = 0; k <= g.idimc[2]; k++)
            for (int j = 0; j <= g.idimc[1]; j++)
                for (int i = 0; i <= g.idimc[0]; i++) {
                    const int ix = ixe + i, iy = iye + j, iz = ize + k;
                    const int nid = ix + g.nxc[0] * iy + g.nxc[0] * g.nxc[1] * iz;
                    double sfi, ds[3];
                    sf_node(g, xyg, ix, iy, iz, sfi, ds);
                    g0 += ds[0] * S.nei[nid];
                    g1 += ds[1] * S.nei[nid];
                    g2 += ds[2] * S.nei[nid];
                }
        const double s2 = 2.0 * P.V[pid];
        g0 *= s2;
        g1 *= s2;
        g2 *= s2;
    }
    S.geup[pid] = g0;
    S.geup[ld + pid] = g1;
    S.geup[2 * ld + pid] = g2;
}

__global__ void k64_shift_dot(int n, size_t ld, ShiftF64 S) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const int i0 = b * kShiftBlock;
    if (i0 >= n) return;
    double s = 0.0;
    for (int i = i0; i < n && i < i0 + kShiftBlock; i++)
        for (int c = 0; c < 3; c++) {
            const double v = S.geup[c * ld + i];
            s += v * v;
        }
    S.part_dot[b] = s;
}

__global__ void k64_shift_b0(int nbe, int nbd, double mtol, ShiftF64 S) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double eu = 0.0, dot = 0.0;
    for (int b = 0; b < nbe; b++) eu += S.part_eu[b];
    for (int b = 0; b < nbd; b++) dot += S.part_dot[b];
    *S.b0 = (dot < mtol) ? 0.0 : (eu / dot);
}

void launch_f64_nodal_vol(hipStream_t s, const GridGeom &g, int npx, int npy, int npz, const GaussOffsets &dec,
                          double volp, double *nvol) {
    hipLaunchKernelGGL(k64_nodal_vol, dim3(cdiv(g.nodec, 128)), dim3(128), 0, s, g, npx, npy, npz, dec, volp, nvol);
}

void launch_f64_g2p(hipStream_t s, const GridGeom &g, const PartF64 &P, const NodesF64 &N, const CellLists &L,
                    const BcDense &bc, const MaterialTable &mt, int scheme, double dt, double mtol, double facl,
                    uint32_t *err, const ShiftF64 *shift, double *njbar, double *contrib) {
    if (scheme != 1)
        hipLaunchKernelGGL(k64_vel_plus_dt_acc, dim3(cdiv(3 * g.nodec, 256)), dim3(256), 0, s, 3 * g.nodec, N.nv,
                           (const double *)N.na, dt);
    if (P.n > 0) hipLaunchKernelGGL(k64_g2p_vel, dim3(cdiv(P.n, 128)), dim3(128), 0, s, g, P, N, scheme, dt);
    if (contrib) {
        if (P.n > 0)
            hipLaunchKernelGGL(k64_contrib<true>, dim3(cdiv(P.n, 128)), dim3(128), 0, s, g, P, L, scheme, 0.0, 0.0,
                               0.0, contrib);
        hipLaunchKernelGGL(k64_musl_gather, dim3(cdiv(g.nodec, 128)), dim3(128), 0, s, g, N, L, bc, mtol, facl,
                           (const double *)contrib);
    } else {
        hipLaunchKernelGGL(k64_musl, dim3(cdiv(g.nodec, 128)), dim3(128), 0, s, g, P, N, L, bc, scheme, mtol, facl);
    }
    const double *geup = nullptr, *b0 = nullptr;
    if (shift && P.n > 0) {
        // update_point_coord_and_stress computes the correction before moving
        // any particle (old positions, old volumes)
        const int nbe = cdiv(g.nodec, kShiftBlock), nbd = cdiv(P.n, kShiftBlock);
        hipLaunchKernelGGL(k64_shift_nei, dim3(cdiv(nbe, 64)), dim3(64), 0, s, g.nodec, (const double *)N.nvof, *shift);
        hipLaunchKernelGGL(k64_shift_geup, dim3(cdiv(P.n, 128)), dim3(128), 0, s, g, P, *shift);
        hipLaunchKernelGGL(k64_shift_dot, dim3(cdiv(nbd, 64)), dim3(64), 0, s, P.n, P.ld, *shift);
        hipLaunchKernelGGL(k64_shift_b0, dim3(1), dim3(64), 0, s, nbe, nbd, mtol, *shift);
        geup = shift->geup;
        b0 = shift->b0;
    }
    const int fbar = njbar != nullptr;
    if (P.n > 0)
        hipLaunchKernelGGL(k64_g2p_update, dim3(cdiv(P.n, 128)), dim3(128), 0, s, g, P, N, mt, dt, err, geup, b0,
                           fbar);
    if (fbar && P.n > 0) {  // var_trans_solid_explicit.cpp:298-301
        hipLaunchKernelGGL(k64_fbar_nodes, dim3(cdiv(g.nodec, 128)), dim3(128), 0, s, g, P, N, L, mtol, njbar);
        hipLaunchKernelGGL(k64_fbar_finish, dim3(cdiv(P.n, 128)), dim3(128), 0, s, g, P, N, mt, dt, err,
                           (const double *)njbar, geup, b0);
    }
}

int launch_f64_maksf(hipStream_t s, const GridGeom &g, int n, const int32_t *cell, const double *xyg, int32_t *nc,
                     double *sf, double *dsf) {
    if (n > 0) hipLaunchKernelGGL(k64_maksf, dim3(cdiv(n, 128)), dim3(128), 0, s, g, n, cell, xyg, nc, sf, dsf);
    return (int)hipGetLastError();
}

void launch_f64_implicit_p2g(hipStream_t s, const GridGeom &g, const PartF64 &P, const NodesF64 &N,
                              const CellLists &L, const BcDense &bc, double mtol, double facl) {
    hipLaunchKernelGGL(k64_implicit_p2g, dim3(cdiv(g.nodec, 256)), dim3(256), 0, s, g, P, N, L, bc, mtol, facl);
}

void launch_f64_implicit_g2p(hipStream_t s, const GridGeom &g, const PartF64 &P, const NodesF64 &N,
                              const MaterialTable &mt, const double *ndispl, const NewmarkF64 &nb, double dt,
                              double gamma_nb, int flip, uint32_t *err) {
    const int n3 = 3 * g.nodec;
    hipLaunchKernelGGL(k64_newmark_nodes, dim3(cdiv(n3, 256)), dim3(256), 0, s, n3, ndispl, N.nv, N.na, nb);
    if (P.n > 0)
        hipLaunchKernelGGL(k64_implicit_g2p, dim3(cdiv(P.n, 256)), dim3(256), 0, s, g, P, N, mt, ndispl, dt, gamma_nb,
                           flip, err);
}

}  // namespace mpm

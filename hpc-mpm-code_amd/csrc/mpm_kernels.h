// Launchers shared between the kernel translation units and the engine.
#pragma once

#include "mpm_common.h"
#include "mpm_mlp.h"

namespace mpm {

// ---------------------------------------------------------------- binning
struct SortBuffers {
    uint32_t *keys;      // output buffers of alternating passes
    uint32_t *vals;
    uint32_t *keys_alt;
    uint32_t *vals_alt;
    uint32_t *hist;      // radix_hist_entries(n)
    uint32_t *bsum;      // >= 4096 entries
};

size_t radix_hist_entries(int n);
void launch_copy16(hipStream_t s, const void *src, void *dst, size_t n16);
// particle-update stream mix over n particles (HBM probe): floats of one
// buffer set, bytes moved per launch, launch
size_t stream_mix_floats(size_t n);
double stream_mix_bytes(size_t n);
void launch_stream_mix(hipStream_t s, float *p, size_t n);
int excl_scan_u32(hipStream_t s, uint32_t *data, int n, uint32_t *bsum);
// Stable LSD radix sort of (keys_in[i], i) over nbits key bits (standalone
// binning entry point). keys_in is not
// modified. Results point into the SortBuffers arrays.
int radix_sort_pairs(hipStream_t s, SortBuffers &b, const uint32_t *keys_in, int n, int nbits, uint32_t **keys_out,
                     uint32_t **vals_out);
void launch_fill_u32(hipStream_t s, uint32_t *p, size_t n, uint32_t v);
void launch_keys_f64(hipStream_t s, const GridGeom &g, const double *x, size_t ld, int n, uint32_t *keys,
                     uint32_t *err);

// Allocation-free binning of device AoS positions (noalloc.hip): all
// intermediates live in a fixed file-scope device array sized for
// kNoAllocMaxPositions positions (about 21 MB per GPU). Asynchronous on s.
constexpr int kNoAllocMaxPositions = 1 << 20;
int noalloc_bin_positions(const GridGeom &g, int n, const double *xyg, int32_t *cell_of, int32_t *order,
                          hipStream_t s);

// Cell lists produced by the binning: particles of cell c are
// perm[cstart[c] .. cend[c]) (cstart == cend for an empty cell); active[] lists
// the non-empty cells in ascending order, ctr->n_active of them; tiles[]
// flags the node tiles next to a non-empty cell.
struct BinCounters {
    uint32_t n_active;
    uint32_t stamp;  // binning generation, advanced on the device (graph replays stay distinct)
    uint32_t bad;    // counts inconsistent with the particle count: this binning placed nothing
};

struct CellLists {
    const uint32_t *perm;
    const uint32_t *cstart;
    const uint32_t *cend;
    const uint32_t *active;
    const uint32_t *astart;  // [n_active + 1]: start of each active cell, then the binned total
    const BinCounters *ctr;
    const uint32_t *tiles;
    const uint32_t *occ;     // one bit per cell: the cell holds particles
    const uint32_t *gs;      // one bit per cell: the cell starts a group of the FP32 cells kernels (P2G)
    uint32_t gsize;          // cells per group of the binning that wrote gs; 1: every cell starts one (no x-merge)
    const uint32_t *gs2;     // the same for the MUSL groups of the stencil kernels (64 cells)
    uint32_t gsize2;
};

struct TileGeom {
    int ntx, nty, ntz, ntiles;
};

__device__ __forceinline__ uint64_t lanemask_lt(int lane) { return (lane == 0) ? 0ull : (~0ull >> (64 - lane)); }

// Run structure of equal keys among the lanes of a wave: returns true for
// the lane that heads its run; *len = run length, *head = head lane.
__device__ __forceinline__ bool wave_runs(uint32_t k, int lane, int *len, int *head) {
    uint32_t prev = (uint32_t)__shfl_up((int)k, 1, 64);
    bool is_head = (lane == 0) || (k != prev);
    uint64_t H = __ballot(is_head);
    // head of this lane's run: highest set bit of H at or below lane
    uint64_t below = H & (lanemask_lt(lane) | (1ull << lane));
    *head = 63 - __clzll((long long)below);
    // next head strictly above this lane's head
    uint64_t above = (*head == 63) ? 0ull : (H & (~0ull << (*head + 1)));
    int nxt = above ? (__ffsll((unsigned long long)above) - 1) : 64;
    *len = nxt - *head;
    return is_head;
}

// Cell counting fused into the FP32 update kernel (the next binning's first
// pass): counts of the new keys, and the clearing of tile flags and counters.
struct BinFuse {
    uint32_t *cnt;  // null: not fused
    int delta;      // 1: cnt holds the counts of the keys being replaced; only moves are applied
    uint64_t *tsum; // with delta: the tile totals of cnt (BinBuffers::bsum), kept up to date as well
    uint32_t *tiles;
    int ntiles;
    BinCounters *ctr;
};

struct BinBuffers {
    uint32_t *cnt;      // [bin_cell_capacity]: zero, or the counts of the current keys (kept counts)
    uint32_t *cstart;   // [bin_cell_capacity]
    uint32_t *cend;     // [bin_cell_capacity]
    uint32_t *active;   // [nelem]
    uint32_t *astart;   // [nelem + 1]
    uint32_t *stamp;    // [nelem], stamp of the binning that marked the cell for the fix-up
    uint32_t *perm;     // [n]
    uint32_t *tiles;    // [ntiles]
    uint32_t *occ;      // [bin_cell_capacity / 32] occupancy bits (zero at allocation)
    uint32_t *gs;       // [bin_cell_capacity / 32] group-start bits (zero at allocation): set for the
                        // active cells whose active-list index is a multiple of gsize
    uint32_t gsize;     // cells per group of the FP32 P2G cells kernels (1: no x-merge)
    uint32_t *gs2;      // [bin_cell_capacity / 32] the same for gsize2 (not written when gsize2 == 0)
    uint32_t gsize2;    // cells per group of the stencil MUSL cells kernels (0: unused)
    uint64_t *bsum;     // [bin_scan_blocks(nelem)] packed (non-empty cells, particles) per tile of
                        // kBinTile cells: written by the reduce pass, or kept with kept counts
    uint32_t *tprev;    // [bin_scan_blocks(nelem)] tile non-empty at the previous binning (zero at allocation)
    BinCounters *ctr;
};

TileGeom make_tile_geom(const GridGeom &g);
size_t bin_scan_blocks(int nelem);
// per-cell arrays (cnt, cstart, cend) must hold this many entries (padding zeroed)
size_t bin_cell_capacity(int nelem);
// Counting-sort binning of keys[0..n) (sentinel nelem = outside the grid).
// count pass alone (per-cell counts of keys, their tile totals, cleared tiles
// and counters)
void launch_bin_count(hipStream_t s, const GridGeom &g, const TileGeom &tg, const uint32_t *keys, int n,
                      BinBuffers &b);
// tile totals (b.bsum) of the per-cell counts b.cnt
void launch_bin_reduce(hipStream_t s, const GridGeom &g, BinBuffers &b);
// counted: the per-cell counts and their tile totals, cleared tiles and
// counters were produced by the previous update kernel (fused counting), so
// the count and reduce passes are skipped.
// keep: the counts are left in place (they are the counts of keys[], which
// the next update kernel then only adjusts, with their tile totals, for the
// particles that change cell); otherwise they are zeroed.
int launch_bin(hipStream_t s, const GridGeom &g, const TileGeom &tg, const uint32_t *keys, int n, BinBuffers &b,
               uint32_t *err, bool counted = false, bool keep = false);

// ------------------------------------------------------ FP64 exact kernels
struct PartF64 {
    double *x;      // [3][ld]
    double *v;      // [3][ld]
    double *m;
    double *V;
    double *V0;
    double *sig;    // [6][ld]
    double *F;      // [9][ld]
    double *J;
    double *trac;   // [3][ld] or null
    int *matid;     // or null
    double *eps_p;  // or null
    double *ddg;    // [9][ld] deformation-gradient increment
    double *C;      // [9][ld] TPIC velocity gradient / APIC B matrix, row-major (null for PIC/FLIP)
    double *Dinv;   // [9][ld] APIC inverse D matrix (null otherwise)
    double *Fbar;   // [9][ld] F-bar deformation gradient (null when F-bar is off)
    double *Jbar;   // [ld] its determinant as projected
    double *detdd;  // [ld] det of the step's deformation-gradient increment (F-bar scratch)
    double *a;      // [3][ld] particle acceleration of the implicit transfers (null until used)
    const uint32_t *key;
    size_t ld;
    int n;
};

struct NodesF64 {
    double *nm;
    double *nvof;
    double *np;  // [3][nodec]
    double *nf;
    double *nv;
    double *na;
};

// contrib: scratch of f64_contrib_doubles(g, ld) doubles for the two-pass
// sums (bit-identical to the one-pass gathers used when it is null)
size_t f64_contrib_doubles(const GridGeom &g, size_t ld);
void launch_f64_p2g(hipStream_t s, const GridGeom &g, const PartF64 &P, const NodesF64 &N, const CellLists &L,
                    const BcDense &bc, int scheme, double fx, double fy, double fz, double mtol, double facl,
                    double *contrib = nullptr);
void launch_f64_solve_grid(hipStream_t s, const GridGeom &g, const NodesF64 &N, const BcDense &bc, double mtol);
// Particle shifting of the reference (DeltaCorrectionParticleShifting), FP64:
// nvol = nodal volume of the filled grid; nei, geup, partial sums and b0 are
// scratch. Global sums are taken in blocks of kShiftBlock consecutive entries.
constexpr int kShiftBlock = 256;
struct ShiftF64 {
    const double *nvol;  // [nodec]
    double *nei;         // [nodec]
    double *geup;        // [3][ld]
    double *part_eu;     // [cdiv(nodec, kShiftBlock)]
    double *part_dot;    // [cdiv(n, kShiftBlock)]
    double *b0;          // [1]
};
struct GaussOffsets {
    double d[6][3];
};
void launch_f64_nodal_vol(hipStream_t s, const GridGeom &g, int npx, int npy, int npz, const GaussOffsets &dec,
                          double volp, double *nvol);
void launch_f64_g2p(hipStream_t s, const GridGeom &g, const PartF64 &P, const NodesF64 &N, const CellLists &L,
                    const BcDense &bc, const MaterialTable &mt, int scheme, double dt, double mtol, double facl,
                    uint32_t *err, const ShiftF64 *shift = nullptr, double *njbar = nullptr, double *contrib = nullptr);
int launch_f64_maksf(hipStream_t s, const GridGeom &g, int n, const int32_t *cell, const double *xyg, int32_t *nc,
                     double *sf, double *dsf);
// Implicit-path transfers (reference SolidMaterialPoint::Particle2Node /
// Node2Particle, PIC and FLIP): node mass, volume, momentum, m a, velocity
// and acceleration with BCs; then, for the nodal displacement increment
// ndispl ([u|v|w] blocks), the Newmark node update with coefficients
// NewmarkBetaParaSet's c[0..5] and the particle update.
struct NewmarkF64 {
    double c[6];
};
void launch_f64_implicit_p2g(hipStream_t s, const GridGeom &g, const PartF64 &P, const NodesF64 &N,
                              const CellLists &L, const BcDense &bc, double mtol, double facl);
void launch_f64_implicit_g2p(hipStream_t s, const GridGeom &g, const PartF64 &P, const NodesF64 &N,
                              const MaterialTable &mt, const double *ndispl, const NewmarkF64 &nb, double dt,
                              double gamma_nb, int flip, uint32_t *err);

// ------------------------------------------------------------ FP32 kernels
struct PartF32 {
    float *r;       // [3][ld] position inside the cell, in cell units: x = xmin + dx (i + r)
    float *v;       // [3][ld]
    float *m;
    float *V0;
    float *V32;     // volume V = J V0 rounded to FP32: written by every update, read by P2G
    double *V;      // FP64 volume and det F: the uploaded values, and those of a particle when it
    double *J;      // leaves the grid; for particles inside, launch_f32_refresh_jv derives them from H
    float *H;       // [9][ld] displacement gradient F - I
    float *sig;     // [6][ld]
    float *eps_p;   // J2 equivalent plastic strain (or null)
    float *trac;    // [3][ld] or null
    int *matid;     // or null
    uint32_t *key;  // cell of each slot
    float *C;       // [9][ld] TPIC velocity gradient / APIC B matrix, row-major (null for PIC/FLIP)
    float *Dinv;    // [9][ld] APIC inverse D matrix (null otherwise)
    // F-bar projection (null when off): the projected gradient as Fbar - I
    // and Jbar - 1, kept from step to step; within a step, the increment
    // dF - I and Jbar det dF - 1 that the projection's node average reads
    float *Hb;      // [9][ld]
    float *jb;
    float *Dd;      // [9][ld]
    float *jm;
    size_t ld;
    int n;
};

struct NodesF32 {
    float *nm;
    float *nv;    // [3][nodec] velocity after the P2G grid update (+dt a for PIC)
    float *na;    // [3][nodec] acceleration (FLIP and diagnostics)
    float *nvm;   // [3][nodec] MUSL velocity
    float *nvp;   // [3][nodec] v + dt a, the node velocity PIC's G2P interpolates
    float *np;    // [3][nodec] momentum (diagnostics, may be null)
    float *nf;    // [3][nodec] force (diagnostics, may be null)
    float *nvof;  // [nodec] volume (diagnostics, may be null)
    float *nje;   // [nodec] F-bar: node average of Jbar det dF, minus 1 (-1 where nvof <= mtol)
};

struct StepConsts {
    float bx, by, bz;  // body force * facl
    float dt;
    double dt64;
    double mtol;
    double facl;
    float m_small;     // below this node mass the division runs in FP64
    int scheme;
    int lanes_per_cell;  // 1, 2, 4 or 8 lanes fold one cell's particles in the cells kernels
    int cells_prefetch;  // cells kernels request the next chunk's indices while a chunk loads
    int cells_grid;      // workgroups of the cells kernels (0 = automatic)
    int mat_dt;          // 1: each particle advances by its material's time step (MaterialDev::dt)
    const double *sh_geup;  // particle shifting (or null): [3][ld] gradient of the excess volume by slot
    const double *sh_b0;    // and the step length b0; the update adds -b0 geup to the position
    int fbar;               // F-bar projection: the update reads the increment of the F-bar pass
};

// Grid spacing constants of the FP32 kernels.
struct NodeTables {
    float inv_dx[3];
    float dx[3];
    float inv_nex;   // 1 / nex and 1 / (nex ney) for the cell-index split
    float inv_nexy;
};

// Partials buffer of the tent kernels: up to 64 floats per cell, as 8 slots
// (x-side of the corner * 4 + y-side + 2 z-side), then the cell, then the
// quantities padded to 8 floats for P2G ({m, px, py, pz, fx, fy, fz}, and the
// corner volume in diagnostic mode) or to 4 for MUSL ({px, py, pz}).
// Left-corner slots hold the x-merged sums; right-corner slots are stored only
// where no merge took them.
constexpr int kPartialsPerCell = 64;
// lanes of one FP32 tent cells-kernel group (one wave): a group holds
// kGroupLanes / lanes-per-cell consecutive active cells
constexpr int kGroupLanes = 64;

// P2G: corner partials per cell + node pass with the grid update (v = p/m,
// a = f/m, BCs). diag != 0 also writes momentum, force and nodal volume.
void launch_f32_p2g(hipStream_t s, const GridGeom &g, const PartF32 &P, const NodesF32 &N, const CellLists &L,
                    const BcDense &bc, const NodeTables &T, const TileGeom &tg, const StepConsts &c, float *partials,
                    int diag, hipEvent_t mid = nullptr);
// G2P velocity fused with the MUSL momentum remap + MUSL node pass.
void launch_f32_g2p_vel_musl(hipStream_t s, const GridGeom &g, const PartF32 &P, const NodesF32 &N,
                             const CellLists &L, const BcDense &bc, const NodeTables &T, const TileGeom &tg,
                             const StepConsts &c, float *partials, int diag, hipEvent_t mid = nullptr,
                             const MaterialTable *mt = nullptr);
// Stencil kernels (stencil32.hip): 2 or 3 control points per axis (idimc 1 or
// 2; combinations (1,1,1), (2,2,2) and (2,1,2)), all four transfer schemes.
// The engine uses them for the quadratic B-spline and for TPIC / APIC; the
// tent-function PIC / FLIP step keeps the kernels above.
bool s32_supported(const GridGeom &g);
int s32_partials_per_cell(const GridGeom &g);
// cells per group of the stencil P2G cells kernel (64 / lanes per cell)
int s32_p2g_cells_per_group(const GridGeom &g);
void launch_s32_p2g(hipStream_t s, const GridGeom &g, const PartF32 &P, const NodesF32 &N, const CellLists &L,
                    const BcDense &bc, const NodeTables &T, const TileGeom &tg, const StepConsts &c, float *partials,
                    int diag, hipEvent_t mid = nullptr);
void launch_s32_g2p_vel_musl(hipStream_t s, const GridGeom &g, const PartF32 &P, const NodesF32 &N,
                             const CellLists &L, const BcDense &bc, const NodeTables &T, const TileGeom &tg,
                             const StepConsts &c, float *partials, int diag, hipEvent_t mid = nullptr);
// Particle shifting correction of the reference (DeltaCorrectionParticleShifting)
// for the FP32 state, before the update kernel: excess nodal volume from the
// P2G volume nvof (FP32) and the nodal volume of the filled grid (FP64), its
// gradient at every slot's position (2 V sum grad N nei, FP64 sums over the
// stencil), the two blocked global sums and b0 into S.
void launch_s32_shift(hipStream_t s, const GridGeom &g, const PartF32 &P, const NodesF32 &N, const NodeTables &T,
                      const ShiftF64 &S, double mtol);
// F-bar projection (fbar_projection.cpp:18-94), before the update: the
// increment dF - I and F of every particle, then the node average of
// Jbar det dF weighted by N V (cell partials and a node pass, as P2G); the
// update (StepConsts::fbar) interpolates it back, scales the increment and
// advances Fbar and Jbar.
void launch_s32_fbar(hipStream_t s, const GridGeom &g, const PartF32 &P, const NodesF32 &N, const CellLists &L,
                     const NodeTables &T, const TileGeom &tg, const StepConsts &c, float *partials);
void launch_s32_g2p_update(hipStream_t s, const GridGeom &g, const PartF32 &P, const NodesF32 &N,
                           const NodeTables &T, const MaterialTable &mt, const StepConsts &c, uint32_t *err,
                           const BinFuse &bf);

// dF, H, J, V, x, stress and the new cell key of every slot.
// J2 constitutive surrogate in the FP32 update: mode 0 = exact update only,
// 1 = exact update and sample logging, 2 = exact update and a queue of the
// plastic trial states, which launch_f32_j2_surrogate then runs through the
// network (residual check, exact result kept on rejection). Each 256-slot
// block of the update kernel compacts its plastic states into its own queue
// region (rows [256 b, 256 b + count_b)) and stores count_b, so the queue is
// filled without any global atomic.
// floats per queued state. J2: slot, 13 features, trial stress (6), E, nu,
// sy, H. Tait: slot, the 2 features, exact pressure, B = rho0 c^2 / gamma,
// zero padding. kSurQueueRow is the larger (the queue's allocation).
constexpr int kSurRowJ2 = 24;
constexpr int kSurRowTait = 8;
constexpr int kSurQueueRow = kSurRowJ2;
constexpr int kSurRegion = 256;   // queue rows per update-kernel block (= its block size)
// Models the surrogate stands in for: the J2 plastic corrector, or the Tait
// pressure (a closure p(J, rate of volume) with the exact equation of state
// as fallback and check).
enum { kSurJ2 = 0, kSurTait = 1 };
struct SurrogateArgs {
    int mode;
    int kind;                      // kSurJ2 / kSurTait
    float tol;                     // mode 2: accepted residual / yield stress
    MlpParams net;                 // mode 2
    const float *net_scaling;      // mode 2: [in_mu | in_isd | out_mu | out_sd], 64 floats
    float *queue;                  // mode 2: [qcap][kSurQueueRow]
    uint32_t *bcount;              // mode 2: queued states per region (one region per update block)
    uint32_t nregions;             // mode 2: update blocks = cdiv(n, kSurRegion)
    uint32_t *rstat;               // mode 2: per region [evaluated, rejected], summed over steps
    uint32_t qcap;                 // mode 2: queue rows (>= nregions * kSurRegion)
    unsigned long long *ctr;       // [2] samples offered to the log ([0], [1]: evaluated, fallbacks of earlier forms)
    float *log_in, *log_tgt;       // mode 1: [cap][16]
    uint32_t *log_key;             // mode 1: step_tag * n + particle id
    uint32_t cap;
    uint32_t sel_threshold;        // mode 1: keep when hash(pid, step) <= threshold
    uint32_t step_tag;
    const int *pid;                // particle id by slot
};

void launch_f32_g2p_update(hipStream_t s, const GridGeom &g, const PartF32 &P, const NodesF32 &N,
                           const NodeTables &T, const MaterialTable &mt, const StepConsts &c, uint32_t *err,
                           const SurrogateArgs &sa, const BinFuse &bf);
// network pass over the queued plastic states of this step (mode 2)
void launch_f32_j2_surrogate(hipStream_t s, const PartF32 &P, const SurrogateArgs &sa);
// J = det(I + H) and V = J V0 in FP64 (the values the update kernel computes)
// into P.J / P.V for every particle inside the grid
void launch_f32_refresh_jv(hipStream_t s, const GridGeom &g, const PartF32 &P);
void launch_f32_permute(hipStream_t s, const PartF32 &src, const PartF32 &dst, const uint32_t *perm,
                        const int *pid_src, int *pid_dst);
void launch_iota_u32(hipStream_t s, uint32_t *p, int n);

}  // namespace mpm

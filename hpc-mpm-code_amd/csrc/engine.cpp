// Host side of the C ABI (include/mpm/mpm_capi.h): device memory, layout
// conversion between the reference's AoS FP64 arrays and the device SoA,
// and the orchestration of the per-pass and whole-step kernels.
#include "mpm_common.h"
#include "mpm_kernels.h"
#include "surrogate_host.h"

#include "../../include/mpm/mpm_capi.h"
#include "../../include/mpm/mpm_ensemble.h"
#include "../../include/mpm/mpm_surrogate.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

namespace mpm {

static thread_local std::string g_last_error;

static int set_error(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int set_error_msg(int code, const char *msg) { return set_error(code, msg); }

// Status of an exception caught at the C boundary (called from a catch
// block): no C++ exception leaves an entry point.
int exception_status() {
    try {
        throw;
    } catch (const std::bad_alloc &) {
        return set_error(MPM_ERR_HOST_MEMORY, "host memory allocation failed");
    } catch (const std::exception &ex) {
        return set_error(MPM_ERR_INTERNAL, std::string("exception: ") + ex.what());
    } catch (...) {
        return set_error(MPM_ERR_INTERNAL, "unknown exception");
    }
}

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
    char buf[512];
    snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file, line);
    g_last_error = buf;
    return MPM_ERR_HIP;
}

static_assert(MPM_NOALLOC_MAX_POSITIONS == kNoAllocMaxPositions, "allocation-free binning limit");

static int bit_length(uint32_t v) {
    int b = 0;
    while (v) {
        b++;
        v >>= 1;
    }
    return b;
}

}  // namespace mpm

using namespace mpm;

enum {
    OPT_FULL_NODE_OUTPUT = 1,
    OPT_LANES_PER_CELL = 2,
    OPT_RELAYOUT_EVERY = 3,
    OPT_GRAPH_BLOCK = 4,
    OPT_FUSED_COUNT = 5,
    OPT_CELLS_PREFETCH = 6,
    OPT_CELLS_GRID = 7,
    OPT_X_MERGE = 8,
    OPT_SURROGATE_FUSED = 9
};

struct mpm_engine {
    GridGeom g{};
    mpm_step_params prm{};
    MaterialTable mt{};
    std::vector<MaterialDev> mat_host;  // host copy of the device material table
    int mat_dt = 0;                     // FP32: per-material time steps (mpm_set_material_dt)
    int prec = MPM_PREC_FP64_EXACT;
    hipStream_t stream = nullptr;
    int device = 0;

    // boundary conditions, host lists and dense device form
    std::vector<int32_t> nbc[4];
    std::vector<double> fbc[4];
    uint8_t *d_bcmask = nullptr;
    double *d_bcval = nullptr;
    bool bc_dirty = true;

    int n = 0;
    size_t ld = 0;
    bool has_trac = false, has_matid = false, has_epsp = false;

    PartF64 P64{};
    NodesF64 N64{};

    PartF32 P32[2]{};
    int *pid[2] = {nullptr, nullptr};
    int cur = 0;
    NodesF32 N32{};
    float *partials = nullptr;
    bool stencil = false;  // FP32: the stencil kernels (quadratic B-spline on some axis, or TPIC / APIC)
    bool stencil_base = false;  // the choice by shape functions and scheme (F-bar also takes the stencil kernels)
    NodeTables T{};
    float m_small = 0.f;
    int lanes_per_cell = 1;  // from the particles per non-empty cell at upload
    bool x_merge = true;     // FP32 tent cells kernels add x-neighbour corner partials within a group
    int cells_prefetch = 1;
    int cells_grid = 0;
    int fp32_outside_at_upload = 0;

    uint32_t *keys64 = nullptr;  // FP64 mode cell keys by particle
    BinBuffers bb{};             // counting-sort binning
    TileGeom tg{};
    uint32_t *perm = nullptr;    // bb.perm (or perm_id right after a relayout)
    uint32_t *perm_id = nullptr;
    uint32_t *d_err = nullptr;
    double *d_ndispl = nullptr;  // [3*nodec] nodal displacement increment of the implicit G2P
    bool bins_valid = false;
    bool keys_valid = false;
    bool counts_ready = false;  // bb.cnt holds the per-cell counts of the current keys and bb.bsum their
                                // tile totals (else the counts are zero)
    bool jv_derived = false;    // FP32: a step ran since the upload; det F and V follow from H at download
    bool fused_count = true;    // FP32: the update kernel counts the new keys for the next binning
    int last_pass = 0;  // 1 = p2g, 2 = g2p
    int steps_since_relayout = 0;
    int full_node_output = 1;
    int relayout_every = 0;
    int active_cells = 0;

    // reference particle shifting (FP64-exact engine)
    bool shifting = false;
    double *d_nvol = nullptr, *d_nei = nullptr, *d_part_eu = nullptr, *d_b0 = nullptr;
    double *d_geup = nullptr, *d_part_dot = nullptr;  // particle-sized
    int geup_n = 0;

    // F-bar projection (FP64-exact engine)
    bool fbar = false;
    double *d_njbar = nullptr;
    // FP64-exact engine: per-particle values of the two-pass P2G / MUSL sums
    // (allocated with the particles; null: one-pass gathers)
    double *d_contrib = nullptr;

    // whole-step graph: blocks of graph_block steps captured once and replayed
    int graph_block = 0;
    int graph_cur = -1, graph_since = -1;  // particle buffer and relayout counter the graph starts from
    hipGraph_t step_graph = nullptr;
    hipGraphExec_t step_exec = nullptr;
    void drop_graph() {
        if (step_exec) (void)hipGraphExecDestroy(step_exec);
        if (step_graph) (void)hipGraphDestroy(step_graph);
        step_exec = nullptr;
        step_graph = nullptr;
    }

    bool timing = false;
    mpm_kernel_times kt{};
    std::vector<hipEvent_t> ev;

    std::vector<void *> allocs;
    std::vector<void *> part_allocs;

    // J2 constitutive surrogate
    struct Surrogate {
        bool logging = false, enabled = false, have_net = false;
        MlpDev net;
        float *log_in = nullptr, *log_tgt = nullptr;
        uint32_t *log_key = nullptr;
        int64_t cap = 0, nsamples = 0, logged_steps = 0;
        unsigned long long *ctr = nullptr;  // evaluated, fallbacks, offered
        float *queue = nullptr;             // plastic trial states of a step, [qcap][kSurQueueRow]
        uint32_t *bcount = nullptr;         // queued states per region of kSurRegion rows
        uint32_t *rstat = nullptr;          // per region: [evaluated, rejected], summed over steps
        int64_t stat_base[2] = {0, 0};      // counts of regions folded in before a reallocation
        int64_t qcap = 0;                   // rows: regions * kSurRegion, one region per update block
        uint32_t thr = 0, step_tag = 0;
        int kind = kSurJ2;  // the model the network stands in for (from the materials)
        bool fused = true;  // the network runs inside the update kernel (else a queue and a network pass)
        bool sorted = false;
        float loss = 0.f, tol = 1e-4f;
        void free_log() {
            for (void *p : {(void *)log_in, (void *)log_tgt, (void *)log_key})
                if (p) (void)hipFree(p);
            log_in = log_tgt = nullptr;
            log_key = nullptr;
            cap = nsamples = 0;
        }
        ~Surrogate() {
            free_log();
            if (ctr) (void)hipFree(ctr);
            if (queue) (void)hipFree(queue);
            if (bcount) (void)hipFree(bcount);
            if (rstat) (void)hipFree(rstat);
            net.release();
        }
    } sur;

    template <typename T>
    int dalloc(T **p, size_t count, bool particle = false) {
        *p = nullptr;
        if (count == 0) return 0;
        hipError_t e = hipMalloc((void **)p, count * sizeof(T));
        if (e != hipSuccess) return hip_fail(e, "hipMalloc", __FILE__, __LINE__);
        (particle ? part_allocs : allocs).push_back((void *)*p);
        return 0;
    }

    // counts left by a fused update kernel that no binning consumed
    void clear_counts() {
        if (counts_ready && bb.cnt) launch_fill_u32(stream, bb.cnt, bin_cell_capacity(g.nelem), 0u);
        counts_ready = false;
    }

    void free_particles() {
        clear_counts();
        if (stream) (void)hipStreamSynchronize(stream);
        for (void *p : part_allocs) (void)hipFree(p);
        part_allocs.clear();
        P64 = PartF64{};
        P32[0] = PartF32{};
        P32[1] = PartF32{};
        pid[0] = pid[1] = nullptr;
        keys64 = nullptr;
        bb.perm = nullptr;
        perm = nullptr;
        perm_id = nullptr;
        d_contrib = nullptr;
        n = 0;
        bins_valid = keys_valid = false;
    }

    ~mpm_engine() {
        drop_graph();
        for (void *p : {(void *)d_geup, (void *)d_part_dot})
            if (p) (void)hipFree(p);
        free_particles();
        for (void *p : allocs) (void)hipFree(p);
        for (auto e : ev) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }

    BcDense bc() const { return BcDense{d_bcmask, d_bcval}; }
    CellLists lists() const {
        return CellLists{perm,  bb.cstart, bb.cend, bb.active, bb.astart, bb.ctr,   bb.tiles,
                         bb.occ, bb.gs,     bb.gsize, bb.gs2,    bb.gsize2};
    }
};

namespace {

int upload_bc(mpm_engine *e) {
    if (!e->bc_dirty) return 0;
    const int nodec = e->g.nodec;
    std::vector<uint8_t> mask(nodec, 0);
    std::vector<double> val(3 * (size_t)nodec, 0.0);
    // BCSetVal order of var_trans_solid_explicit.cpp:91-96: u, v, w, then rigid on all comps
    for (int l = 0; l < 4; l++) {
        for (size_t i = 0; i < e->nbc[l].size(); i++) {
            int nd = e->nbc[l][i];
            for (int c = 0; c < 3; c++) {
                if (l < 3 && c != l) continue;
                mask[nd] |= (uint8_t)(1u << c);
                val[(size_t)c * nodec + nd] = e->fbc[l][i];
            }
        }
    }
    MPM_HIP_CHECK(hipMemcpyAsync(e->d_bcmask, mask.data(), nodec, hipMemcpyHostToDevice, e->stream));
    MPM_HIP_CHECK(hipMemcpyAsync(e->d_bcval, val.data(), val.size() * sizeof(double), hipMemcpyHostToDevice,
                                 e->stream));
    MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    e->bc_dirty = false;
    return 0;
}

double facl_of(const mpm_engine *e, int istep) {
    // work/src_solid/explicit/main.cpp:43-48
    if (istep <= e->prm.nlstep) return (1.0 / (double)e->prm.nlstep) * (double)istep;
    return 1.0;
}

StepConsts consts(const mpm_engine *e, double facl) {
    StepConsts c{};
    c.bx = (float)(e->prm.bb[0] * facl);
    c.by = (float)(e->prm.bb[1] * facl);
    c.bz = (float)(e->prm.bb[2] * facl);
    c.dt = (float)e->prm.dt;
    c.dt64 = e->prm.dt;
    c.mtol = e->prm.mtol;
    c.facl = facl;
    c.m_small = e->m_small;
    c.scheme = e->prm.scheme;
    c.lanes_per_cell = e->lanes_per_cell;
    c.cells_prefetch = e->cells_prefetch;
    c.cells_grid = e->cells_grid;
    c.mat_dt = e->mat_dt;
    return c;
}

int check_err(mpm_engine *e) {
    uint32_t h = 0;
    MPM_HIP_CHECK(hipMemcpyAsync(&h, e->d_err, sizeof(uint32_t), hipMemcpyDeviceToHost, e->stream));
    MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    if ((h & kErrOutOfGrid) || e->fp32_outside_at_upload)
        return set_error(MPM_ERR_OUT_OF_GRID, "a particle is outside the background grid");
    if (h & kErrNonFinite) return set_error(MPM_ERR_NONFINITE, "non-finite particle state");
    if (h & kErrJacobian) return set_error(MPM_ERR_JACOBIAN, "det F <= 0");
    if (h & kErrCellOverflow) return set_error(MPM_ERR_UNSUPPORTED, "more than 4096 particles in one cell");
    if (h & kErrBinCounts) return set_error(MPM_ERR_INTERNAL, "binning: cell counts exceed the particle count");
    return 0;
}

int clear_err(mpm_engine *e) {
    launch_fill_u32(e->stream, e->d_err, 1, 0u);
    return 0;
}

// keys (if needed) + radix sort + cell ranges
int do_bin(mpm_engine *e) {
    const int n = e->n;
    if (!e->keys_valid) {
        if (e->prec == MPM_PREC_FP64_EXACT)
            launch_keys_f64(e->stream, e->g, e->P64.x, e->ld, n, e->keys64, e->d_err);
        e->keys_valid = true;  // FP32 keys are uploaded and then kept by the update kernel
    }
    const uint32_t *keys = (e->prec == MPM_PREC_FP64_EXACT) ? e->keys64 : e->P32[e->cur].key;
    // FP32 with fused counting keeps the counts: the update kernel then moves
    // only the counts of particles that change cell
    const bool keep = e->prec == MPM_PREC_FP32 && e->fused_count;
    // group-start bits for the cells kernels' x-merge: the tent kernels group 64 / lanes-per-cell
    // cells (P2G and MUSL alike); the stencil kernels 64 / z-layer lanes per cell in P2G and 64 in MUSL
    const bool xm = e->x_merge && e->prec == MPM_PREC_FP32;
    if (e->stencil) {
        e->bb.gsize = xm ? (uint32_t)s32_p2g_cells_per_group(e->g) : 1u;
        e->bb.gsize2 = xm ? (uint32_t)kGroupLanes : 1u;
    } else {
        e->bb.gsize = xm ? (uint32_t)(kGroupLanes / std::max(1, e->lanes_per_cell)) : 1u;
        e->bb.gsize2 = 0u;
    }
    int rc = launch_bin(e->stream, e->g, e->tg, keys, n, e->bb, e->d_err, e->counts_ready, keep);
    e->counts_ready = keep;
    if (rc) return set_error(MPM_ERR_HIP, "binning launch failed");
    e->perm = e->bb.perm;
    if (e->prec == MPM_PREC_FP32 && e->relayout_every > 0 && ++e->steps_since_relayout >= e->relayout_every) {
        // physically reorder particle slots into the sorted order
        const int o = 1 - e->cur;
        if (!e->P32[o].r || !e->pid[o]) return set_error(MPM_ERR_INTERNAL, "relayout without a second particle buffer");
        launch_f32_permute(e->stream, e->P32[e->cur], e->P32[o], e->perm, e->pid[e->cur], e->pid[o]);
        e->cur = o;
        launch_iota_u32(e->stream, e->perm_id, n);
        e->perm = e->perm_id;
        e->steps_since_relayout = 0;
    }
    e->bins_valid = true;
    MPM_HIP_CHECK(hipGetLastError());
    return 0;
}

// Scratch of the FP64 two-pass sums (allocated with the particles; null when
// that allocation failed: the one-pass gathers run, with the same results).
double *f64_contrib(mpm_engine *e) { return e->d_contrib; }

int do_p2g(mpm_engine *e, double facl, int diag, hipEvent_t mid = nullptr) {
    if (e->shifting || e->fbar) diag = 1;  // shifting and F-bar read the P2G nodal volume
    if (e->prec == MPM_PREC_FP64_EXACT) {
        launch_f64_p2g(e->stream, e->g, e->P64, e->N64, e->lists(), e->bc(), e->prm.scheme, e->prm.bb[0] * facl,
                       e->prm.bb[1] * facl, e->prm.bb[2] * facl, e->prm.mtol, facl, f64_contrib(e));
    } else if (e->stencil) {
        launch_s32_p2g(e->stream, e->g, e->P32[e->cur], e->N32, e->lists(), e->bc(), e->T, e->tg, consts(e, facl),
                       e->partials, diag, mid);
    } else {
        launch_f32_p2g(e->stream, e->g, e->P32[e->cur], e->N32, e->lists(), e->bc(), e->T, e->tg, consts(e, facl),
                       e->partials, diag, mid);
    }
    e->last_pass = 1;
    MPM_HIP_CHECK(hipGetLastError());
    return 0;
}

int do_solve(mpm_engine *e) {
    if (e->prec == MPM_PREC_FP64_EXACT) launch_f64_solve_grid(e->stream, e->g, e->N64, e->bc(), e->prm.mtol);
    // FP32: the acceleration is formed in the P2G node pass
    MPM_HIP_CHECK(hipGetLastError());
    return 0;
}

// Surrogate queue rows for n particles: one region of kSurRegion rows per update block.
static int64_t sur_queue_rows(int64_t n) { return (int64_t)cdiv((int)n, kSurRegion) * kSurRegion; }

// Particle-sized buffers of the shifting correction (gradient by particle
// or slot, per-block sums), sized for the current upload.
int shift_buffers(mpm_engine *e, ShiftF64 *sh) {
    if (e->geup_n != e->n) {
        for (double **p : {&e->d_geup, &e->d_part_dot}) {
            if (*p) (void)hipFree(*p);
            *p = nullptr;
        }
        MPM_HIP_CHECK(hipMalloc((void **)&e->d_geup, 3 * (size_t)e->ld * sizeof(double)));
        MPM_HIP_CHECK(hipMalloc((void **)&e->d_part_dot, (size_t)cdiv(e->n, kShiftBlock) * sizeof(double)));
        e->geup_n = e->n;
    }
    *sh = ShiftF64{e->d_nvol, e->d_nei, e->d_geup, e->d_part_eu, e->d_part_dot, e->d_b0};
    return 0;
}

int do_g2p(mpm_engine *e, double facl, int diag, hipEvent_t mid, hipEvent_t mid2 = nullptr) {
    if (e->prec == MPM_PREC_FP64_EXACT) {
        ShiftF64 sh{};
        if (e->shifting) {
            int rc = shift_buffers(e, &sh);
            if (rc) return rc;
        }
        launch_f64_g2p(e->stream, e->g, e->P64, e->N64, e->lists(), e->bc(), e->mt, e->prm.scheme, e->prm.dt,
                       e->prm.mtol, facl, e->d_err, e->shifting ? &sh : nullptr, e->fbar ? e->d_njbar : nullptr,
                       f64_contrib(e));
        e->keys_valid = false;
    } else if (e->stencil) {
        StepConsts c = consts(e, facl);
        launch_s32_g2p_vel_musl(e->stream, e->g, e->P32[e->cur], e->N32, e->lists(), e->bc(), e->T, e->tg, c,
                                e->partials, diag, mid2);
        if (e->shifting) {
            ShiftF64 sh{};
            int rc = shift_buffers(e, &sh);
            if (rc) return rc;
            launch_s32_shift(e->stream, e->g, e->P32[e->cur], e->N32, e->T, sh, e->prm.mtol);
            c.sh_geup = sh.geup;
            c.sh_b0 = sh.b0;
        }
        if (e->fbar) {  // var_trans_solid_explicit.cpp:298-302
            launch_s32_fbar(e->stream, e->g, e->P32[e->cur], e->N32, e->lists(), e->T, e->tg, c, e->partials);
            c.fbar = 1;
        }
        if (mid) MPM_HIP_CHECK(hipEventRecord(mid, e->stream));
        if (!e->fused_count) e->clear_counts();
        BinFuse bf{e->fused_count ? e->bb.cnt : nullptr, e->counts_ready ? 1 : 0, e->bb.bsum, e->bb.tiles,
                   e->tg.ntiles, e->bb.ctr};
        launch_s32_g2p_update(e->stream, e->g, e->P32[e->cur], e->N32, e->T, e->mt, c, e->d_err, bf);
        if (bf.cnt && !bf.delta) launch_bin_reduce(e->stream, e->g, e->bb);
        e->jv_derived = true;
        e->counts_ready = e->fused_count;
        e->keys_valid = true;
    } else {
        StepConsts c = consts(e, facl);
        launch_f32_g2p_vel_musl(e->stream, e->g, e->P32[e->cur], e->N32, e->lists(), e->bc(), e->T, e->tg, c,
                                e->partials, diag, mid2, &e->mt);
        if (e->shifting) {
            ShiftF64 sh{};
            int rc = shift_buffers(e, &sh);
            if (rc) return rc;
            launch_s32_shift(e->stream, e->g, e->P32[e->cur], e->N32, e->T, sh, e->prm.mtol);
            c.sh_geup = sh.geup;
            c.sh_b0 = sh.b0;
        }
        if (mid) MPM_HIP_CHECK(hipEventRecord(mid, e->stream));
        SurrogateArgs sa{};
        auto &su = e->sur;
        if (su.logging && su.log_in) {
            // the sample key step_tag * n + id must fit 32 bits
            if ((uint64_t)(su.step_tag + 1) * (uint64_t)e->n <= 0xFFFFFFFFull) {
                sa.mode = 1;
                sa.kind = su.kind;
                sa.log_in = su.log_in;
                sa.log_tgt = su.log_tgt;
                sa.log_key = su.log_key;
                sa.cap = (uint32_t)su.cap;
                sa.sel_threshold = su.thr;
                sa.step_tag = su.step_tag++;
                su.logged_steps++;
            }
        } else if (su.enabled && su.have_net && su.queue && su.qcap >= sur_queue_rows(e->n)) {
            sa.mode = su.fused ? 3 : 2;
            sa.kind = su.kind;
            sa.tol = su.tol;
            sa.net = su.net.view();
            sa.net_scaling = su.net.scaling;
            sa.queue = su.queue;
            sa.bcount = su.bcount;
            sa.rstat = su.rstat;
            sa.nregions = (uint32_t)cdiv(e->n, kSurRegion);
            sa.qcap = (uint32_t)su.qcap;
        }
        sa.ctr = su.ctr;
        sa.pid = e->pid[e->cur];
        // count the new keys for the next binning: adjust kept counts of the
        // current keys, or count from zero
        if (!e->fused_count) e->clear_counts();
        BinFuse bf{e->fused_count ? e->bb.cnt : nullptr, e->counts_ready ? 1 : 0, e->bb.bsum, e->bb.tiles,
                   e->tg.ntiles, e->bb.ctr};
        launch_f32_g2p_update(e->stream, e->g, e->P32[e->cur], e->N32, e->T, e->mt, c, e->d_err, sa, bf);
        // counted from zero: the tile totals of the new counts (kept counts
        // carry theirs along)
        if (bf.cnt && !bf.delta) launch_bin_reduce(e->stream, e->g, e->bb);
        e->jv_derived = true;
        if (sa.mode == 2) launch_f32_j2_surrogate(e->stream, e->P32[e->cur], sa);
        e->counts_ready = e->fused_count;
        e->keys_valid = true;  // the update kernel writes the new keys
    }
    e->bins_valid = false;
    e->last_pass = 2;
    MPM_HIP_CHECK(hipGetLastError());
    return 0;
}

int require_particles(mpm_engine *e) {
    if (!e) return set_error(MPM_ERR_INVALID_ARGUMENT, "null engine");
    if (e->n <= 0) return set_error(MPM_ERR_INVALID_ARGUMENT, "no particles uploaded");
    return 0;
}

// F-bar arrays of one FP32 particle buffer: Fbar - I and Jbar - 1 (moved by
// the relayout), the step's increment and Jbar det dF - 1.
int alloc_f32_fbar(mpm_engine *e, PartF32 &P) {
    const size_t ld = e->ld;
    int rc = 0;
    if (!P.Hb && ((rc = e->dalloc(&P.Hb, 9 * ld, true)) || (rc = e->dalloc(&P.jb, ld, true)) ||
                  (rc = e->dalloc(&P.Dd, 9 * ld, true)) || (rc = e->dalloc(&P.jm, ld, true))))
        return rc;
    return 0;
}

// One FP32 particle buffer (slot arrays of every field the upload carries);
// buffer 1 is the target of the relayout permutation.
int alloc_f32_buffer(mpm_engine *e, int b) {
    const size_t ld = e->ld;
    PartF32 &P = e->P32[b];
    int rc = 0;
    if ((rc = e->dalloc(&P.r, 3 * ld, true)) || (rc = e->dalloc(&P.v, 3 * ld, true)) || (rc = e->dalloc(&P.m, ld, true)) ||
        (rc = e->dalloc(&P.V0, ld, true)) || (rc = e->dalloc(&P.V, ld, true)) ||
        (rc = e->dalloc(&P.V32, ld, true)) || (rc = e->dalloc(&P.J, ld, true)) ||
        (rc = e->dalloc(&P.H, 9 * ld, true)) || (rc = e->dalloc(&P.sig, 6 * ld, true)) ||
        (rc = e->dalloc(&P.key, ld, true)) || (rc = e->dalloc(&e->pid[b], ld, true)))
        return rc;
    if (e->has_trac && (rc = e->dalloc(&P.trac, 3 * ld, true))) return rc;
    if (e->has_matid && (rc = e->dalloc(&P.matid, ld, true))) return rc;
    if (e->has_epsp && (rc = e->dalloc(&P.eps_p, ld, true))) return rc;
    if (e->prm.scheme >= MPM_SCHEME_TPIC && (rc = e->dalloc(&P.C, 9 * ld, true))) return rc;
    if (e->prm.scheme == MPM_SCHEME_APIC && (rc = e->dalloc(&P.Dinv, 9 * ld, true))) return rc;
    if (e->fbar && (rc = alloc_f32_fbar(e, P))) return rc;
    P.ld = ld;
    P.n = e->n;
    return 0;
}

// A relayout needs the second FP32 buffer. An engine whose particles were
// uploaded with relayout off has only the first; allocate the second before
// the interval becomes positive.
int ensure_relayout_buffer(mpm_engine *e) {
    if (e->prec != MPM_PREC_FP32 || e->n <= 0) return 0;
    const int o = 1 - e->cur;
    if (e->P32[o].r && e->pid[o]) return 0;
    return alloc_f32_buffer(e, o);
}

int alloc_particles(mpm_engine *e, int n) {
    e->free_particles();
    e->n = n;
    // rows of the structure-of-arrays fields padded to 64 entries, so that
    // every field starts on a cache-line boundary whatever n is (unaligned
    // rows cost the update kernel 12% more time)
    e->ld = ((size_t)n + 63) & ~(size_t)63;
    const size_t ld = e->ld;
    int rc = 0;
    if (e->prec == MPM_PREC_FP64_EXACT) {
        PartF64 &P = e->P64;
        if ((rc = e->dalloc(&P.x, 3 * ld, true)) || (rc = e->dalloc(&P.v, 3 * ld, true)) ||
            (rc = e->dalloc(&P.m, ld, true)) || (rc = e->dalloc(&P.V, ld, true)) ||
            (rc = e->dalloc(&P.V0, ld, true)) || (rc = e->dalloc(&P.sig, 6 * ld, true)) ||
            (rc = e->dalloc(&P.F, 9 * ld, true)) || (rc = e->dalloc(&P.J, ld, true)) ||
            (rc = e->dalloc(&P.ddg, 9 * ld, true)) || (rc = e->dalloc(&e->keys64, ld, true)))
            return rc;
        if (e->has_trac && (rc = e->dalloc(&P.trac, 3 * ld, true))) return rc;
        if (e->has_matid && (rc = e->dalloc(&P.matid, ld, true))) return rc;
        if (e->has_epsp && (rc = e->dalloc(&P.eps_p, ld, true))) return rc;
        if (e->prm.scheme >= MPM_SCHEME_TPIC && (rc = e->dalloc(&P.C, 9 * ld, true))) return rc;
        if (e->prm.scheme == MPM_SCHEME_APIC && (rc = e->dalloc(&P.Dinv, 9 * ld, true))) return rc;
        if (e->fbar && ((rc = e->dalloc(&P.Fbar, 9 * ld, true)) || (rc = e->dalloc(&P.Jbar, ld, true)) ||
                        (rc = e->dalloc(&P.detdd, ld, true))))
            return rc;
        P.key = e->keys64;
        P.ld = ld;
        P.n = n;
        // optional: a failed allocation leaves the one-pass gathers in use
        const size_t cbytes = f64_contrib_doubles(e->g, ld) * sizeof(double);
        if (hipMalloc((void **)&e->d_contrib, cbytes) == hipSuccess) {
            e->part_allocs.push_back((void *)e->d_contrib);
        } else {
            (void)hipGetLastError();
            e->d_contrib = nullptr;
        }
    } else {
        int nbuf = e->relayout_every > 0 ? 2 : 1;
        for (int b = 0; b < nbuf; b++)
            if ((rc = alloc_f32_buffer(e, b))) return rc;
        e->cur = 0;
        if ((rc = e->dalloc(&e->perm_id, ld, true))) return rc;
    }
    if ((rc = e->dalloc(&e->bb.perm, ld, true))) return rc;
    return 0;
}

template <typename T>
int h2d(mpm_engine *e, T *dst, const std::vector<T> &src) {
    if (src.empty()) return 0;
    MPM_HIP_CHECK(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, e->stream));
    return 0;
}

template <typename T>
int d2h(mpm_engine *e, std::vector<T> &dst, const T *src, size_t count) {
    dst.resize(count);
    if (count == 0) return 0;
    MPM_HIP_CHECK(hipMemcpyAsync(dst.data(), src, count * sizeof(T), hipMemcpyDeviceToHost, e->stream));
    return 0;
}

int validate_grid(const mpm_grid_desc *g, bool allow_quadratic) {
    for (int i = 0; i < 3; i++) {
        if (g->nelem[i] <= 0) return set_error(MPM_ERR_INVALID_ARGUMENT, "nelem must be positive");
        if (!(g->xmax[i] > g->xmin[i])) return set_error(MPM_ERR_INVALID_ARGUMENT, "xmax must exceed xmin");
        if (g->idimc[i] != 1 && !(allow_quadratic && g->idimc[i] == 2))
            return set_error(MPM_ERR_UNSUPPORTED, allow_quadratic ? "shape-function order must be 1 or 2"
                                                                  : "shape-function order must be 1 here");
    }
    double ne = (double)g->nelem[0] * g->nelem[1] * g->nelem[2];
    double nn = (double)(g->nelem[0] + g->idimc[0]) * (g->nelem[1] + g->idimc[1]) * (g->nelem[2] + g->idimc[2]);
    if (ne >= 2147483647.0 || nn >= 2147483647.0) return set_error(MPM_ERR_INVALID_ARGUMENT, "grid too large");
    return 0;
}

GridGeom make_geom(const mpm_grid_desc *gd) {
    GridGeom g{};
    for (int i = 0; i < 3; i++) {
        g.xmin[i] = gd->xmin[i];
        g.xmax[i] = gd->xmax[i];
        g.dx[i] = (gd->xmax[i] - gd->xmin[i]) / (double)gd->nelem[i];  // datain_para.cpp:87
        g.nex[i] = gd->nelem[i];
        g.idimc[i] = gd->idimc[i];
        g.nxc[i] = gd->nelem[i] + gd->idimc[i];
    }
    g.nelem = g.nex[0] * g.nex[1] * g.nex[2];
    g.nodec = g.nxc[0] * g.nxc[1] * g.nxc[2];
    return g;
}

}  // namespace

extern "C" {

const char *mpm_version(void) { return "mpm-mi355x 0.1.0 gfx950"; }

const char *mpm_last_error(void) { return g_last_error.c_str(); }

int mpm_engine_create(const mpm_grid_desc *grid, const mpm_material_desc *mats, int32_t nmat,
                      const mpm_step_params *params, mpm_engine **out) try {
    if (!grid || !params || !out || (nmat > 0 && !mats))
        return set_error(MPM_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (nmat < 1 || nmat > kMaxMat) return set_error(MPM_ERR_INVALID_ARGUMENT, "nmat must be in 1..1024");
    int rc = validate_grid(grid, true);
    if (rc) return rc;
    if (!(params->dt > 0.0)) return set_error(MPM_ERR_INVALID_ARGUMENT, "dt must be positive");
    if (params->scheme < MPM_SCHEME_PIC || params->scheme > MPM_SCHEME_APIC)
        return set_error(MPM_ERR_INVALID_ARGUMENT, "unknown transfer scheme");
    {
        const GridGeom g0 = make_geom(grid);
        if (params->precision == MPM_PREC_FP32 && !s32_supported(g0))
            return set_error(MPM_ERR_UNSUPPORTED,
                             "the FP32 engine takes idimc (1,1,1), (2,2,2) or (2,1,2); the FP64-exact engine any of 1 "
                             "and 2 per axis");
    }
    if (params->precision != MPM_PREC_FP64_EXACT && params->precision != MPM_PREC_FP32)
        return set_error(MPM_ERR_INVALID_ARGUMENT, "unknown precision");
    if (params->nlstep < 1) return set_error(MPM_ERR_INVALID_ARGUMENT, "nlstep must be >= 1");
    for (int i = 0; i < nmat; i++) {
        int m = mats[i].model;
        if (!(m == -1 || m == 0 || m == 1 || m == 2 || m == 3 || m == 10 || m == 20))
            return set_error(MPM_ERR_UNSUPPORTED, "unknown constitutive model");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(MPM_ERR_NO_DEVICE, "no HIP device");

    mpm_engine *e = new mpm_engine();
    e->g = make_geom(grid);
    e->prm = *params;
    e->prec = params->precision;
    e->stencil = e->prec == MPM_PREC_FP32 &&
                 (grid->idimc[0] != 1 || grid->idimc[1] != 1 || grid->idimc[2] != 1 || params->scheme >= MPM_SCHEME_TPIC);
    e->stencil_base = e->stencil;
    e->relayout_every = params->precision == MPM_PREC_FP32 ? std::max(0, (int)params->relayout_every) : 0;
    e->mt.nmat = nmat;
    std::vector<MaterialDev> mh(nmat);
    for (int i = 0; i < nmat; i++) {
        mh[i].model = mats[i].model;
        mh[i].dt = 0.f;
        for (int k = 0; k < 8; k++) mh[i].prop[k] = mats[i].prop[k];
    }
    hipError_t he = hipGetDevice(&e->device);
    {
        MaterialDev *dm = nullptr;
        int r = e->dalloc(&dm, nmat);
        if (r) {
            delete e;
            return r;
        }
        he = hipMemcpy(dm, mh.data(), sizeof(MaterialDev) * nmat, hipMemcpyHostToDevice);
        e->mt.m = dm;
        e->mat_host = mh;
    }
    if (he == hipSuccess) he = hipGetDevice(&e->device);
    if (he == hipSuccess) he = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (he != hipSuccess) {
        int r = hip_fail(he, "hipStreamCreate", __FILE__, __LINE__);
        delete e;
        return r;
    }
    const GridGeom &g = e->g;
    const size_t nodec = (size_t)g.nodec, nelem = (size_t)g.nelem;
    e->tg = make_tile_geom(e->g);
    BinBuffers &b = e->bb;
    const size_t ncap = bin_cell_capacity(e->g.nelem);
    if ((rc = e->dalloc(&e->d_bcmask, nodec)) || (rc = e->dalloc(&e->d_bcval, 3 * nodec)) ||
        (rc = e->dalloc(&b.cnt, ncap)) || (rc = e->dalloc(&b.cstart, ncap)) || (rc = e->dalloc(&b.cend, ncap)) ||
        (rc = e->dalloc(&b.active, nelem)) || (rc = e->dalloc(&b.astart, nelem + 1)) ||
        (rc = e->dalloc(&b.stamp, nelem)) ||
        (rc = e->dalloc(&b.tiles, (size_t)e->tg.ntiles)) ||
        (rc = e->dalloc(&b.bsum, bin_scan_blocks(e->g.nelem))) ||
        (rc = e->dalloc(&b.tprev, bin_scan_blocks(e->g.nelem))) || (rc = e->dalloc(&b.ctr, 1)) ||
        (rc = e->dalloc(&b.occ, ncap / 32)) || (rc = e->dalloc(&b.gs, ncap / 32)) ||
        (rc = e->dalloc(&b.gs2, ncap / 32)) ||
        (rc = e->dalloc(&e->d_err, 1))) {
        delete e;
        return rc;
    }
    if (bin_scan_blocks(e->g.nelem) > 4096) {
        delete e;
        return set_error(MPM_ERR_INVALID_ARGUMENT, "grid too large for the binning scan (> 16.7M cells)");
    }
    he = hipMemset(b.cnt, 0, ncap * sizeof(uint32_t));
    if (he == hipSuccess) he = hipMemset(b.stamp, 0, nelem * sizeof(uint32_t));
    if (he == hipSuccess) he = hipMemset(b.cstart, 0, ncap * sizeof(uint32_t));
    if (he == hipSuccess) he = hipMemset(b.cend, 0, ncap * sizeof(uint32_t));
    if (he == hipSuccess) he = hipMemset(b.occ, 0, ncap / 32 * sizeof(uint32_t));
    if (he == hipSuccess) he = hipMemset(b.gs, 0, ncap / 32 * sizeof(uint32_t));
    if (he == hipSuccess) he = hipMemset(b.gs2, 0, ncap / 32 * sizeof(uint32_t));
    if (he == hipSuccess) he = hipMemset(b.tprev, 0, bin_scan_blocks(e->g.nelem) * sizeof(uint32_t));
    if (he == hipSuccess) he = hipMemset(b.ctr, 0, sizeof(BinCounters));
    if (he != hipSuccess) {
        int r = hip_fail(he, "hipMemset", __FILE__, __LINE__);
        delete e;
        return r;
    }
    if (e->prec == MPM_PREC_FP64_EXACT) {
        NodesF64 &N = e->N64;
        if ((rc = e->dalloc(&N.nm, nodec)) || (rc = e->dalloc(&N.nvof, nodec)) || (rc = e->dalloc(&N.np, 3 * nodec)) ||
            (rc = e->dalloc(&N.nf, 3 * nodec)) || (rc = e->dalloc(&N.nv, 3 * nodec)) ||
            (rc = e->dalloc(&N.na, 3 * nodec))) {
            delete e;
            return rc;
        }
    } else {
        NodesF32 &N = e->N32;
        if ((rc = e->dalloc(&N.nm, nodec)) || (rc = e->dalloc(&N.nvof, nodec)) || (rc = e->dalloc(&N.np, 3 * nodec)) ||
            (rc = e->dalloc(&N.nf, 3 * nodec)) || (rc = e->dalloc(&N.nv, 3 * nodec)) ||
            (rc = e->dalloc(&N.na, 3 * nodec)) || (rc = e->dalloc(&N.nvm, 3 * nodec)) ||
            (rc = e->dalloc(&N.nvp, 3 * nodec)) ||
            (rc = e->dalloc(&e->partials, (size_t)(e->stencil ? std::max(kPartialsPerCell, s32_partials_per_cell(g))
                                                                  : kPartialsPerCell) *
                                              nelem))) {
            delete e;
            return rc;
        }
        for (int a = 0; a < 3; a++) {
            e->T.inv_dx[a] = (float)(1.0 / g.dx[a]);
            e->T.dx[a] = (float)g.dx[a];
        }
        e->T.inv_nex = (float)(1.0 / (double)g.nex[0]);
        e->T.inv_nexy = (float)(1.0 / ((double)g.nex[0] * g.nex[1]));
    }
    *out = e;
    return 0;
} catch (...) {
    return mpm::exception_status();
}

// Sums the per-region network tallies (evaluated, rejected) on the host.
static int sur_region_tally(mpm_engine *e, int64_t out[2]) {
    auto &su = e->sur;
    out[0] = su.stat_base[0];
    out[1] = su.stat_base[1];
    if (!su.rstat) return 0;
    const size_t nr = (size_t)(su.qcap / kSurRegion);
    std::vector<uint32_t> h(2 * nr);
    MPM_HIP_CHECK(hipMemcpyAsync(h.data(), su.rstat, h.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, e->stream));
    MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    for (size_t i = 0; i < nr; i++) {
        out[0] += h[2 * i];
        out[1] += h[2 * i + 1];
    }
    return 0;
}

// (Re)allocates the surrogate queue, its per-region counts and tallies for n
// particles; tallies of the old regions are folded into the base counts.
static int sur_alloc_queue(mpm_engine *e, int64_t n) {
    auto &su = e->sur;
    const int64_t rows = sur_queue_rows(n);
    if (su.qcap >= rows && su.bcount && su.rstat) return 0;
    int64_t t[2];
    int rc = sur_region_tally(e, t);
    if (rc) return rc;
    su.stat_base[0] = t[0];
    su.stat_base[1] = t[1];
    for (void *p : {(void *)su.queue, (void *)su.bcount, (void *)su.rstat})
        if (p) (void)hipFree(p);
    su.queue = nullptr;
    su.bcount = su.rstat = nullptr;
    su.qcap = 0;
    const size_t nr = (size_t)(rows / kSurRegion);
    MPM_HIP_CHECK(hipMalloc((void **)&su.queue, (size_t)rows * kSurQueueRow * sizeof(float)));
    MPM_HIP_CHECK(hipMalloc((void **)&su.bcount, nr * sizeof(uint32_t)));
    MPM_HIP_CHECK(hipMalloc((void **)&su.rstat, 2 * nr * sizeof(uint32_t)));
    MPM_HIP_CHECK(hipMemsetAsync(su.rstat, 0, 2 * nr * sizeof(uint32_t), e->stream));
    MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    su.qcap = rows;
    return 0;
}

static int sur_require(mpm_engine *e) {
    if (!e) return set_error(MPM_ERR_INVALID_ARGUMENT, "null engine");
    if (e->prec != MPM_PREC_FP32)
        return set_error(MPM_ERR_UNSUPPORTED, "the constitutive surrogate runs in the FP32 engine only");
    if (e->stencil)
        return set_error(MPM_ERR_UNSUPPORTED,
                         "the constitutive surrogate runs with the tent functions and PIC / FLIP only");
    if (!e->sur.ctr) {
        MPM_HIP_CHECK(hipMalloc((void **)&e->sur.ctr, 4 * sizeof(unsigned long long)));
        MPM_HIP_CHECK(hipMemsetAsync(e->sur.ctr, 0, 4 * sizeof(unsigned long long), e->stream));
    }
    if (!e->sur.net.master) {
        int rc = e->sur.net.init();
        if (rc) return rc;
    }
    // J2 plasticity if any material has it, else the Tait pressure closure
    bool j2 = false, tait = false;
    for (const MaterialDev &m : e->mat_host) {
        j2 = j2 || m.model == MPM_MODEL_J2;
        tait = tait || m.model == MPM_MODEL_TAIT;
    }
    if (!j2 && !tait)
        return set_error(MPM_ERR_UNSUPPORTED, "the surrogate stands in for J2 plasticity or the Tait pressure");
    e->sur.kind = j2 ? kSurJ2 : kSurTait;
    return 0;
}

int mpm_surrogate_collect(mpm_engine *e, int32_t on, int64_t capacity, double rate) try {
    int rc = sur_require(e);
    if (rc) return rc;
    e->drop_graph();
    auto &su = e->sur;
    MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    if (!on) {
        su.logging = false;
        return 0;
    }
    if (capacity <= 0 || capacity >= (1ll << 31) || !(rate > 0.0))
        return set_error(MPM_ERR_INVALID_ARGUMENT, "capacity must be in (0, 2^31) and rate > 0");
    if (capacity != su.cap) {
        su.free_log();
        MPM_HIP_CHECK(hipMalloc((void **)&su.log_in, (size_t)capacity * 16 * sizeof(float)));
        MPM_HIP_CHECK(hipMalloc((void **)&su.log_tgt, (size_t)capacity * 16 * sizeof(float)));
        MPM_HIP_CHECK(hipMalloc((void **)&su.log_key, (size_t)capacity * sizeof(uint32_t)));
        su.cap = capacity;
    }
    MPM_HIP_CHECK(hipMemsetAsync(su.ctr, 0, 4 * sizeof(unsigned long long), e->stream));
    su.thr = rate >= 1.0 ? 0xFFFFFFFFu : (uint32_t)(rate * 4294967296.0);
    su.step_tag = 0;
    su.logged_steps = 0;
    su.nsamples = 0;
    su.sorted = false;
    su.logging = true;
    return 0;
} catch (...) {
    return mpm::exception_status();
}

static int sur_samples(mpm_engine *e, int64_t *n) {
    unsigned long long c[4];
    MPM_HIP_CHECK(hipMemcpyAsync(c, e->sur.ctr, sizeof c, hipMemcpyDeviceToHost, e->stream));
    MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    *n = (int64_t)std::min<unsigned long long>(c[2], (unsigned long long)e->sur.cap);
    return 0;
}

// order the logged samples by (step, particle id): the log is filled with
// atomics, so its raw order varies from run to run
static int sur_sort(mpm_engine *e) {
    auto &su = e->sur;
    if (su.sorted) return 0;
    int64_t n = 0;
    int rc = sur_samples(e, &n);
    if (rc) return rc;
    su.nsamples = n;
    if (n == 0) return set_error(MPM_ERR_INVALID_ARGUMENT, "no samples were logged (no plastic J2 trial states)");
    SortBuffers sb{};
    float *tmp = nullptr;
    hipError_t he = hipMalloc((void **)&sb.keys, n * sizeof(uint32_t));
    if (he == hipSuccess) he = hipMalloc((void **)&sb.vals, n * sizeof(uint32_t));
    if (he == hipSuccess) he = hipMalloc((void **)&sb.keys_alt, n * sizeof(uint32_t));
    if (he == hipSuccess) he = hipMalloc((void **)&sb.vals_alt, n * sizeof(uint32_t));
    if (he == hipSuccess) he = hipMalloc((void **)&sb.hist, radix_hist_entries((int)n) * sizeof(uint32_t));
    if (he == hipSuccess) he = hipMalloc((void **)&sb.bsum, 4096 * sizeof(uint32_t));
    if (he == hipSuccess) he = hipMalloc((void **)&tmp, (size_t)n * 16 * sizeof(float));
    uint32_t *sk = nullptr, *sv = nullptr;
    if (he == hipSuccess && radix_sort_pairs(e->stream, sb, su.log_key, (int)n, 32, &sk, &sv)) he = hipErrorUnknown;
    if (he == hipSuccess) {
        launch_gather_rows(e->stream, su.log_in, sv, n, tmp);
        he = hipMemcpyAsync(su.log_in, tmp, (size_t)n * 16 * sizeof(float), hipMemcpyDeviceToDevice, e->stream);
    }
    if (he == hipSuccess) {
        launch_gather_rows(e->stream, su.log_tgt, sv, n, tmp);
        he = hipMemcpyAsync(su.log_tgt, tmp, (size_t)n * 16 * sizeof(float), hipMemcpyDeviceToDevice, e->stream);
    }
    if (he == hipSuccess) he = hipMemcpyAsync(su.log_key, sk, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, e->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
    for (void *p : {(void *)sb.keys, (void *)sb.vals, (void *)sb.keys_alt, (void *)sb.vals_alt, (void *)sb.hist,
                    (void *)sb.bsum, (void *)tmp})
        if (p) (void)hipFree(p);
    if (he != hipSuccess) return hip_fail(he, "surrogate sample sort", __FILE__, __LINE__);
    su.sorted = true;
    return 0;
}

int mpm_surrogate_fit(mpm_engine *e, const mpm_train_params *tp, float *final_loss) try {
    int rc = sur_require(e);
    if (rc) return rc;
    e->drop_graph();
    if (!tp) return set_error(MPM_ERR_INVALID_ARGUMENT, "null training parameters");
    auto &su = e->sur;
    if (!su.log_in) return set_error(MPM_ERR_INVALID_ARGUMENT, "no sample log (call mpm_surrogate_collect)");
    su.logging = false;
    if ((rc = sur_sort(e))) return rc;
    float scaling[64];
    if ((rc = mlp_fit_scaling(su.log_in, su.log_tgt, su.nsamples, scaling, e->stream))) return rc;
    std::vector<float> params(kMlpNParams);
    mlp_init_params(params.data(), tp->seed);
    if ((rc = mlp_set(su.net, params.data(), scaling, e->stream))) return rc;
    std::vector<float> el;
    if ((rc = mlp_train(su.net, su.log_in, su.log_tgt, su.nsamples, *tp, el, e->stream))) return rc;
    su.loss = el.empty() ? 0.f : el.back();
    su.have_net = true;
    if (final_loss) *final_loss = su.loss;
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_surrogate_enable(mpm_engine *e, int32_t on, double tol) try {
    int rc = sur_require(e);
    if (rc) return rc;
    e->drop_graph();
    if (on && !e->sur.have_net) return set_error(MPM_ERR_INVALID_ARGUMENT, "no trained or loaded network");
    if (on && !(tol >= 0.0)) return set_error(MPM_ERR_INVALID_ARGUMENT, "tolerance must be >= 0");
    MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    e->sur.enabled = on != 0;
    if (on) {
        e->sur.tol = (float)std::min(tol, 1e30);
        e->sur.logging = false;
        rc = sur_alloc_queue(e, e->n);
        if (rc) return rc;
    }
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_surrogate_reset_counters(mpm_engine *e) try {
    int rc = sur_require(e);
    if (rc) return rc;
    unsigned long long c[4];
    MPM_HIP_CHECK(hipMemcpyAsync(c, e->sur.ctr, sizeof c, hipMemcpyDeviceToHost, e->stream));
    MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    c[0] = c[1] = 0;  // keep the log count
    MPM_HIP_CHECK(hipMemcpyAsync(e->sur.ctr, c, sizeof c, hipMemcpyHostToDevice, e->stream));
    auto &su = e->sur;
    su.stat_base[0] = su.stat_base[1] = 0;
    if (su.rstat)
        MPM_HIP_CHECK(hipMemsetAsync(su.rstat, 0, 2 * (size_t)(su.qcap / kSurRegion) * sizeof(uint32_t), e->stream));
    MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_surrogate_stats_get(mpm_engine *e, mpm_surrogate_stats *out) try {
    int rc = sur_require(e);
    if (rc) return rc;
    if (!out) return set_error(MPM_ERR_INVALID_ARGUMENT, "null output");
    unsigned long long c[4];
    MPM_HIP_CHECK(hipMemcpyAsync(c, e->sur.ctr, sizeof c, hipMemcpyDeviceToHost, e->stream));
    MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    out->samples = (int64_t)std::min<unsigned long long>(c[2], (unsigned long long)e->sur.cap);
    out->logged_steps = e->sur.logged_steps;
    int64_t t[2];
    if ((rc = sur_region_tally(e, t))) return rc;
    out->evaluated = (int64_t)c[0] + t[0];
    out->fallbacks = (int64_t)c[1] + t[1];
    out->loss = e->sur.loss;
    out->tol = e->sur.tol;
    out->offered = (int64_t)c[2];
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_surrogate_get_params(mpm_engine *e, float *params, float *scaling) try {
    int rc = sur_require(e);
    if (rc) return rc;
    if (!e->sur.have_net) return set_error(MPM_ERR_INVALID_ARGUMENT, "no trained or loaded network");
    return mlp_get(e->sur.net, params, scaling, e->stream);
} catch (...) {
    return mpm::exception_status();
}

int mpm_surrogate_set_params(mpm_engine *e, const float *params, const float *scaling) try {
    int rc = sur_require(e);
    if (rc) return rc;
    e->drop_graph();
    if (!params || !scaling) return set_error(MPM_ERR_INVALID_ARGUMENT, "null argument");
    MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    if ((rc = mlp_set(e->sur.net, params, scaling, e->stream))) return rc;
    e->sur.have_net = true;
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_surrogate_download_samples(mpm_engine *e, float *in, float *tgt, int64_t max, int64_t *n) try {
    int rc = sur_require(e);
    if (rc) return rc;
    if (!n || max < 0 || (max > 0 && (!in || !tgt))) return set_error(MPM_ERR_INVALID_ARGUMENT, "bad arguments");
    if (!e->sur.log_in) {
        *n = 0;
        return 0;
    }
    MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    if (!e->sur.logging && (rc = sur_sort(e))) return rc;
    int64_t have = 0;
    if ((rc = sur_samples(e, &have))) return rc;
    const int64_t k = std::min(have, max);
    if (k > 0) {
        MPM_HIP_CHECK(hipMemcpyAsync(in, e->sur.log_in, (size_t)k * 16 * sizeof(float), hipMemcpyDeviceToHost,
                                     e->stream));
        MPM_HIP_CHECK(hipMemcpyAsync(tgt, e->sur.log_tgt, (size_t)k * 16 * sizeof(float), hipMemcpyDeviceToHost,
                                     e->stream));
        MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    }
    *n = k;
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_set_shifting(mpm_engine *e, int32_t on, const int32_t *npxye) try {
    if (!e) return set_error(MPM_ERR_INVALID_ARGUMENT, "null engine");
    e->drop_graph();
    if (!on) {
        e->shifting = false;
        return 0;
    }
    if (!npxye) return set_error(MPM_ERR_INVALID_ARGUMENT, "null points-per-cell");
    for (int a = 0; a < 3; a++)
        if (npxye[a] < 1 || npxye[a] > 6) return set_error(MPM_ERR_INVALID_ARGUMENT, "points per cell must be 1..6");
    int rc;
    const size_t nodec = (size_t)e->g.nodec;
    if (!e->d_nvol &&
        ((rc = e->dalloc(&e->d_nvol, nodec)) || (rc = e->dalloc(&e->d_nei, nodec)) ||
         (rc = e->dalloc(&e->d_part_eu, (size_t)cdiv(e->g.nodec, kShiftBlock))) || (rc = e->dalloc(&e->d_b0, 1))))
        return rc;
    // GaussianDistribution offsets (module/mesh.cpp:12-54)
    GaussOffsets dec{};
    for (int a = 0; a < 3; a++) {
        const double h = e->g.dx[a];
        double *d = &dec.d[0][a];
        auto set = [&](int i, double v) { d[3 * i] = v; };
        switch (npxye[a]) {
        case 1: set(0, 0.0); break;
        case 2: set(0, -0.25 * h); set(1, 0.25 * h); break;
        case 3: set(0, -1.0 / 3.0 * h); set(1, 0.0); set(2, 1.0 / 3.0 * h); break;
        case 4: set(0, -0.375 * h); set(1, -0.125 * h); set(2, 0.125 * h); set(3, 0.375 * h); break;
        case 5: set(0, -0.4 * h); set(1, -0.2 * h); set(2, 0.0); set(3, 0.2 * h); set(4, 0.4 * h); break;
        default:
            set(0, -5.0 / 12.0 * h); set(1, -3.0 / 12.0 * h); set(2, -1.0 / 12.0 * h);
            set(3, 1.0 / 12.0 * h); set(4, 3.0 / 12.0 * h); set(5, 5.0 / 12.0 * h);
            break;
        }
    }
    const double volp = e->g.dx[0] * e->g.dx[1] * e->g.dx[2] / (double)(npxye[0] * npxye[1] * npxye[2]);
    launch_f64_nodal_vol(e->stream, e->g, npxye[0], npxye[1], npxye[2], dec, volp, e->d_nvol);
    MPM_HIP_CHECK(hipGetLastError());
    MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    e->shifting = true;
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_download_nodal_volume(mpm_engine *e, double *nvol) try {
    if (!e || !nvol) return set_error(MPM_ERR_INVALID_ARGUMENT, "null argument");
    if (!e->d_nvol) return set_error(MPM_ERR_INVALID_ARGUMENT, "no nodal volume (call mpm_set_shifting)");
    MPM_HIP_CHECK(hipMemcpyAsync(nvol, e->d_nvol, sizeof(double) * e->g.nodec, hipMemcpyDeviceToHost, e->stream));
    MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    return 0;
} catch (...) {
    return mpm::exception_status();
}

// F-bar state at the start of a run: Fbar = I, Jbar = 1 (datain_para.cpp:122)
static int fbar_reset(mpm_engine *e) {
    const size_t ld = e->ld;
    std::vector<double> F(9 * ld, 0.0), J(ld, 1.0);
    for (size_t i = 0; i < ld; i++) F[0 * ld + i] = F[4 * ld + i] = F[8 * ld + i] = 1.0;
    int rc;
    if ((rc = h2d(e, e->P64.Fbar, F)) || (rc = h2d(e, e->P64.Jbar, J))) return rc;
    MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    return 0;
}

// FP32 F-bar state at the start of a run: Fbar - I = 0, Jbar - 1 = 0, in
// the current particle buffer
static int fbar_reset32(mpm_engine *e) {
    const PartF32 &P = e->P32[e->cur];
    MPM_HIP_CHECK(hipMemsetAsync(P.Hb, 0, 9 * e->ld * sizeof(float), e->stream));
    MPM_HIP_CHECK(hipMemsetAsync(P.jb, 0, e->ld * sizeof(float), e->stream));
    MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    return 0;
}

int mpm_set_fbar(mpm_engine *e, int32_t on) try {
    if (!e) return set_error(MPM_ERR_INVALID_ARGUMENT, "null engine");
    if (on && e->prec == MPM_PREC_FP32 && (e->sur.enabled || e->sur.logging || e->mat_dt))
        return set_error(MPM_ERR_UNSUPPORTED,
                         "the FP32 F-bar projection runs without the constitutive surrogate and per-material time steps");
    e->drop_graph();
    e->bins_valid = false;  // the kernels (and their cell groups) may change
    if (!on) {
        e->fbar = false;
        e->stencil = e->stencil_base;
        return 0;
    }
    int rc;
    if (e->prec == MPM_PREC_FP32) {
        // the stencil kernels carry the projection (the tent functions too)
        if (!e->N32.nje && (rc = e->dalloc(&e->N32.nje, (size_t)e->g.nodec))) return rc;
        e->fbar = true;
        e->stencil = true;
        if (e->n > 0) {
            for (int b = 0; b < 2; b++)
                if (e->P32[b].r && (rc = alloc_f32_fbar(e, e->P32[b]))) return rc;
            if ((rc = fbar_reset32(e))) return rc;
        }
        return 0;
    }
    if (!e->d_njbar && (rc = e->dalloc(&e->d_njbar, (size_t)e->g.nodec))) return rc;
    e->fbar = true;
    if (e->n > 0) {
        PartF64 &P = e->P64;
        const size_t ld = e->ld;
        if (!P.Fbar && ((rc = e->dalloc(&P.Fbar, 9 * ld, true)) || (rc = e->dalloc(&P.Jbar, ld, true)) ||
                        (rc = e->dalloc(&P.detdd, ld, true))))
            return rc;
        if ((rc = fbar_reset(e))) return rc;
    }
    return 0;
} catch (...) {
    return mpm::exception_status();
}

// FP32: slot s holds particle pid[s]; Fbar = I + Hb, Jbar = 1 + jb.
static int fbar_copy32(mpm_engine *e, double *fbar, double *jbar, bool to_device) {
    int rc = 0;
    const PartF32 &P = e->P32[e->cur];
    if (!P.Hb) return set_error(MPM_ERR_INVALID_ARGUMENT, "F-bar is off");
    const size_t ld = e->ld;
    const int n = e->n;
    std::vector<int> pidv;
    std::vector<float> hb, jb;
    if ((rc = d2h(e, pidv, e->pid[e->cur], ld)) || (rc = d2h(e, hb, (const float *)P.Hb, 9 * ld)) ||
        (rc = d2h(e, jb, (const float *)P.jb, ld)))
        return rc;
    MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    for (int s = 0; s < n; s++) {
        const size_t i = (size_t)pidv[s];
        for (int q = 0; q < 9; q++) {
            const double eye = (q == 0 || q == 4 || q == 8) ? 1.0 : 0.0;
            if (fbar && to_device) hb[q * ld + s] = (float)(fbar[9 * i + q] - eye);
            if (fbar && !to_device) fbar[9 * i + q] = (double)hb[q * ld + s] + eye;
        }
        if (jbar && to_device) jb[s] = (float)(jbar[i] - 1.0);
        if (jbar && !to_device) jbar[i] = 1.0 + (double)jb[s];
    }
    if (to_device) {
        if ((fbar && (rc = h2d(e, P.Hb, hb))) || (jbar && (rc = h2d(e, P.jb, jb)))) return rc;
        MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    }
    return 0;
}

static int fbar_copy(mpm_engine *e, double *fbar, double *jbar, bool to_device) {
    int rc = require_particles(e);
    if (rc) return rc;
    if (e->prec == MPM_PREC_FP32) return fbar_copy32(e, fbar, jbar, to_device);
    PartF64 &P = e->P64;
    if (e->prec != MPM_PREC_FP64_EXACT || !P.Fbar) return set_error(MPM_ERR_INVALID_ARGUMENT, "F-bar is off");
    const size_t ld = e->ld;
    const int n = e->n;
    if (fbar) {
        std::vector<double> soa(9 * ld);
        if (to_device) {
            for (int i = 0; i < n; i++)
                for (int q = 0; q < 9; q++) soa[q * ld + i] = fbar[9 * (size_t)i + q];
            if ((rc = h2d(e, P.Fbar, soa))) return rc;
        } else {
            if ((rc = d2h(e, soa, (const double *)P.Fbar, 9 * ld))) return rc;
        }
        MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
        if (!to_device)
            for (int i = 0; i < n; i++)
                for (int q = 0; q < 9; q++) fbar[9 * (size_t)i + q] = soa[q * ld + i];
    }
    if (jbar) {
        if (to_device)
            MPM_HIP_CHECK(hipMemcpyAsync(P.Jbar, jbar, n * sizeof(double), hipMemcpyHostToDevice, e->stream));
        else
            MPM_HIP_CHECK(hipMemcpyAsync(jbar, P.Jbar, n * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    }
    return 0;
}

int mpm_download_fbar(mpm_engine *e, double *def_grad_bar, double *det_def_grad_bar) try {
    return fbar_copy(e, def_grad_bar, det_def_grad_bar, false);
} catch (...) {
    return mpm::exception_status();
}

int mpm_upload_fbar(mpm_engine *e, const double *def_grad_bar, const double *det_def_grad_bar) try {
    if (e) e->drop_graph();
    return fbar_copy(e, const_cast<double *>(def_grad_bar), const_cast<double *>(det_def_grad_bar), true);
} catch (...) {
    return mpm::exception_status();
}

// TPIC velocity gradient / APIC B matrix and APIC inverse D, host [n][3][3]
// FP32 engine: slot arrays in the engine's current particle order, host
// arrays in the original particle order (slot s holds particle pid[s]).
static int affine_copy_f32(mpm_engine *e, double *cmat, double *dinv, bool to_device) {
    const PartF32 &P = e->P32[e->cur];
    if (!P.C) return set_error(MPM_ERR_INVALID_ARGUMENT, "the engine's transfer scheme is not TPIC or APIC");
    if (dinv && !P.Dinv) return set_error(MPM_ERR_INVALID_ARGUMENT, "inverse D exists under APIC only");
    const size_t ld = e->ld;
    const int n = e->n;
    int rc = 0;
    std::vector<int> pidv;
    if ((rc = d2h(e, pidv, (const int *)e->pid[e->cur], ld))) return rc;
    MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    double *host[2] = {cmat, dinv};
    float *dev[2] = {P.C, P.Dinv};
    for (int b = 0; b < 2; b++) {
        if (!host[b]) continue;
        std::vector<float> soa(9 * ld, 0.f);
        if (to_device) {
            for (int s = 0; s < n; s++)
                for (int q = 0; q < 9; q++) soa[q * ld + s] = (float)host[b][9 * (size_t)pidv[s] + q];
            if ((rc = h2d(e, dev[b], soa))) return rc;
            MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
        } else {
            if ((rc = d2h(e, soa, (const float *)dev[b], 9 * ld))) return rc;
            MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
            for (int s = 0; s < n; s++)
                for (int q = 0; q < 9; q++) host[b][9 * (size_t)pidv[s] + q] = (double)soa[q * ld + s];
        }
    }
    return 0;
}

static int affine_copy(mpm_engine *e, double *cmat, double *dinv, bool to_device) {
    int rc = require_particles(e);
    if (rc) return rc;
    if (e->prec == MPM_PREC_FP32) return affine_copy_f32(e, cmat, dinv, to_device);
    const PartF64 &P = e->P64;
    if (e->prec != MPM_PREC_FP64_EXACT || !P.C)
        return set_error(MPM_ERR_INVALID_ARGUMENT, "the engine's transfer scheme is not TPIC or APIC");
    if (dinv && !P.Dinv) return set_error(MPM_ERR_INVALID_ARGUMENT, "inverse D exists under APIC only");
    const size_t ld = e->ld;
    const int n = e->n;
    double *host[2] = {cmat, dinv};
    double *dev[2] = {P.C, P.Dinv};
    for (int b = 0; b < 2; b++) {
        if (!host[b]) continue;
        std::vector<double> soa(9 * ld);
        if (to_device) {
            for (int i = 0; i < n; i++)
                for (int q = 0; q < 9; q++) soa[q * ld + i] = host[b][9 * (size_t)i + q];
            if ((rc = h2d(e, dev[b], soa))) return rc;
            MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
        } else {
            if ((rc = d2h(e, soa, (const double *)dev[b], 9 * ld))) return rc;
            MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
            for (int i = 0; i < n; i++)
                for (int q = 0; q < 9; q++) host[b][9 * (size_t)i + q] = soa[q * ld + i];
        }
    }
    return 0;
}

int mpm_download_affine(mpm_engine *e, double *cmat, double *dinv) try {
    return affine_copy(e, cmat, dinv, false);
} catch (...) {
    return mpm::exception_status();
}

int mpm_upload_affine(mpm_engine *e, const double *cmat, const double *dinv) try {
    if (e) e->drop_graph();
    return affine_copy(e, const_cast<double *>(cmat), const_cast<double *>(dinv), true);
} catch (...) {
    return mpm::exception_status();
}

int mpm_get_engine_info(mpm_engine *e, mpm_engine_info *out) try {
    if (!e || !out) return set_error(MPM_ERR_INVALID_ARGUMENT, "null argument");
    out->n = e->n;
    out->nelem = e->g.nelem;
    out->nodec = e->g.nodec;
    out->precision = e->prec;
    out->scheme = e->prm.scheme;
    out->shifting = e->shifting ? 1 : 0;
    out->fbar = e->fbar ? 1 : 0;
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_engine_destroy(mpm_engine *eng) try {
    if (!eng) return 0;
    (void)hipStreamSynchronize(eng->stream);
    delete eng;
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_set_bc(mpm_engine *e, int32_t comp, const int32_t *nbc, const double *fbc, int32_t ibc) try {
    if (!e || comp < 0 || comp > 3 || ibc < 0 || (ibc > 0 && !nbc))
        return set_error(MPM_ERR_INVALID_ARGUMENT, "bad BC arguments");
    for (int i = 0; i < ibc; i++)
        if (nbc[i] < 0 || nbc[i] >= e->g.nodec) return set_error(MPM_ERR_INVALID_ARGUMENT, "BC node out of range");
    e->nbc[comp].assign(nbc, nbc + ibc);
    if (fbc)
        e->fbc[comp].assign(fbc, fbc + ibc);
    else
        e->fbc[comp].assign(ibc, 0.0);
    e->bc_dirty = true;
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_upload_particles(mpm_engine *e, const mpm_particles_view *p) try {
    if (e) e->drop_graph();
    if (!e || !p || p->n <= 0 || !p->coord || !p->mass || !p->vol0)
        return set_error(MPM_ERR_INVALID_ARGUMENT, "particle view needs n > 0, coord, mass, vol0");
    if (p->n >= 2147483647LL) return set_error(MPM_ERR_INVALID_ARGUMENT, "too many particles");
    const int n = (int)p->n;
    if (p->matid)
        for (int i = 0; i < n; i++)
            if (p->matid[i] < 0 || p->matid[i] >= e->mt.nmat)
                return set_error(MPM_ERR_INVALID_ARGUMENT, "matid out of range");
    e->has_trac = p->trac_force != nullptr;
    e->has_matid = p->matid != nullptr;
    e->has_epsp = true;
    int rc = alloc_particles(e, n);
    if (rc) return rc;
    if (e->sur.enabled) {  // the surrogate queue holds every particle
        rc = sur_alloc_queue(e, n);
        if (rc) return rc;
    }
    const size_t ld = e->ld;
    auto F_at = [&](int i, int q) -> double {
        if (p->def_grad) return p->def_grad[9 * (size_t)i + q];
        return (q == 0 || q == 4 || q == 8) ? 1.0 : 0.0;
    };
    double msum = 0.0;
    for (int i = 0; i < n; i++) msum += p->mass[i];
    e->m_small = (float)(1e-2 * msum / (double)n);
    {
        // particles per non-empty cell picks how many lanes fold one cell: a
        // 256-particle LDS chunk should cover the 64 / L cells of a group
        const GridGeom &g = e->g;
        std::vector<uint8_t> seen((size_t)g.nelem, 0);
        long long nact = 0;
        for (int i = 0; i < n; i++) {
            double q[3];
            int ok = 1, idx[3];
            for (int a = 0; a < 3; a++) {
                q[a] = std::floor((p->coord[3 * (size_t)i + a] - g.xmin[a]) / g.dx[a]);
                if (!(q[a] >= 0.0 && q[a] < (double)g.nex[a])) ok = 0;
                idx[a] = ok ? (int)q[a] : 0;
            }
            if (!ok) continue;
            size_t c = (size_t)idx[0] + (size_t)g.nex[0] * ((size_t)idx[1] + (size_t)g.nex[1] * idx[2]);
            if (!seen[c]) {
                seen[c] = 1;
                nact++;
            }
        }
        double ppc = nact > 0 ? (double)n / (double)nact : 1.0;
        int lpc = 1;
        while (lpc < 8 && 64.0 * ppc / lpc > 512.0) lpc *= 2;
        e->lanes_per_cell = lpc;
    }
    if (e->prec == MPM_PREC_FP64_EXACT) {
        std::vector<double> x(3 * ld), v(3 * ld, 0.0), sig(6 * ld, 0.0), F(9 * ld), m(ld), V(ld), V0(ld), J(ld),
            tr, ep(ld, 0.0);
        for (int i = 0; i < n; i++) {
            for (int a = 0; a < 3; a++) {
                x[a * ld + i] = p->coord[3 * (size_t)i + a];
                if (p->vel) v[a * ld + i] = p->vel[3 * (size_t)i + a];
            }
            if (p->stress)
                for (int q = 0; q < 6; q++) sig[q * ld + i] = p->stress[6 * (size_t)i + q];
            for (int q = 0; q < 9; q++) F[q * ld + i] = F_at(i, q);
            m[i] = p->mass[i];
            V0[i] = p->vol0[i];
            V[i] = p->vol ? p->vol[i] : p->vol0[i];
            J[i] = p->det_def_grad ? p->det_def_grad[i] : 1.0;
            if (p->eps_p) ep[i] = p->eps_p[i];
        }
        PartF64 &P = e->P64;
        if ((rc = h2d(e, P.x, x)) || (rc = h2d(e, P.v, v)) || (rc = h2d(e, P.sig, sig)) || (rc = h2d(e, P.F, F)) ||
            (rc = h2d(e, P.m, m)) || (rc = h2d(e, P.V, V)) || (rc = h2d(e, P.V0, V0)) || (rc = h2d(e, P.J, J)) ||
            (rc = h2d(e, P.eps_p, ep)))
            return rc;
        if (P.Fbar && (rc = fbar_reset(e))) return rc;
        // TPIC/APIC matrices start at zero (solid_material_point.cpp:36-43)
        if (P.C || P.Dinv) {
            std::vector<double> zero(9 * ld, 0.0);
            if ((P.C && (rc = h2d(e, P.C, zero))) || (P.Dinv && (rc = h2d(e, P.Dinv, zero)))) return rc;
            MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
        }
        if (e->has_trac) {
            tr.assign(3 * ld, 0.0);
            for (int i = 0; i < n; i++)
                for (int a = 0; a < 3; a++) tr[a * ld + i] = p->trac_force[3 * (size_t)i + a];
            if ((rc = h2d(e, P.trac, tr))) return rc;
        }
        if (e->has_matid) {
            std::vector<int> mi(p->matid, p->matid + n);
            if ((rc = h2d(e, P.matid, mi))) return rc;
        }
    } else {
        std::vector<float> rr(3 * ld), v(3 * ld, 0.f), sig(6 * ld, 0.f), H(9 * ld), m(ld), V0(ld), V32(ld), ep(ld, 0.f),
            tr;
        std::vector<double> V(ld), J(ld);
        std::vector<int> pidv(ld);
        std::vector<uint32_t> keys(ld);
        const GridGeom &g = e->g;
        int outside = 0;
        for (int i = 0; i < n; i++) {
            // cell with the reference's FP64 formula (material_point.cpp:59-65), offset inside it
            int ijk[3];
            bool ok = true;
            for (int a = 0; a < 3; a++) {
                double xv = p->coord[3 * (size_t)i + a];
                double q = std::floor((xv - g.xmin[a]) / g.dx[a]);
                if (!(q >= 0.0 && q < (double)g.nex[a])) {
                    ok = false;
                    q = 0.0;
                }
                ijk[a] = (int)q;
                float ra = (float)((xv - (g.xmin[a] + g.dx[a] * (double)ijk[a])) / g.dx[a]);
                if (ra >= 1.f) ra = std::nextafter(1.f, 0.f);
                if (ra < 0.f) ra = 0.f;
                rr[a * ld + i] = ra;
                if (p->vel) v[a * ld + i] = (float)p->vel[3 * (size_t)i + a];
            }
            if (!ok) outside++;
            keys[i] = ok ? (uint32_t)(ijk[0] + ijk[1] * g.nex[0] + ijk[2] * g.nex[0] * g.nex[1]) : (uint32_t)g.nelem;
            if (p->stress)
                for (int q = 0; q < 6; q++) sig[q * ld + i] = (float)p->stress[6 * (size_t)i + q];
            for (int q = 0; q < 9; q++)
                H[q * ld + i] = (float)(F_at(i, q) - ((q == 0 || q == 4 || q == 8) ? 1.0 : 0.0));
            m[i] = (float)p->mass[i];
            V0[i] = (float)p->vol0[i];
            V[i] = p->vol ? p->vol[i] : p->vol0[i];
            V32[i] = (float)V[i];
            J[i] = p->det_def_grad ? p->det_def_grad[i] : 1.0;
            if (p->eps_p) ep[i] = (float)p->eps_p[i];
            pidv[i] = i;
        }
        PartF32 &P = e->P32[0];
        if ((rc = h2d(e, P.r, rr)) || (rc = h2d(e, P.v, v)) || (rc = h2d(e, P.key, keys)) ||
            (rc = h2d(e, P.sig, sig)) || (rc = h2d(e, P.H, H)) || (rc = h2d(e, P.m, m)) || (rc = h2d(e, P.V0, V0)) ||
            (rc = h2d(e, P.V, V)) || (rc = h2d(e, P.V32, V32)) || (rc = h2d(e, P.J, J)) ||
            (rc = h2d(e, P.eps_p, ep)) || (rc = h2d(e, e->pid[0], pidv)))
            return rc;
        e->fp32_outside_at_upload = outside;
        // TPIC/APIC matrices start at zero (solid_material_point.cpp:36-43)
        if (P.C || P.Dinv) {
            std::vector<float> zero(9 * ld, 0.f);
            if ((P.C && (rc = h2d(e, P.C, zero))) || (P.Dinv && (rc = h2d(e, P.Dinv, zero)))) return rc;
            MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
        }
        if (P.Hb && (rc = fbar_reset32(e))) return rc;
        if (e->has_trac) {
            tr.assign(3 * ld, 0.f);
            for (int i = 0; i < n; i++)
                for (int a = 0; a < 3; a++) tr[a * ld + i] = (float)p->trac_force[3 * (size_t)i + a];
            if ((rc = h2d(e, P.trac, tr))) return rc;
        }
        if (e->has_matid) {
            std::vector<int> mi(p->matid, p->matid + n);
            if ((rc = h2d(e, P.matid, mi))) return rc;
        }
    }
    MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    e->keys_valid = e->prec == MPM_PREC_FP32;
    e->jv_derived = false;
    e->clear_counts();
    e->bins_valid = false;
    e->steps_since_relayout = 0;
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_download_particles(mpm_engine *e, mpm_particles_view *p) try {
    int rc = require_particles(e);
    if (rc) return rc;
    if (!p || p->n != e->n) return set_error(MPM_ERR_INVALID_ARGUMENT, "view size differs from the engine's");
    const int n = e->n;
    const size_t ld = e->ld;
    if (e->prec == MPM_PREC_FP64_EXACT) {
        const PartF64 &P = e->P64;
        std::vector<double> x, v, sig, F, V, J, ep, m, V0, tr;
        std::vector<int> mid;
        if ((rc = d2h(e, x, P.x, 3 * ld)) || (rc = d2h(e, v, P.v, 3 * ld)) || (rc = d2h(e, sig, P.sig, 6 * ld)) ||
            (rc = d2h(e, F, P.F, 9 * ld)) || (rc = d2h(e, V, P.V, ld)) || (rc = d2h(e, J, P.J, ld)) ||
            (rc = d2h(e, ep, P.eps_p, ld)) || (rc = d2h(e, m, P.m, ld)) || (rc = d2h(e, V0, P.V0, ld)))
            return rc;
        if (P.matid && (rc = d2h(e, mid, (const int *)P.matid, ld))) return rc;
        if (P.trac && (rc = d2h(e, tr, (const double *)P.trac, 3 * ld))) return rc;
        MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
        for (int i = 0; i < n; i++) {
            if (p->mass) p->mass[i] = m[i];
            if (p->vol0) p->vol0[i] = V0[i];
            if (p->matid) p->matid[i] = mid.empty() ? 0 : mid[i];
            if (p->trac_force)
                for (int a = 0; a < 3; a++) p->trac_force[3 * (size_t)i + a] = tr.empty() ? 0.0 : tr[a * ld + i];
            for (int a = 0; a < 3; a++) {
                if (p->coord) p->coord[3 * (size_t)i + a] = x[a * ld + i];
                if (p->vel) p->vel[3 * (size_t)i + a] = v[a * ld + i];
            }
            if (p->stress)
                for (int q = 0; q < 6; q++) p->stress[6 * (size_t)i + q] = sig[q * ld + i];
            if (p->def_grad)
                for (int q = 0; q < 9; q++) p->def_grad[9 * (size_t)i + q] = F[q * ld + i];
            if (p->vol) p->vol[i] = V[i];
            if (p->det_def_grad) p->det_def_grad[i] = J[i];
            if (p->eps_p && !ep.empty()) p->eps_p[i] = ep[i];
        }
    } else {
        const PartF32 &P = e->P32[e->cur];
        if (e->jv_derived) launch_f32_refresh_jv(e->stream, e->g, P);
        std::vector<float> rr, v, sig, H, ep, m, V0, tr;
        std::vector<double> V, J;
        std::vector<int> pidv, mid;
        std::vector<uint32_t> keys;
        if ((rc = d2h(e, rr, P.r, 3 * ld)) || (rc = d2h(e, v, P.v, 3 * ld)) || (rc = d2h(e, keys, P.key, ld)) ||
            (rc = d2h(e, sig, P.sig, 6 * ld)) || (rc = d2h(e, H, P.H, 9 * ld)) || (rc = d2h(e, V, P.V, ld)) ||
            (rc = d2h(e, J, P.J, ld)) || (rc = d2h(e, ep, P.eps_p, ld)) || (rc = d2h(e, pidv, e->pid[e->cur], ld)) ||
            (rc = d2h(e, m, P.m, ld)) || (rc = d2h(e, V0, P.V0, ld)))
            return rc;
        if (P.matid && (rc = d2h(e, mid, (const int *)P.matid, ld))) return rc;
        if (P.trac && (rc = d2h(e, tr, (const float *)P.trac, 3 * ld))) return rc;
        MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
        const GridGeom &g = e->g;
        for (int s = 0; s < n; s++) {
            const size_t i = (size_t)pidv[s];
            const uint32_t k = keys[s] < (uint32_t)g.nelem ? keys[s] : 0u;
            const int ijk[3] = {(int)(k % (uint32_t)g.nex[0]), (int)((k / (uint32_t)g.nex[0]) % (uint32_t)g.nex[1]),
                                (int)(k / ((uint32_t)g.nex[0] * (uint32_t)g.nex[1]))};
            for (int a = 0; a < 3; a++) {
                if (p->coord)
                    p->coord[3 * i + a] = g.xmin[a] + g.dx[a] * ((double)ijk[a] + (double)rr[a * ld + s]);
                if (p->vel) p->vel[3 * i + a] = v[a * ld + s];
            }
            if (p->stress)
                for (int q = 0; q < 6; q++) p->stress[6 * i + q] = sig[q * ld + s];
            if (p->def_grad)
                for (int q = 0; q < 9; q++)
                    p->def_grad[9 * i + q] = (double)H[q * ld + s] + ((q == 0 || q == 4 || q == 8) ? 1.0 : 0.0);
            if (p->vol) p->vol[i] = V[s];
            if (p->det_def_grad) p->det_def_grad[i] = J[s];
            if (p->eps_p && !ep.empty()) p->eps_p[i] = ep[s];
            if (p->mass) p->mass[i] = m[s];
            if (p->vol0) p->vol0[i] = V0[s];
            if (p->matid) p->matid[i] = mid.empty() ? 0 : mid[s];
            if (p->trac_force)
                for (int a = 0; a < 3; a++) p->trac_force[3 * i + a] = tr.empty() ? 0.0 : (double)tr[a * ld + s];
        }
    }
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_download_nodes(mpm_engine *e, mpm_nodes_view *nv) try {
    if (!e || !nv) return set_error(MPM_ERR_INVALID_ARGUMENT, "null argument");
    if (nv->nodec != e->g.nodec) return set_error(MPM_ERR_INVALID_ARGUMENT, "nodec differs from the engine's");
    const size_t nodec = (size_t)e->g.nodec;
    int rc = 0;
    if (e->prec == MPM_PREC_FP64_EXACT) {
        const NodesF64 &N = e->N64;
        struct {
            double *dst;
            const double *src;
            size_t cnt;
        } items[] = {{nv->nmass, N.nm, nodec},  {nv->nvof, N.nvof, nodec},     {nv->nmome, N.np, 3 * nodec},
                     {nv->nforce, N.nf, 3 * nodec}, {nv->nvel, N.nv, 3 * nodec}, {nv->naccel, N.na, 3 * nodec}};
        for (auto &it : items)
            if (it.dst) MPM_HIP_CHECK(hipMemcpyAsync(it.dst, it.src, it.cnt * sizeof(double), hipMemcpyDeviceToHost,
                                                     e->stream));
        MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    } else {
        const NodesF32 &N = e->N32;
        const float *vel_src = (e->last_pass == 2) ? N.nvm : N.nv;
        struct {
            double *dst;
            const float *src;
            size_t cnt;
        } items[] = {{nv->nmass, N.nm, nodec},  {nv->nvof, N.nvof, nodec},     {nv->nmome, N.np, 3 * nodec},
                     {nv->nforce, N.nf, 3 * nodec}, {nv->nvel, vel_src, 3 * nodec}, {nv->naccel, N.na, 3 * nodec}};
        std::vector<float> tmp;
        for (auto &it : items) {
            if (!it.dst) continue;
            if ((rc = d2h(e, tmp, it.src, it.cnt))) return rc;
            MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
            for (size_t i = 0; i < it.cnt; i++) it.dst[i] = tmp[i];
        }
    }
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_bin(mpm_engine *e) try {
    int rc = require_particles(e);
    if (rc) return rc;
    clear_err(e);
    if ((rc = do_bin(e))) return rc;
    return check_err(e);
} catch (...) {
    return mpm::exception_status();
}

int mpm_p2g(mpm_engine *e, double facl) try {
    int rc = require_particles(e);
    if (rc) return rc;
    if (!e->bins_valid) return set_error(MPM_ERR_INVALID_ARGUMENT, "call mpm_bin before mpm_p2g");
    if ((rc = upload_bc(e))) return rc;
    if ((rc = do_p2g(e, facl, 1))) return rc;
    MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_solve_grid(mpm_engine *e) try {
    if (!e) return set_error(MPM_ERR_INVALID_ARGUMENT, "null engine");
    int rc = upload_bc(e);
    if (rc) return rc;
    if ((rc = do_solve(e))) return rc;
    MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_g2p(mpm_engine *e, double facl) try {
    int rc = require_particles(e);
    if (rc) return rc;
    if (!e->bins_valid) return set_error(MPM_ERR_INVALID_ARGUMENT, "call mpm_bin/mpm_p2g before mpm_g2p");
    if ((rc = upload_bc(e))) return rc;
    clear_err(e);
    if ((rc = do_g2p(e, facl, 1, nullptr))) return rc;
    return check_err(e);
} catch (...) {
    return mpm::exception_status();
}

// ---------------------------------------------------------------- implicit path
// Particle accelerations of the implicit transfers (FP64-exact engine),
// allocated as zero on first use after an upload.
static int implicit_require(mpm_engine *e) {
    int rc = require_particles(e);
    if (rc) return rc;
    if (e->prec != MPM_PREC_FP64_EXACT)
        return set_error(MPM_ERR_UNSUPPORTED, "the implicit transfers run in the FP64-exact engine");
    if (e->prm.scheme != MPM_SCHEME_PIC && e->prm.scheme != MPM_SCHEME_FLIP)
        return set_error(MPM_ERR_UNSUPPORTED, "the implicit transfers support PIC and FLIP");
    if (!e->P64.a) {
        if ((rc = e->dalloc(&e->P64.a, 3 * e->ld, true))) return rc;
        MPM_HIP_CHECK(hipMemsetAsync(e->P64.a, 0, 3 * e->ld * sizeof(double), e->stream));
        MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    }
    return 0;
}

static int accel_copy(mpm_engine *e, double *accel, bool to_device) {
    if (!accel) return set_error(MPM_ERR_INVALID_ARGUMENT, "null acceleration array");
    int rc = implicit_require(e);
    if (rc) return rc;
    const size_t ld = e->ld;
    const int n = e->n;
    std::vector<double> soa(3 * ld);
    if (to_device) {
        for (int i = 0; i < n; i++)
            for (int c = 0; c < 3; c++) soa[c * ld + i] = accel[3 * (size_t)i + c];
        if ((rc = h2d(e, e->P64.a, soa))) return rc;
        MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    } else {
        if ((rc = d2h(e, soa, (const double *)e->P64.a, 3 * ld))) return rc;
        MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
        for (int i = 0; i < n; i++)
            for (int c = 0; c < 3; c++) accel[3 * (size_t)i + c] = soa[c * ld + i];
    }
    return 0;
}

int mpm_upload_accel(mpm_engine *e, const double *accel) try {
    if (e) e->drop_graph();
    return accel_copy(e, const_cast<double *>(accel), true);
} catch (...) {
    return mpm::exception_status();
}

int mpm_download_accel(mpm_engine *e, double *accel) try {
    return accel_copy(e, accel, false);
} catch (...) {
    return mpm::exception_status();
}

int mpm_implicit_p2g(mpm_engine *e, double facl) try {
    int rc = implicit_require(e);
    if (rc) return rc;
    if (!e->bins_valid) return set_error(MPM_ERR_INVALID_ARGUMENT, "call mpm_bin before mpm_implicit_p2g");
    if ((rc = upload_bc(e))) return rc;
    clear_err(e);
    launch_f64_implicit_p2g(e->stream, e->g, e->P64, e->N64, e->lists(), e->bc(), e->prm.mtol, facl);
    MPM_HIP_CHECK(hipGetLastError());
    e->last_pass = 1;
    return check_err(e);
} catch (...) {
    return mpm::exception_status();
}

int mpm_implicit_g2p(mpm_engine *e, const double *ndispl, const double *nb_para, double gamma_nb) try {
    int rc = implicit_require(e);
    if (rc) return rc;
    if (!ndispl || !nb_para) return set_error(MPM_ERR_INVALID_ARGUMENT, "null ndispl or Newmark coefficients");
    if (e->shifting || e->fbar)
        return set_error(MPM_ERR_UNSUPPORTED, "the implicit G2P runs with particle shifting and F-bar off");
    if (!e->bins_valid) return set_error(MPM_ERR_INVALID_ARGUMENT, "call mpm_bin before mpm_implicit_g2p");
    const size_t n3 = 3 * (size_t)e->g.nodec;
    if (!e->d_ndispl && (rc = e->dalloc(&e->d_ndispl, n3, false))) return rc;
    MPM_HIP_CHECK(hipMemcpyAsync(e->d_ndispl, ndispl, n3 * sizeof(double), hipMemcpyHostToDevice, e->stream));
    NewmarkF64 nb{};
    for (int q = 0; q < 6; q++) nb.c[q] = nb_para[q];
    e->drop_graph();
    clear_err(e);
    launch_f64_implicit_g2p(e->stream, e->g, e->P64, e->N64, e->mt, e->d_ndispl, nb, e->prm.dt, gamma_nb,
                            e->prm.scheme == MPM_SCHEME_FLIP ? 1 : 0, e->d_err);
    MPM_HIP_CHECK(hipGetLastError());
    e->keys_valid = false;  // the positions moved
    e->bins_valid = false;
    e->last_pass = 2;
    return check_err(e);
} catch (...) {
    return mpm::exception_status();
}

int mpm_download_bins(mpm_engine *e, int32_t *idepf, int32_t *idp2p, int32_t *numep, int32_t *cell_of) try {
    int rc = require_particles(e);
    if (rc) return rc;
    if (!e->bins_valid) return set_error(MPM_ERR_INVALID_ARGUMENT, "no valid binning (call mpm_bin)");
    const int n = e->n, nelem = e->g.nelem;
    std::vector<uint32_t> perm, cs, ce, keys;
    std::vector<int> pidv;
    if ((rc = d2h(e, perm, e->perm, n)) || (rc = d2h(e, cs, e->bb.cstart, nelem)) ||
        (rc = d2h(e, ce, e->bb.cend, nelem)))
        return rc;
    const uint32_t *kd = (e->prec == MPM_PREC_FP64_EXACT) ? e->keys64 : e->P32[e->cur].key;
    if ((rc = d2h(e, keys, kd, n))) return rc;
    if (e->prec == MPM_PREC_FP32) {
        if ((rc = d2h(e, pidv, e->pid[e->cur], n))) return rc;
    }
    MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    auto to_pid = [&](uint32_t slot) -> int { return pidv.empty() ? (int)slot : pidv[slot]; };
    if (idp2p)
        for (int i = 0; i < n; i++) idp2p[i] = -1;
    // The reference's lists run through a cell's particles in index order.
    // The binning orders a cell by slot, which is index order until the FP32
    // engine relays the slots out; then each cell is put back in index order.
    std::vector<int> pids;
    for (int c = 0; c < nelem; c++) {
        uint32_t a = cs[c], b = ce[c];
        if (numep) numep[c] = (int32_t)(b - a);
        pids.clear();
        for (uint32_t s = a; s < b; s++) pids.push_back(to_pid(perm[s]));
        if (!pidv.empty()) std::sort(pids.begin(), pids.end());
        if (idepf) idepf[c] = pids.empty() ? -1 : pids[0];
        if (idp2p)
            for (size_t q = 0; q + 1 < pids.size(); q++) idp2p[pids[q]] = pids[q + 1];
    }
    if (cell_of)
        for (int s = 0; s < n; s++) cell_of[to_pid((uint32_t)s)] = keys[s] >= (uint32_t)nelem ? -1 : (int32_t)keys[s];
    return 0;
} catch (...) {
    return mpm::exception_status();
}

// Before recording a graph: make the per-cell counts of the current keys
// ready (as after an update kernel), so that every recorded block starts and
// ends in the same state.
static void prime_counts(mpm_engine *e) {
    if (e->prec == MPM_PREC_FP32 && e->fused_count && !e->counts_ready) {
        launch_bin_count(e->stream, e->g, e->tg, e->P32[e->cur].key, e->n, e->bb);
        e->counts_ready = true;
    }
}

// Enqueues nsteps whole steps on the engine's stream: kernels only (no
// host synchronisation, no memset), so a stream capture can record it.
// ev: 7 events per step for per-kernel timing, or null.
static int enqueue_steps(mpm_engine *e, int32_t istep0, int32_t nsteps, bool diag_last, hipEvent_t *evs) {
    int rc;
    const int nev = 7;
    for (int s = 0; s < nsteps; s++) {
        const int istep = istep0 + s;
        const double facl = facl_of(e, istep);
        const int diag = (diag_last && s == nsteps - 1) ? 1 : 0;
        hipEvent_t *ev = evs ? &evs[(size_t)nev * s] : nullptr;
        // events: 0 | bin | 1 | p2g cells | 2 | p2g nodes | 3 | musl cells | 4 | musl nodes | 5 | update | 6
        if (ev) MPM_HIP_CHECK(hipEventRecord(ev[0], e->stream));
        if ((rc = do_bin(e))) return rc;
        if (ev) MPM_HIP_CHECK(hipEventRecord(ev[1], e->stream));
        if ((rc = do_p2g(e, facl, diag, ev ? ev[2] : nullptr))) return rc;
        if ((rc = do_solve(e))) return rc;
        if (ev) MPM_HIP_CHECK(hipEventRecord(ev[3], e->stream));
        if ((rc = do_g2p(e, facl, diag, ev ? ev[5] : nullptr, ev ? ev[4] : nullptr))) return rc;
        if (ev) MPM_HIP_CHECK(hipEventRecord(ev[6], e->stream));
    }
    return 0;
}

int mpm_step(mpm_engine *e, int32_t istep0, int32_t nsteps) try {
    int rc = require_particles(e);
    if (rc) return rc;
    if (nsteps < 0) return set_error(MPM_ERR_INVALID_ARGUMENT, "nsteps < 0");
    if ((rc = upload_bc(e))) return rc;
    clear_err(e);
    const int nev = 7;
    if (e->timing) {
        size_t need = (size_t)nev * nsteps;
        while (e->ev.size() < need) {
            hipEvent_t ev;
            MPM_HIP_CHECK(hipEventCreate(&ev));
            e->ev.push_back(ev);
        }
    }
    const int B = e->graph_block;
    const bool ramp_over = e->prm.nlstep <= 1 || istep0 > e->prm.nlstep;
    const bool relayout_ok = !(e->prec == MPM_PREC_FP32 && e->relayout_every > 0) || B % (2 * e->relayout_every) == 0;
    if (B > 0 && !e->timing && ramp_over && relayout_ok && !e->sur.logging && nsteps > B) {
        // all but the last step in whole graph blocks; the rest eagerly (the
        // last step may write the diagnostic node fields)
        const int nblk = (nsteps - 1) / B;
        if (e->step_exec && (e->graph_cur != e->cur || e->graph_since != e->steps_since_relayout)) e->drop_graph();
        if (e->step_exec) prime_counts(e);
        if (!e->step_exec) {
            if (e->prec == MPM_PREC_FP64_EXACT) e->keys_valid = false;
            prime_counts(e);
            const int cur0 = e->cur, since0 = e->steps_since_relayout;
            MPM_HIP_CHECK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
            rc = enqueue_steps(e, istep0, B, false, nullptr);
            hipGraph_t g = nullptr;
            hipError_t he = hipStreamEndCapture(e->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (he != hipSuccess) return hip_fail(he, "hipStreamEndCapture", __FILE__, __LINE__);
            e->step_graph = g;
            if (e->cur != cur0 || e->steps_since_relayout != since0 ||
                (e->prec == MPM_PREC_FP32 && e->fused_count && !e->counts_ready)) {
                e->drop_graph();
                return set_error(MPM_ERR_INVALID_ARGUMENT, "graph block does not return to the same particle buffer");
            }
            MPM_HIP_CHECK(hipGraphInstantiate(&e->step_exec, e->step_graph, nullptr, nullptr, 0));
            e->graph_cur = cur0;
            e->graph_since = since0;
        }
        if (nblk > 0 && e->prec == MPM_PREC_FP32) e->jv_derived = true;
        for (int b = 0; b < nblk; b++) MPM_HIP_CHECK(hipGraphLaunch(e->step_exec, e->stream));
        if (e->prec == MPM_PREC_FP64_EXACT) e->keys_valid = false;
        const int done = nblk * B;
        if ((rc = enqueue_steps(e, istep0 + done, nsteps - done, e->full_node_output != 0, nullptr))) return rc;
    } else if ((rc = enqueue_steps(e, istep0, nsteps, e->full_node_output != 0, e->timing ? e->ev.data() : nullptr))) {
        return rc;
    }
    MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    if (e->timing) {
        mpm_kernel_times kt{};
        const bool f32 = e->prec == MPM_PREC_FP32;
        for (int s = 0; s < nsteps; s++) {
            hipEvent_t *ev = &e->ev[(size_t)nev * s];
            float t[6] = {0, 0, 0, 0, 0, 0};
            if (f32) {
                for (int k = 0; k < 6; k++) MPM_HIP_CHECK(hipEventElapsedTime(&t[k], ev[k], ev[k + 1]));
            } else {
                MPM_HIP_CHECK(hipEventElapsedTime(&t[0], ev[0], ev[1]));
                MPM_HIP_CHECK(hipEventElapsedTime(&t[1], ev[1], ev[3]));
                MPM_HIP_CHECK(hipEventElapsedTime(&t[5], ev[3], ev[6]));
            }
            kt.bin_ms += t[0];
            kt.p2g_cells_ms += t[1];
            kt.p2g_nodes_ms += t[2];
            kt.musl_cells_ms += t[3];
            kt.musl_nodes_ms += t[4];
            kt.g2p_upd_ms += t[5];
        }
        kt.p2g_ms = kt.p2g_cells_ms + kt.p2g_nodes_ms;
        kt.g2p_vel_ms = kt.musl_cells_ms + kt.musl_nodes_ms;
        kt.total_ms = kt.bin_ms + kt.p2g_ms + kt.g2p_vel_ms + kt.g2p_upd_ms;
        kt.steps = nsteps;
        e->kt = kt;
    }
    return check_err(e);
} catch (...) {
    return mpm::exception_status();
}

int mpm_sync(mpm_engine *e) try {
    if (!e) return set_error(MPM_ERR_INVALID_ARGUMENT, "null engine");
    MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_get_kernel_times(mpm_engine *e, mpm_kernel_times *out) try {
    if (!e || !out) return set_error(MPM_ERR_INVALID_ARGUMENT, "null argument");
    *out = e->kt;
    if (e->bins_valid || e->n > 0) {
        // count of non-empty cells of the last binning
        std::vector<uint32_t> ce, cs;
        int rc = d2h(e, ce, e->bb.cend, (size_t)e->g.nelem);
        if (!rc) rc = d2h(e, cs, e->bb.cstart, (size_t)e->g.nelem);
        if (rc) return rc;
        MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
        const GridGeom &g = e->g;
        std::vector<uint8_t> act((size_t)g.nodec, 0);
        int cnt = 0;
        for (int c = 0; c < g.nelem; c++) {
            if (ce[c] <= cs[c]) continue;
            cnt++;
            int iz = c / (g.nex[0] * g.nex[1]);
            int iy = (c / g.nex[0]) % g.nex[1];
            int ix = c % g.nex[0];
            for (int k = 0; k <= g.idimc[2]; k++)
                for (int j = 0; j <= g.idimc[1]; j++)
                    for (int i = 0; i <= g.idimc[0]; i++)
                        act[(size_t)(ix + i) + (size_t)g.nxc[0] * (iy + j) + (size_t)g.nxc[0] * g.nxc[1] * (iz + k)] = 1;
        }
        int na = 0;
        for (uint8_t a : act) na += a;
        out->active_cells = cnt;
        out->active_nodes = na;
    }
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_set_timing(mpm_engine *e, int32_t enabled) try {
    if (!e) return set_error(MPM_ERR_INVALID_ARGUMENT, "null engine");
    e->timing = enabled != 0;
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_set_option(mpm_engine *e, int32_t key, int64_t value) try {
    if (!e) return set_error(MPM_ERR_INVALID_ARGUMENT, "null engine");
    e->drop_graph();
    switch (key) {
    case OPT_FULL_NODE_OUTPUT: e->full_node_output = value ? 1 : 0; return 0;
    case OPT_LANES_PER_CELL:
        if (value != 1 && value != 2 && value != 4 && value != 8)
            return set_error(MPM_ERR_INVALID_ARGUMENT, "lanes per cell must be 1, 2, 4 or 8");
        e->lanes_per_cell = (int)value;
        e->bins_valid = false;  // the group-start bits of the binning follow the cells kernels' groups
        return 0;
    case OPT_RELAYOUT_EVERY:
        if (value < 0 || value > (1 << 20)) return set_error(MPM_ERR_INVALID_ARGUMENT, "relayout interval out of range");
        if (e->prec != MPM_PREC_FP32) return value ? set_error(MPM_ERR_UNSUPPORTED, "relayout is an FP32 option") : 0;
        if (value > 0) {
            int rc = ensure_relayout_buffer(e);
            if (rc) return rc;
        }
        e->relayout_every = (int)value;
        e->steps_since_relayout = 0;
        return 0;
    case OPT_X_MERGE:
        e->x_merge = value != 0;
        e->bins_valid = false;  // the group-start bits are written by the binning
        return 0;
    case OPT_SURROGATE_FUSED:
        e->sur.fused = value != 0;
        return 0;
    case OPT_CELLS_PREFETCH:
        e->drop_graph();
        e->cells_prefetch = value ? 1 : 0;
        return 0;
    case OPT_CELLS_GRID:
        if (value < 0 || value > (1 << 20)) return set_error(MPM_ERR_INVALID_ARGUMENT, "cells grid out of range");
        e->drop_graph();
        e->cells_grid = (int)value;
        return 0;
    case OPT_FUSED_COUNT:
        e->drop_graph();  // recorded steps assume the counting mode they were recorded with
        e->clear_counts();
        e->fused_count = value != 0;
        return 0;
    case OPT_GRAPH_BLOCK:
        if (value < 0 || value > 4096) return set_error(MPM_ERR_INVALID_ARGUMENT, "graph block out of range");
        e->graph_block = (int)value;
        return 0;
    default: return set_error(MPM_ERR_INVALID_ARGUMENT, "unknown option");
    }
} catch (...) {
    return mpm::exception_status();
}

int mpm_set_material_dt(mpm_engine *e, const double *dt) try {
    if (!e) return set_error(MPM_ERR_INVALID_ARGUMENT, "null engine");
    if (dt && (e->prec != MPM_PREC_FP32 || e->stencil))
        return set_error(MPM_ERR_UNSUPPORTED,
                         "per-material time steps run in the FP32 engine with the tent functions and PIC / FLIP");
    std::vector<MaterialDev> mh = e->mat_host;
    for (size_t i = 0; i < mh.size(); i++) {
        if (dt && !(dt[i] > 0.0 && dt[i] < 1e30)) return set_error(MPM_ERR_INVALID_ARGUMENT, "time steps must be positive");
        mh[i].dt = dt ? (float)dt[i] : 0.f;
    }
    e->drop_graph();
    MPM_HIP_CHECK(hipMemcpyAsync((void *)e->mt.m, mh.data(), sizeof(MaterialDev) * mh.size(), hipMemcpyHostToDevice,
                                 e->stream));
    MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    e->mat_host = mh;
    e->mat_dt = dt ? 1 : 0;
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_maksf_batch(const mpm_grid_desc *grid, int64_t n, const int32_t *cell, const double *xyg, int32_t *nc,
                    double *sf, double *dsf) try {
    if (!grid || n < 0 || (n > 0 && (!cell || !xyg || !nc || !sf || !dsf)))
        return set_error(MPM_ERR_INVALID_ARGUMENT, "null argument");
    int rc = validate_grid(grid, true);
    if (rc) return rc;
    if (n == 0) return 0;
    if (n >= (1LL << 31) / 27) return set_error(MPM_ERR_INVALID_ARGUMENT, "batch too large");
    GridGeom g = make_geom(grid);
    for (int64_t i = 0; i < n; i++)
        if (cell[i] < 0 || cell[i] >= g.nelem) return set_error(MPM_ERR_INVALID_ARGUMENT, "cell out of range");
    const int nen = (g.idimc[0] + 1) * (g.idimc[1] + 1) * (g.idimc[2] + 1);
    int32_t *dc = nullptr, *dnc = nullptr;
    double *dx = nullptr, *dsfd = nullptr, *ddsf = nullptr;
    hipStream_t s = nullptr;
    auto cleanup = [&]() {
        (void)hipFree(dc);
        (void)hipFree(dnc);
        (void)hipFree(dx);
        (void)hipFree(dsfd);
        (void)hipFree(ddsf);
        if (s) (void)hipStreamDestroy(s);
    };
    hipError_t he = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (he == hipSuccess) he = hipMalloc((void **)&dc, n * sizeof(int32_t));
    if (he == hipSuccess) he = hipMalloc((void **)&dnc, n * nen * sizeof(int32_t));
    if (he == hipSuccess) he = hipMalloc((void **)&dx, 3 * n * sizeof(double));
    if (he == hipSuccess) he = hipMalloc((void **)&dsfd, n * nen * sizeof(double));
    if (he == hipSuccess) he = hipMalloc((void **)&ddsf, 3 * n * nen * sizeof(double));
    if (he == hipSuccess) he = hipMemcpyAsync(dc, cell, n * sizeof(int32_t), hipMemcpyHostToDevice, s);
    if (he == hipSuccess) he = hipMemcpyAsync(dx, xyg, 3 * n * sizeof(double), hipMemcpyHostToDevice, s);
    if (he == hipSuccess) he = (hipError_t)launch_f64_maksf(s, g, (int)n, dc, dx, dnc, dsfd, ddsf);
    if (he == hipSuccess) he = hipMemcpyAsync(nc, dnc, n * nen * sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (he == hipSuccess) he = hipMemcpyAsync(sf, dsfd, n * nen * sizeof(double), hipMemcpyDeviceToHost, s);
    if (he == hipSuccess) he = hipMemcpyAsync(dsf, ddsf, 3 * n * nen * sizeof(double), hipMemcpyDeviceToHost, s);
    if (he == hipSuccess) he = hipStreamSynchronize(s);
    if (he != hipSuccess) rc = hip_fail(he, "mpm_maksf_batch", __FILE__, __LINE__);
    cleanup();
    return rc;
} catch (...) {
    return mpm::exception_status();
}

int mpm_bin_positions(const mpm_grid_desc *grid, int64_t n, const double *xyg, int32_t *cell_of, int32_t *order) try {
    if (!grid || n < 0 || (n > 0 && (!xyg || !cell_of || !order)))
        return set_error(MPM_ERR_INVALID_ARGUMENT, "null argument");
    int rc = validate_grid(grid, false);
    if (rc) return rc;
    if (n == 0) return 0;
    if (n >= (1LL << 31)) return set_error(MPM_ERR_INVALID_ARGUMENT, "too many positions");
    GridGeom g = make_geom(grid);
    const size_t ld = (size_t)n;
    std::vector<double> x(3 * ld);
    for (size_t i = 0; i < ld; i++)
        for (int a = 0; a < 3; a++) x[a * ld + i] = xyg[3 * i + a];
    double *dx = nullptr;
    uint32_t *keys = nullptr, *err = nullptr;
    SortBuffers sb{};
    hipStream_t s = nullptr;
    hipError_t he = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (he == hipSuccess) he = hipMalloc((void **)&dx, 3 * ld * sizeof(double));
    if (he == hipSuccess) he = hipMalloc((void **)&keys, ld * sizeof(uint32_t));
    if (he == hipSuccess) he = hipMalloc((void **)&err, sizeof(uint32_t));
    if (he == hipSuccess) he = hipMalloc((void **)&sb.keys, ld * sizeof(uint32_t));
    if (he == hipSuccess) he = hipMalloc((void **)&sb.vals, ld * sizeof(uint32_t));
    if (he == hipSuccess) he = hipMalloc((void **)&sb.keys_alt, ld * sizeof(uint32_t));
    if (he == hipSuccess) he = hipMalloc((void **)&sb.vals_alt, ld * sizeof(uint32_t));
    if (he == hipSuccess) he = hipMalloc((void **)&sb.hist, radix_hist_entries((int)n) * sizeof(uint32_t));
    if (he == hipSuccess) he = hipMalloc((void **)&sb.bsum, 4096 * sizeof(uint32_t));
    if (he == hipSuccess) he = hipMemcpyAsync(dx, x.data(), 3 * ld * sizeof(double), hipMemcpyHostToDevice, s);
    uint32_t *sk = nullptr, *sv = nullptr;
    std::vector<uint32_t> kh(ld), vh(ld);
    if (he == hipSuccess) {
        launch_fill_u32(s, err, 1, 0u);
        launch_keys_f64(s, g, dx, ld, (int)n, keys, err);
        if (radix_sort_pairs(s, sb, keys, (int)n, bit_length((uint32_t)g.nelem), &sk, &sv)) he = hipErrorUnknown;
    }
    if (he == hipSuccess) he = hipMemcpyAsync(kh.data(), keys, ld * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
    if (he == hipSuccess) he = hipMemcpyAsync(vh.data(), sv, ld * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
    if (he == hipSuccess) he = hipStreamSynchronize(s);
    if (he != hipSuccess) rc = hip_fail(he, "mpm_bin_positions", __FILE__, __LINE__);
    (void)hipFree(dx);
    (void)hipFree(keys);
    (void)hipFree(err);
    (void)hipFree(sb.keys);
    (void)hipFree(sb.vals);
    (void)hipFree(sb.keys_alt);
    (void)hipFree(sb.vals_alt);
    (void)hipFree(sb.hist);
    (void)hipFree(sb.bsum);
    if (s) (void)hipStreamDestroy(s);
    if (rc) return rc;
    for (size_t i = 0; i < ld; i++) {
        cell_of[i] = kh[i] >= (uint32_t)g.nelem ? -1 : (int32_t)kh[i];
        order[i] = (int32_t)vh[i];
    }
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_bin_positions_noalloc(const mpm_grid_desc *grid, int64_t n, const double *xyg, int32_t *cell_of,
                              int32_t *order, void *stream) try {
    if (!grid || n < 0 || (n > 0 && (!xyg || !cell_of || !order)))
        return set_error(MPM_ERR_INVALID_ARGUMENT, "null argument");
    int rc = validate_grid(grid, false);
    if (rc) return rc;
    if (n > MPM_NOALLOC_MAX_POSITIONS)
        return set_error(MPM_ERR_INVALID_ARGUMENT, "more positions than the allocation-free binning holds");
    if (n == 0) return 0;
    const GridGeom g = make_geom(grid);
    const int e = noalloc_bin_positions(g, (int)n, xyg, cell_of, order, (hipStream_t)stream);
    if (e) return hip_fail((hipError_t)e, "mpm_bin_positions_noalloc", __FILE__, __LINE__);
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_stream_mix_probe(int64_t n, int32_t iters, double *gbps) try {
    if (n < 256 || n > (int64_t)1 << 28 || iters < 1 || !gbps)
        return set_error(MPM_ERR_INVALID_ARGUMENT, "bad probe arguments");
    // four buffer sets used in turn: more than the last-level cache holds, so
    // every launch streams from HBM
    constexpr int kSets = 4;
    float *p[kSets] = {nullptr, nullptr, nullptr, nullptr};
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    const size_t nf = stream_mix_floats((size_t)n);
    hipError_t he = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    for (int k = 0; k < kSets && he == hipSuccess; k++) {
        he = hipMalloc((void **)&p[k], nf * sizeof(float));
        if (he == hipSuccess) he = hipMemsetAsync(p[k], 0, nf * sizeof(float), s);
    }
    if (he == hipSuccess) he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    if (he == hipSuccess) {
        for (int k = 0; k < kSets; k++) launch_stream_mix(s, p[k], (size_t)n);  // warm-up
        he = hipEventRecord(e0, s);
    }
    if (he == hipSuccess) {
        for (int i = 0; i < iters; i++) launch_stream_mix(s, p[i % kSets], (size_t)n);
        he = hipEventRecord(e1, s);
    }
    if (he == hipSuccess) he = hipEventSynchronize(e1);
    float ms = 0.f;
    if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
    if (he != hipSuccess) rc = hip_fail(he, "mpm_stream_mix_probe", __FILE__, __LINE__);
    else *gbps = stream_mix_bytes((size_t)n) * iters / (ms * 1e-3) / 1e9;
    for (float *q : p)
        if (q) (void)hipFree(q);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    return rc;
} catch (...) {
    return mpm::exception_status();
}

int mpm_bandwidth_probe(int64_t nbytes, int32_t iters, double *gbps) try {
    if (nbytes < 4096 || iters < 1 || !gbps) return set_error(MPM_ERR_INVALID_ARGUMENT, "bad probe arguments");
    const size_t n16 = (size_t)nbytes / 16;
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    hipError_t he = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (he == hipSuccess) he = hipMalloc(&a, n16 * 16);
    if (he == hipSuccess) he = hipMalloc(&b, n16 * 16);
    if (he == hipSuccess) he = hipMemsetAsync(a, 1, n16 * 16, s);
    if (he == hipSuccess) he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    if (he == hipSuccess) {
        launch_copy16(s, a, b, n16);  // warm-up
        he = hipEventRecord(e0, s);
    }
    if (he == hipSuccess) {
        for (int i = 0; i < iters; i++) launch_copy16(s, (i & 1) ? b : a, (i & 1) ? a : b, n16);
        he = hipEventRecord(e1, s);
    }
    if (he == hipSuccess) he = hipEventSynchronize(e1);
    float ms = 0.f;
    if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
    if (he != hipSuccess) rc = hip_fail(he, "mpm_bandwidth_probe", __FILE__, __LINE__);
    else *gbps = 2.0 * (double)n16 * 16.0 * iters / (ms * 1e-3) / 1e9;
    (void)hipFree(a);
    (void)hipFree(b);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    return rc;
} catch (...) {
    return mpm::exception_status();
}

}  // extern "C"

// ------------------------------------------------------------------ ensembles
struct mpm_ensemble {
    std::vector<mpm_engine *> eng;
    std::vector<int32_t> bucket;
    hipStream_t s0 = nullptr;
    hipEvent_t fork = nullptr, t0 = nullptr, t1 = nullptr;
    std::vector<hipEvent_t> join;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int block_steps = 0;

    ~mpm_ensemble() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto ev : join) (void)hipEventDestroy(ev);
        for (auto ev : {fork, t0, t1})
            if (ev) (void)hipEventDestroy(ev);
        if (s0) (void)hipStreamDestroy(s0);
    }
};

namespace {

// fork: every scene's stream waits for s0; join: s0 waits for every scene
int ens_fork(mpm_ensemble *x, const std::vector<int> &idx) {
    MPM_HIP_CHECK(hipEventRecord(x->fork, x->s0));
    for (int i : idx) MPM_HIP_CHECK(hipStreamWaitEvent(x->eng[i]->stream, x->fork, 0));
    return 0;
}

int ens_join(mpm_ensemble *x, const std::vector<int> &idx) {
    for (int i : idx) {
        MPM_HIP_CHECK(hipEventRecord(x->join[i], x->eng[i]->stream));
        MPM_HIP_CHECK(hipStreamWaitEvent(x->s0, x->join[i], 0));
    }
    return 0;
}

int ens_prepare(mpm_ensemble *x) {
    int rc;
    for (auto *e : x->eng) {
        if ((rc = require_particles(e)) || (rc = upload_bc(e))) return rc;
        clear_err(e);
        prime_counts(e);  // recorded blocks start from counted keys
    }
    for (auto *e : x->eng) MPM_HIP_CHECK(hipStreamSynchronize(e->stream));
    return 0;
}

int ens_check(mpm_ensemble *x) {
    MPM_HIP_CHECK(hipStreamSynchronize(x->s0));
    for (size_t i = 0; i < x->eng.size(); i++) {
        int rc = check_err(x->eng[i]);
        if (rc) {
            std::string m = "scene " + std::to_string(i) + ": " + g_last_error;
            return set_error(rc, m);
        }
    }
    return 0;
}

int ens_eager(mpm_ensemble *x, const std::vector<int> &idx, int32_t istep0, int32_t nsteps, double *ms) {
    int rc;
    MPM_HIP_CHECK(hipEventRecord(x->t0, x->s0));
    if ((rc = ens_fork(x, idx))) return rc;
    for (int i : idx)
        if ((rc = enqueue_steps(x->eng[i], istep0, nsteps, false, nullptr))) return rc;
    if ((rc = ens_join(x, idx))) return rc;
    MPM_HIP_CHECK(hipEventRecord(x->t1, x->s0));
    MPM_HIP_CHECK(hipEventSynchronize(x->t1));
    float t = 0.f;
    MPM_HIP_CHECK(hipEventElapsedTime(&t, x->t0, x->t1));
    if (ms) *ms = t;
    return 0;
}

}  // namespace

extern "C" {

int mpm_ensemble_create(mpm_engine *const *engines, int32_t n, const int32_t *bucket, mpm_ensemble **out) try {
    if (!engines || n <= 0 || !out) return set_error(MPM_ERR_INVALID_ARGUMENT, "bad ensemble arguments");
    *out = nullptr;
    auto *x = new mpm_ensemble();
    for (int i = 0; i < n; i++) {
        if (!engines[i]) {
            delete x;
            return set_error(MPM_ERR_INVALID_ARGUMENT, "null engine in ensemble");
        }
        x->eng.push_back(engines[i]);
        x->bucket.push_back(bucket ? bucket[i] : 0);
    }
    hipError_t he = hipSetDevice(x->eng[0]->device);
    if (he == hipSuccess) he = hipStreamCreateWithFlags(&x->s0, hipStreamNonBlocking);
    if (he == hipSuccess) he = hipEventCreateWithFlags(&x->fork, hipEventDisableTiming);
    if (he == hipSuccess) he = hipEventCreate(&x->t0);
    if (he == hipSuccess) he = hipEventCreate(&x->t1);
    x->join.assign(n, nullptr);
    for (int i = 0; i < n && he == hipSuccess; i++) he = hipEventCreateWithFlags(&x->join[i], hipEventDisableTiming);
    if (he != hipSuccess) {
        int r = hip_fail(he, "ensemble create", __FILE__, __LINE__);
        delete x;
        return r;
    }
    *out = x;
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_ensemble_destroy(mpm_ensemble *x) try {
    if (!x) return 0;
    if (x->s0) (void)hipStreamSynchronize(x->s0);
    delete x;
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_ensemble_step(mpm_ensemble *x, int32_t istep0, int32_t nsteps, double *elapsed_ms) try {
    if (!x || nsteps < 0) return set_error(MPM_ERR_INVALID_ARGUMENT, "bad arguments");
    int rc = ens_prepare(x);
    if (rc) return rc;
    std::vector<int> all(x->eng.size());
    for (size_t i = 0; i < all.size(); i++) all[i] = (int)i;
    if ((rc = ens_eager(x, all, istep0, nsteps, elapsed_ms))) return rc;
    return ens_check(x);
} catch (...) {
    return mpm::exception_status();
}

int mpm_ensemble_capture(mpm_ensemble *x, int32_t istep0, int32_t block_steps) try {
    if (!x || block_steps <= 0) return set_error(MPM_ERR_INVALID_ARGUMENT, "bad arguments");
    for (auto *e : x->eng) {
        if (istep0 <= e->prm.nlstep && e->prm.nlstep > 1)
            return set_error(MPM_ERR_INVALID_ARGUMENT, "a captured block must start after the load ramp (nlstep)");
        if (e->sur.logging) return set_error(MPM_ERR_INVALID_ARGUMENT, "sample logging cannot be captured");
        if (e->prec == MPM_PREC_FP32 && e->relayout_every > 0 && block_steps % (2 * e->relayout_every) != 0)
            return set_error(MPM_ERR_INVALID_ARGUMENT,
                             "block_steps must be a multiple of twice every scene's relayout interval");
    }
    int rc = ens_prepare(x);
    if (rc) return rc;
    if (x->exec) {
        (void)hipGraphExecDestroy(x->exec);
        x->exec = nullptr;
    }
    if (x->graph) {
        (void)hipGraphDestroy(x->graph);
        x->graph = nullptr;
    }
    std::vector<int> all(x->eng.size());
    for (size_t i = 0; i < all.size(); i++) all[i] = (int)i;
    // host-side step state that a replay must find unchanged
    std::vector<int> cur0, since0;
    for (auto *e : x->eng) {
        prime_counts(e);
        if (e->prec == MPM_PREC_FP64_EXACT) e->keys_valid = false;  // every replay recomputes the FP64 cell keys
        cur0.push_back(e->cur);
        since0.push_back(e->steps_since_relayout);
    }
    MPM_HIP_CHECK(hipStreamBeginCapture(x->s0, hipStreamCaptureModeThreadLocal));
    rc = ens_fork(x, all);
    for (int i : all)
        if (!rc) rc = enqueue_steps(x->eng[i], istep0, block_steps, false, nullptr);
    if (!rc) rc = ens_join(x, all);
    hipGraph_t g = nullptr;
    hipError_t he = hipStreamEndCapture(x->s0, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (he != hipSuccess) return hip_fail(he, "hipStreamEndCapture", __FILE__, __LINE__);
    x->graph = g;
    for (size_t i = 0; i < x->eng.size(); i++)
        if (x->eng[i]->cur != cur0[i] || x->eng[i]->steps_since_relayout != since0[i])
            return set_error(MPM_ERR_INVALID_ARGUMENT, "captured block does not return to the same particle buffer");
    MPM_HIP_CHECK(hipGraphInstantiate(&x->exec, x->graph, nullptr, nullptr, 0));
    x->block_steps = block_steps;
    // the capture did not run anything: the state is where it was
    return 0;
} catch (...) {
    return mpm::exception_status();
}

int mpm_ensemble_launch(mpm_ensemble *x, int32_t nlaunch, double *elapsed_ms) try {
    if (!x || nlaunch < 0) return set_error(MPM_ERR_INVALID_ARGUMENT, "bad arguments");
    if (!x->exec) return set_error(MPM_ERR_INVALID_ARGUMENT, "no captured graph");
    int rc = ens_prepare(x);
    if (rc) return rc;
    MPM_HIP_CHECK(hipEventRecord(x->t0, x->s0));
    for (auto *e : x->eng)
        if (nlaunch > 0 && e->prec == MPM_PREC_FP32) e->jv_derived = true;
    for (int k = 0; k < nlaunch; k++) MPM_HIP_CHECK(hipGraphLaunch(x->exec, x->s0));
    MPM_HIP_CHECK(hipEventRecord(x->t1, x->s0));
    MPM_HIP_CHECK(hipEventSynchronize(x->t1));
    float t = 0.f;
    MPM_HIP_CHECK(hipEventElapsedTime(&t, x->t0, x->t1));
    if (elapsed_ms) *elapsed_ms = t;
    for (auto *e : x->eng) {
        e->bins_valid = false;
        e->keys_valid = true;
    }
    return ens_check(x);
} catch (...) {
    return mpm::exception_status();
}

int mpm_ensemble_autotune(mpm_ensemble *x, int32_t istep0, int32_t trial_steps, mpm_tune_result *results,
                          int32_t max_results, int32_t *nresults) try {
    if (!x || trial_steps <= 0 || (max_results > 0 && !results) || !nresults)
        return set_error(MPM_ERR_INVALID_ARGUMENT, "bad arguments");
    int rc = ens_prepare(x);
    if (rc) return rc;
    std::vector<int32_t> ids = x->bucket;
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    const int lpcs[3] = {1, 2, 4}, rels[3] = {0, 8, 16};
    int nres = 0;
    for (int32_t b : ids) {
        std::vector<int> idx;
        bool f32 = true;
        for (size_t i = 0; i < x->eng.size(); i++)
            if (x->bucket[i] == b) {
                idx.push_back((int)i);
                f32 = f32 && x->eng[i]->prec == MPM_PREC_FP32;
            }
        if (f32)
            for (int i : idx)
                if ((rc = ensure_relayout_buffer(x->eng[i]))) return rc;
        double best = 1e300, worst = 0.0;
        int best_l = 1, best_r = 0, step = istep0;
        for (int li = 0; li < 3; li++)
            for (int ri = 0; ri < 3; ri++) {
                for (int i : idx) {
                    auto *e = x->eng[i];
                    if (f32) {
                        e->lanes_per_cell = lpcs[li];
                        e->relayout_every = rels[ri];
                        e->steps_since_relayout = 0;
                    }
                }
                double ms = 0.0;
                if ((rc = ens_eager(x, idx, step, trial_steps, &ms))) return rc;
                step += trial_steps;
                const double per = ms / trial_steps;
                if (per < best) {
                    best = per;
                    best_l = lpcs[li];
                    best_r = rels[ri];
                }
                worst = std::max(worst, per);
            }
        for (int i : idx) {
            auto *e = x->eng[i];
            if (f32) {
                e->lanes_per_cell = best_l;
                e->relayout_every = best_r;
                e->steps_since_relayout = 0;
            }
        }
        if (nres < max_results) {
            results[nres].bucket = b;
            results[nres].lanes_per_cell = f32 ? best_l : 0;
            results[nres].relayout_every = f32 ? best_r : 0;
            results[nres].scenes = (int32_t)idx.size();
            results[nres].ms_per_step = best;
            results[nres].ms_per_step_worst = worst;
        }
        nres++;
    }
    *nresults = std::min(nres, max_results);
    return ens_check(x);
} catch (...) {
    return mpm::exception_status();
}

}  // extern "C"


// C ABI and host runtime of the XPBD cloth compute path (include/xpbd.h).
//
// This is the MI355X counterpart of the reference's render-world systems:
// init_cloth_sim_jacobi (buffer creation/upload, cloth_jacobi.rs:213-569),
// run_cloth_sim_jacobi / run_cloth_sim_gauss_seidel (dispatch sequence,
// cloth_jacobi.rs:675-916, cloth_compute.rs:1272-1522) and the reset path
// (cloth_jacobi.rs:918-944).  Device state lives in HBM for the handle's
// lifetime; a frame is enqueued as kernels on one stream and, when enabled,
// replayed from a captured hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/xpbd.h"
#include "kernels.hpp"
#include "mesh.hpp"
#include "mlp.hpp"

using namespace xpbd;

struct xpbd_mesh {
    Mesh m;
};

// The warm-start network: f32 master parameters, Adam state, bf16 operand
// copies and training scratch, all on one device and stream.
struct xpbd_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    float *P = nullptr, *m = nullptr, *v = nullptr, *partial = nullptr, *grad = nullptr, *loss = nullptr;
    uint32_t loss_cap = 0;
    uint16_t* bf = nullptr;
    float out_scale = 1.0f;
    uint32_t step = 0;
    std::string msg;
};

namespace {

constexpr float REFERENCE_FRAME_DELTA_SECS = 1.0f / 60.0f;   // cloth_compute.rs:65
constexpr uint32_t SUBSTEPS = 36;                             // cloth_compute.rs:59
constexpr uint32_t INNER_ITERS = 22;                          // cloth_compute.rs:61
constexpr float THICKNESS = 0.04f;                            // cloth_compute.rs:72
constexpr float DEFAULT_COLL_SCALE = 0.38f;                   // cloth_compute.rs:74
constexpr float DEFAULT_LINEAR_AIR_DRAG_PER_SEC = 1.25f;      // cloth_compute.rs:80

// Coloured GS forms: one persistent launch with grid barriers, one launch per
// colour and iteration, one persistent launch ordered by per-tile flags.
constexpr uint32_t GS_FORM_BARRIERS = 0, GS_FORM_COLOURS = 1, GS_FORM_TILES = 2;
// smallest GS flag tile (particles); the device tile arrays are sized for it
constexpr uint32_t GS_TILE_MIN = 64;
// fewest rows of an LDS grid tile for GS; its arrays are sized for it
constexpr uint32_t GSG_MIN_ROWS = 8;

uint32_t next_pow2(uint32_t v) {
    uint32_t p = 1;
    while (p < v && p < 0x80000000u) p <<= 1;
    return p;
}

}  // namespace

struct xpbd_cloth {
    xpbd_sim_config cfg{};
    xpbd_sim_uniforms u{};
    uint32_t n = 0, e = 0, t = 0, nslices = 0, num_batches = 0;
    std::vector<float> init_pos4;     // reset pose, w = inverse mass
    std::vector<uint32_t> batch_offsets;
    double rest_sq_sum = 0.0;
    bool has_ground = false;

    hipStream_t stream = nullptr;
    bool own_stream = false;
    int device = 0;

    std::vector<void*> allocs;
    float4 *X = nullptr, *V = nullptr, *A = nullptr, *B = nullptr, *V2 = nullptr, *rest4 = nullptr;
    float4* sP = nullptr;   // collision: positions + cell key in sorted order
    uint32_t* slot_rec = nullptr;   // collision: table slot filled at each sorted index (EMPTY_KEY: none)
    uint32_t *soff = nullptr, *swid = nullptr, *nbj = nullptr;
    float *nbr = nullptr, *nbc = nullptr;
    uint32_t* nbpk = nullptr;       // packed adjacency (nullptr: plain layout in use)
    uint32_t* nbtile = nullptr;     // packed adjacency with LDS-tile offsets (nullptr: no grid tiles)
    uint32_t tile_cols = 0, tile_rows = 0;
    // constraint-shared grid tiles (layout built at create when it applies;
    // used when js_on)
    SharedTiles js{};
    bool js_ok = false, js_tuned = false;
    uint32_t js_on = 0;
    float2* pairs = nullptr;
    uint32_t num_pairs = 0;
    uint8_t* ground = nullptr;
    uint2* gs_ij = nullptr;
    float2* gs_rc = nullptr;
    // the constraints again, ordered by lower endpoint (neighbouring loads):
    // the warm-start gate's residual
    uint2* near_ij = nullptr;
    float* near_rest = nullptr;
    uint32_t* gs_off = nullptr;
    uint32_t *keys = nullptr, *ids = nullptr, *keys2 = nullptr, *ids2 = nullptr, *hist = nullptr;
    uint4* table = nullptr;
    uint32_t table_slots = 0;
    uint32_t* epoch = nullptr;     // one cloth: the direct group table's pass tag (nullptr: hashed table)
    uint32_t* ovf = nullptr;       // collision overflow list (count, done, n sorted indices)
    float* aabb = nullptr;         // bounds partials, 8 floats per producing block / cloth
    uint32_t aabb_slots = 0;
    uint2* cspan = nullptr;         // batches: bounds pieces (particle range of one cloth each)
    uint32_t* cpfirst = nullptr;    // batches: first piece of each cloth (nclothes + 1)
    uint32_t ncspans = 0;
    CollGrid* grids = nullptr;     // one per cloth
    uint32_t nclothes = 1, local_bits = 32;
    uint32_t coll_digit_bits = 8;  // the collision sort's digits: 9 (3 passes) for one cloth, else 8
    uint32_t* cloth_of = nullptr;  // particle -> cloth (nullptr for one cloth)
    uint32_t* cloth_off = nullptr; // C + 1 vertex offsets
    uint32_t *tris = nullptr, *vt_off = nullptr, *vt_tri = nullptr;
    uint32_t *barrier = nullptr, *err = nullptr;
    double* resid = nullptr;       // per-block partial sums of the reported residual

    float4 *render_pos = nullptr, *render_nrm = nullptr;
    // imported render buffers (xpbd_import_render_buffers_fd): the imports
    // and their mapped ranges, released at the next bind or at destroy
    hipExternalMemory_t ext_mem[2] = {nullptr, nullptr};
    void* ext_ptr[2] = {nullptr, nullptr};

    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool graph_valid = false;

    uint32_t jacobi_block = 256, remap = 1, fused_predict = 0;
    uint32_t gs_block = 256, gs_blocks = 0;
    uint32_t gs_form_tuned = 0;     // the GS form the autotuner measured fastest (GS_FORM_*)
    uint32_t gs_form = 0;           // the GS form substeps run (GS_FORM_*)
    // GS with per-tile flags: the tiles in use, their grid, the device
    // capacity of the tile arrays (tiles of at least GS_TILE_MIN particles),
    // and the colour-sorted rows to plan tiles from (held while the handle is
    // created, read back from gs_ij / gs_rc for a later re-plan)
    GsTiles gt{};
    bool gt_ok = false;
    uint32_t gt_blocks = 0, gt_cap = 0, gt_deps = 0;
    std::vector<uint2> gs_ij_host;
    // ... and on LDS-resident grid tiles (one grid cloth): the plan, whether
    // it is usable (gg_ok) and used by the flag form (gg_on), the cloth's
    // grid, the capacity of coff / flag (tiles of GSG_MIN_ROWS rows)
    GsGrid gg{};
    bool gg_ok = false, gg_on = false;
    // Jacobi on LDS-resident grid tiles (k_jacobi_resident): the plan, whether
    // it can run (jr_ok), the autotuner's finding, whether substeps use it
    JResident jr{};
    bool jr_ok = false, jr_tuned = false;
    uint32_t jr_on = 0, jr_cap = 0;
    uint32_t gg_cols = 0, gg_rows = 0, gg_cap = 0;
    std::vector<float2> gs_rc_host;
    uint32_t pj_blocks = 0;        // persistent Jacobi grid (co-resident blocks)

    // fast correctly rounded division / sqrt in the Jacobi gather: the static
    // inputs' ranges (host-checked) and per-buffer "tame" flags of A and B
    float max_inv_mass = 0.f, max_comp = 0.f;
    bool mass_floor = false, comp_nonneg = false;   // fast-division preconditions (see fast_division_active)
    bool rest_tame = true;
    uint8_t *tameA = nullptr, *tameB = nullptr;   // per slice: 1 = coordinates in range
    uint8_t *tameW = nullptr, *tameS = nullptr;   // the same for the warm start's W0 / sP iterates
    uint2* srange = nullptr;                      // per slice: first / last neighbour slice

    // learned warm start (grid meshes): 8 grid neighbours and per-vertex
    // feature constants, built at create when the mesh carries cloth columns
    bool warm_ok = false;
    int32_t* nbr8 = nullptr;
    float4* vaux = nullptr;
    float4* W0 = nullptr;          // warm-start iterate
    uint32_t* warm_flag = nullptr; // [0] fallback flag, [1] warm substeps, [2] fallbacks
    uint32_t* warm_acc = nullptr;  // the gate's accumulators (warm_gate_words())
    xpbd_mlp* warm = nullptr;
    uint32_t warm_iters = 2;
    float warm_threshold = 1e-4f;

    // instrumented frames (xpbd_profile_frames): events around every launch of
    // one kernel kind, enqueued eagerly
    struct Probe {
        int which = -1;
        std::vector<hipEvent_t> ev;   // pairs
        size_t used = 0;
    } probe;

    std::string msg;
};

namespace {

int fail(xpbd_cloth* h, int code, const std::string& m) {
    if (h) h->msg = m;
    return code;
}

#define HIPC(h, expr)                                                                                        \
    do {                                                                                                     \
        hipError_t _e = (expr);                                                                              \
        if (_e != hipSuccess)                                                                                \
            return fail((h), _e == hipErrorOutOfMemory ? XPBD_E_NOMEM : XPBD_E_HIP,                        \
                        std::string(#expr) + ": " + hipGetErrorString(_e));                                 \
    } while (0)

template <class T>
int dalloc(xpbd_cloth* h, T** p, size_t count) {
    void* q = nullptr;
    size_t bytes = std::max<size_t>(count * sizeof(T), 256);
    HIPC(h, hipMalloc(&q, bytes));
    h->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return XPBD_OK;
}

template <class T>
int upload(xpbd_cloth* h, T* dst, const T* src, size_t count) {
    if (count == 0) return XPBD_OK;
    HIPC(h, hipMemcpy(dst, src, count * sizeof(T), hipMemcpyHostToDevice));
    return XPBD_OK;
}

#define TRY(expr)                      \
    do {                               \
        int _r = (expr);               \
        if (_r != XPBD_OK) return _r;  \
    } while (0)

DevParams dev_params(const xpbd_cloth* h) {
    const xpbd_sim_uniforms& u = h->u;
    DevParams p{};
    p.dt = u.dt;
    p.inv_dt = u.inv_dt;
    p.inv_dt_sq = u.inv_dt * u.inv_dt;                 // ClothSimParamsGpu::pack, cloth_compute.rs:325
    p.omega = u.jacobi_omega;
    p.gx = u.gravity[0];
    p.gy = u.gravity[1];
    p.gz = u.gravity[2];
    p.damp = std::exp(-u.linear_drag_per_sec * u.dt);  // xpbd_cpu.rs:402 (f32 expf)
    p.grab_x = u.grab_target[0];
    p.grab_y = u.grab_target[1];
    p.grab_z = u.grab_target[2];
    p.grab_k = u.grab_stiffness;
    p.grab_idx = u.grab_idx;
    p.grab_active = u.grab_active;
    p.thickness = u.thickness;
    p.coll_scale = u.coll_scale;
    p.n = h->n;
    p.nslices = h->nslices;
    return p;
}

// The shorter correctly rounded operations are exact for tame iterates when
// inverse masses <= 2^40, compliance / dt^2 <= 2^40 and every rest length is
// 0 or in [2^-30, 2^30] (the ranges the kernels' proof assumes).
bool fast_division_active(const xpbd_cloth* h) {
    if (!h->cfg.fast_division) return false;
    const float two40 = 1099511627776.0f;
    const float inv_dt_sq = h->u.inv_dt * h->u.inv_dt;
    return h->rest_tame && h->mass_floor && h->comp_nonneg && std::isfinite(h->max_inv_mass) &&
           h->max_inv_mass <= two40 && std::isfinite(inv_dt_sq) &&
           double(h->max_comp) * double(inv_dt_sq) <= double(two40);
}

uint8_t* tame_slot(xpbd_cloth* h, const float4* buf) {
    if (buf == h->A) return h->tameA;
    if (buf == h->B) return h->tameB;
    if (buf == h->W0) return h->tameW;
    if (buf == h->sP) return h->tameS;
    return nullptr;
}

JacobiArgs jacobi_args(const xpbd_cloth* h, const DevParams& p) {
    JacobiArgs a{};
    a.sell = Sell{h->soff, h->swid, h->nbj, h->nbr, h->nbc, h->nbpk, h->pairs, h->nbtile, h->tile_cols, h->tile_rows};
    if (h->js_on) a.sh = h->js;
    a.ground = h->has_ground ? h->ground : nullptr;
    a.aabb = h->aabb;
    a.p = p;
    a.remap = h->remap;
    a.fastdiv = fast_division_active(h) ? 1u : 0u;
    a.num_pairs = h->num_pairs;
    a.srange = h->srange;
    return a;
}

float coll_cell(const xpbd_cloth* h) {
    return h->cfg.collision_cell > 0.f ? std::max(h->cfg.collision_cell, h->u.thickness) : h->u.thickness;
}

// Collision chain on iterate F: grid, keys, sort, table, sweep (which also
// applies the ground plane and the velocity update).  Replaces the
// reference's clear_atomics .. collide_apply sequence (cloth_jacobi.rs:774-868).
// Cell grids and keys of the current iterate: one cloth's key kernel derives
// the grid from the nparts bounds partials its last iteration wrote; many
// cloths get their grids from bounds pieces first.
void enqueue_keys(xpbd_cloth* h, const float4* F, uint32_t nparts, float cell, hipStream_t s) {
    const bool one = h->nclothes == 1;
    if (!one)
        launch_coll_grids(F, h->cspan, h->ncspans, h->cpfirst, h->nclothes, h->aabb, cell, h->local_bits, h->grids,
                          s);
    launch_coll_keys(F, h->n, h->grids, h->cloth_of, h->keys, h->ids, h->table, h->slot_rec, h->hist,
                     h->coll_digit_bits, one ? h->aabb : nullptr, nparts, cell, h->epoch, s);
}

// Instrumented frames: bracket a launch of kind `which` with a pair of events.
struct ProbeScope {
    xpbd_cloth* h;
    bool on;
    ProbeScope(xpbd_cloth* hh, int which) : h(hh), on(hh->probe.which == which && hh->probe.used + 2 <= hh->probe.ev.size()) {
        if (on) (void)hipEventRecord(h->probe.ev[h->probe.used], h->stream);
    }
    ~ProbeScope() {
        if (on) {
            (void)hipEventRecord(h->probe.ev[h->probe.used + 1], h->stream);
            h->probe.used += 2;
        }
    }
};

void enqueue_collision(xpbd_cloth* h, const DevParams& p, const float4* F, uint32_t nparts, float4* xout,
                       float4* vout) {
    hipStream_t s = h->stream;
    enqueue_keys(h, F, nparts, coll_cell(h), s);
    const bool alt = launch_radix_sort(h->keys, h->ids, h->keys2, h->ids2, h->n, h->hist, s, true, h->coll_digit_bits,
                                       F, h->sP);
    launch_coll_table(alt ? h->keys2 : h->keys, h->n, h->ovf, h->table, h->table_slots - 1, h->slot_rec, h->epoch, s);
    CollArgs c{};
    c.F = F;
    c.sP = h->sP;
    c.rest4 = h->rest4;
    c.prev = h->X;
    c.xout = xout;
    c.vout = vout;
    c.ground = h->has_ground ? h->ground : nullptr;
    c.grids = h->grids;
    c.cloth_of = h->cloth_of;
    c.skeys = alt ? h->keys2 : h->keys;
    c.sids = alt ? h->ids2 : h->ids;
    c.table = h->table;
    c.table_mask = h->table_slots - 1;
    c.epoch = h->epoch;
    c.p = p;
    c.ovf = h->ovf;
    c.xcd_seg = SWEEP_XCD_SEG;
    ProbeScope ps(h, XPBD_KERNEL_COLLISION_SWEEP);
    launch_coll_sweep(c, s);
}

// Coloured GS as one launch per colour and iteration: predict X,V -> W, the
// colours, then either the state update (no collision) or the bounds
// partials of W for the collision pass (nparts is set).  The same steps as
// one persistent k_gs_substep, with launch boundaries for its grid barriers.
void enqueue_gs_colours(xpbd_cloth* h, const GsArgs& g, bool collision, uint32_t& nparts, hipStream_t s) {
    launch_predict(g.X, g.V, g.W, g.p, s);
    if (g.num_batches) launch_gs_colours(g, h->batch_offsets, g.iterations, s);
    if (collision) nparts = launch_aabb_of(g.W, h->n, h->aabb, s);
    else launch_finalize(g.W, g.xout, g.vout, g.ground, g.p, s);
}

// The GS part of a substep in the handle's form (gs_form); g.final_write must
// be !collision.  With collision, nparts = the bounds partials written.
void enqueue_gs(xpbd_cloth* h, const GsArgs& g, bool collision, uint32_t& nparts, hipStream_t s) {
    if (h->gs_form == GS_FORM_COLOURS) {
        enqueue_gs_colours(h, g, collision, nparts, s);
    } else if (h->gs_form == GS_FORM_TILES && h->gg_on) {
        launch_gs_grid(g, h->gg, s);
        nparts = h->gg.tx * h->gg.ty;
    } else if (h->gs_form == GS_FORM_TILES) {
        launch_gs_tiles(g, h->gt, h->gt_blocks, s);
        nparts = h->gt_blocks;
    } else {
        launch_zero_u32(h->barrier, GRID_SYNC_WORDS, s);
        launch_gs(g, h->gs_blocks, h->gs_block, s);
        nparts = h->gs_blocks;
    }
}

// One substep (cloth_jacobi.rs:702-883 / cloth_compute.rs:1313-1488).  The
// state is always (X, V) at substep boundaries; A/B hold the iterates.
void enqueue_substep(xpbd_cloth* h, bool collision) {
    hipStream_t s = h->stream;
    const DevParams p = dev_params(h);
    const uint32_t iters = std::max(h->cfg.solve_inner_iterations, 1u);
    const float4* F = nullptr;
    uint32_t nparts = 0;
    if (h->cfg.solver == XPBD_SOLVER_GAUSS_SEIDEL) {
        GsArgs g{};
        g.X = h->X;
        g.V = h->V;
        g.W = h->A;
        g.xout = h->X;
        g.vout = h->V;
        g.ground = h->has_ground ? h->ground : nullptr;
        g.aabb = h->aabb;
        g.ij = h->gs_ij;
        g.rc = h->gs_rc;
        g.batch_offsets = h->gs_off;
        g.num_batches = h->e ? h->num_batches : 0;
        g.iterations = iters;
        g.barrier = h->barrier;
        g.error = h->err;
        g.final_write = collision ? 0u : 1u;
        g.p = p;
        {
            ProbeScope ps(h, XPBD_KERNEL_GS_SUBSTEP);
            enqueue_gs(h, g, collision, nparts, s);
        }
        F = h->A;
    } else if (h->warm) {
        // learned warm start: predict -> (features, MLP) x0 = xp + MLP -> warm
        // iterations.  With a finite threshold, the relative residual of the
        // result gates one persistent plain substep (predict + I iterations)
        // that does nothing when the warm result is accepted.
        xpbd_mlp* m = h->warm;
        JacobiArgs a = jacobi_args(h, p);
        launch_predict(h->X, h->V, h->A, p, s);
        WarmFeatArgs fa{};
        fa.xp = h->A;
        fa.vel = h->V;
        fa.rest4 = h->rest4;
        fa.nbr8 = h->nbr8;
        fa.vaux = h->vaux;
        fa.dt = h->u.dt;
        fa.n = h->n;
        // x0 lands where the ping-pong (W0 / sP) ends in W0
        float4* cur = (h->warm_iters % 2) ? h->sP : h->W0;
        float4* nxt = (h->warm_iters % 2) ? h->W0 : h->sP;
        // (without a gate the launch counts the warm substep itself); x0's
        // slice tame bytes let the iterations take the short exact division
        const bool fast = a.fastdiv != 0;
        fa.tame_out = fast ? tame_slot(h, cur) : nullptr;
        launch_warm_infer(fa, m->P, m->bf, 1.0f / m->out_scale, cur,
                          std::isfinite(h->warm_threshold) ? nullptr : h->warm_flag, s);
        for (uint32_t k = 0; k < h->warm_iters; ++k) {
            a.src = cur;
            a.dst = nxt;
            a.tame_in = fast ? tame_slot(h, cur) : nullptr;
            a.tame_out = fast ? tame_slot(h, nxt) : nullptr;
            ProbeScope ps(h, XPBD_KERNEL_JACOBI_ITER);
            launch_jacobi(0, a, h->jacobi_block, s);
            std::swap(cur, nxt);
        }
        if (std::isfinite(h->warm_threshold)) {
            // gate (one launch, also clears the barrier words), then the
            // fallback: a persistent plain substep that runs only if the gate
            // rejected the warm result; the state comes from either
            launch_warm_gate(h->W0, h->near_ij, h->near_rest, h->e, h->rest_sq_sum, h->warm_threshold,
                             h->warm_acc, h->warm_flag, h->barrier, s);
            JacobiArgs g = jacobi_args(h, p);
            PersistArgs q{};
            q.X = h->X;
            q.V = h->V;
            q.A = h->A;
            q.B = h->B;
            q.iterations = iters;
            q.final_write = 0;
            q.barrier = h->barrier;
            q.error = h->err;
            q.gate = h->warm_flag;
            if (g.fastdiv) {
                q.tameA = h->tameA;
                q.tameB = h->tameB;
            }
            launch_jacobi_persistent(g, q, h->pj_blocks, s);
            float4* alt = (iters % 2 == 0) ? h->A : h->B;
            if (collision) {
                launch_copy4_if(alt, h->W0, h->n, h->warm_flag, s);
                F = h->W0;
                if (h->nclothes == 1) nparts = launch_aabb_of(F, h->n, h->aabb, s);
            } else {
                launch_finalize_sel(h->W0, alt, h->warm_flag, h->X, h->V, h->has_ground ? h->ground : nullptr, p,
                                    s);
            }
        } else if (collision) {
            F = h->W0;
            if (h->nclothes == 1) nparts = launch_aabb_of(F, h->n, h->aabb, s);
        } else {
            launch_finalize(h->W0, h->X, h->V, h->has_ground ? h->ground : nullptr, p, s);
        }
    } else if (h->jr_on) {
        // predict and all iterations in one launch on LDS-resident tiles; the
        // result lands in X,V (no collision) or, with one bounds partial per
        // tile, in A (even I) / B (odd I) for the collision pass.
        JacobiArgs a = jacobi_args(h, p);
        PersistArgs q{};
        q.X = h->X;
        q.V = h->V;
        q.A = h->A;
        q.B = h->B;
        q.iterations = iters;
        q.final_write = collision ? 0u : 1u;
        q.barrier = h->barrier;
        q.error = h->err;
        {
            ProbeScope ps(h, XPBD_KERNEL_JACOBI_PERSISTENT);
            launch_jacobi_resident(a, q, h->jr, s);
        }
        if (collision) {
            F = (iters % 2 == 0) ? h->A : h->B;
            nparts = (h->tile_cols / 64) * h->jr.ty;
        }
    } else if (h->cfg.persistent_jacobi == 1) {
        // predict and all iterations in one launch; the result lands in X,V
        // (no collision) or in A (even I) / B (odd I) for the collision pass.
        launch_zero_u32(h->barrier, GRID_SYNC_WORDS, s);
        JacobiArgs a = jacobi_args(h, p);
        a.vel = nullptr;
        PersistArgs q{};
        q.X = h->X;
        q.V = h->V;
        q.A = h->A;
        q.B = h->B;
        q.iterations = iters;
        q.final_write = collision ? 0u : 1u;
        q.barrier = h->barrier;
        q.error = h->err;
        if (a.fastdiv) {
            q.tameA = h->tameA;
            q.tameB = h->tameB;
        }
        {
            ProbeScope ps(h, XPBD_KERNEL_JACOBI_PERSISTENT);
            launch_jacobi_persistent(a, q, h->pj_blocks, s);
        }
        if (collision) {
            F = (iters % 2 == 0) ? h->A : h->B;
            if (h->nclothes == 1) nparts = launch_aabb_of(F, h->n, h->aabb, s);
        }
    } else if (!h->fused_predict) {
        // predict -> A, then iterations ping-pong A/B; the last one writes the
        // state in place (no collision) or the iterate F the collision reads.
        JacobiArgs a = jacobi_args(h, p);
        const bool fast = a.fastdiv != 0;
        launch_predict(h->X, h->V, h->A, p, s, fast ? tame_slot(h, h->A) : nullptr);
        float4* cur = h->A;
        float4* nxt = h->B;
        for (uint32_t k = 0; k < iters; ++k) {
            a.src = cur;
            a.tame_in = fast ? tame_slot(h, cur) : nullptr;
            a.tame_out = nullptr;
            if (k + 1 < iters) {
                a.dst = nxt;
                a.tame_out = fast ? tame_slot(h, nxt) : nullptr;
                ProbeScope ps(h, XPBD_KERNEL_JACOBI_ITER);
                launch_jacobi(0, a, h->jacobi_block, s);
                std::swap(cur, nxt);
            } else if (collision) {
                a.dst = nxt;
                nparts = launch_jacobi(h->nclothes == 1 ? JM_AABB : 0u, a, h->jacobi_block, s);
                F = nxt;
            } else {
                a.dst = h->X;
                a.vout = h->V;
                launch_jacobi(JM_FINAL, a, h->jacobi_block, s);
            }
        }
    } else {
        JacobiArgs a = jacobi_args(h, p);
        if (iters == 1) {
            a.src = h->X;
            a.vel = h->V;
            a.dst = h->A;
            nparts = launch_jacobi(collision ? (JM_PREDICT | JM_AABB) : JM_PREDICT, a, h->jacobi_block, s);
            if (!collision) launch_finalize(h->A, h->X, h->V, a.ground, p, s);
            F = h->A;
        } else {
            a.src = h->X;
            a.vel = h->V;
            a.dst = h->A;
            launch_jacobi(JM_PREDICT, a, h->jacobi_block, s);
            float4* cur = h->A;
            float4* nxt = h->B;
            for (uint32_t k = 1; k + 1 < iters; ++k) {
                a.src = cur;
                a.dst = nxt;
                ProbeScope ps(h, XPBD_KERNEL_JACOBI_ITER);
                launch_jacobi(0, a, h->jacobi_block, s);
                std::swap(cur, nxt);
            }
            a.src = cur;
            if (collision) {
                a.dst = nxt;
                nparts = launch_jacobi(JM_AABB, a, h->jacobi_block, s);
                F = nxt;
            } else {
                a.dst = h->X;
                a.vout = h->V;
                launch_jacobi(JM_FINAL, a, h->jacobi_block, s);
            }
        }
    }
    if (collision) enqueue_collision(h, p, F, nparts, h->X, h->V);
}

bool collision_this_substep(const xpbd_cloth* h, uint32_t si) {
    if (!(h->u.coll_scale > 1e-12f)) return false;   // cloth_jacobi.rs:699
    const uint32_t stride = std::max(h->cfg.collision_every_n_substeps, 1u);
    return si % stride == stride - 1;                  // cloth_jacobi.rs:703-704
}

void enqueue_frame(xpbd_cloth* h) {
    const uint32_t S = std::max(h->cfg.solve_substeps, 1u);
    for (uint32_t si = 0; si < S; ++si) enqueue_substep(h, collision_this_substep(h, si));
    if (h->render_pos || h->render_nrm)
        launch_publish(h->X, h->n, h->tris, h->vt_off, h->vt_tri, h->render_pos, h->render_nrm, h->stream);
}

void drop_graph(xpbd_cloth* h) {
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    h->gexec = nullptr;
    h->graph = nullptr;
    h->graph_valid = false;
}

// Unmap and release imported render buffers (the caller has synchronised).
void release_imports(xpbd_cloth* h) {
    for (int k = 0; k < 2; ++k) {
        if (h->ext_ptr[k]) (void)hipFree(h->ext_ptr[k]);
        if (h->ext_mem[k]) (void)hipDestroyExternalMemory(h->ext_mem[k]);
        if (h->render_pos == h->ext_ptr[k] && h->ext_ptr[k]) h->render_pos = nullptr;
        if (h->render_nrm == h->ext_ptr[k] && h->ext_ptr[k]) h->render_nrm = nullptr;
        h->ext_ptr[k] = nullptr;
        h->ext_mem[k] = nullptr;
    }
}

int check_device_flag(xpbd_cloth* h) {
    uint32_t flag = 0;
    HIPC(h, hipMemcpyAsync(&flag, h->err, sizeof(flag), hipMemcpyDeviceToHost, h->stream));
    HIPC(h, hipStreamSynchronize(h->stream));
    if (flag) {
        uint32_t zero = 0;
        HIPC(h, hipMemcpy(h->err, &zero, sizeof(zero), hipMemcpyHostToDevice));
        return fail(h, XPBD_E_STATE, "grid barrier wait exceeded its bound (blocks not co-resident)");
    }
    return XPBD_OK;
}

// Sliced-ELL packing of the reference CSR (neighbor_offsets/other/rest/
// compliance, mesh_prep.rs:588-634; pack_neighbor_gpu, cloth_jacobi.rs:95-111).
// Indices stay u32 (the reference packs them as f32, exact only below 2^24).
void build_sell(const Mesh& m, std::vector<uint32_t>& soff, std::vector<uint32_t>& swid, std::vector<uint32_t>& sj,
                std::vector<float>& sr, std::vector<float>& sc) {
    const size_t n = m.num_particles();
    const size_t ns = (n + 63) / 64;
    soff.assign(ns + 1, 0);
    swid.assign(ns, 0);
    for (size_t s = 0; s < ns; ++s) {
        uint32_t w = 0;
        for (size_t l = 0; l < 64 && s * 64 + l < n; ++l) {
            const size_t i = s * 64 + l;
            w = std::max(w, m.neighbor_offsets[i + 1] - m.neighbor_offsets[i]);
        }
        swid[s] = w;
        soff[s + 1] = soff[s] + w * 64;
    }
    const size_t total = soff[ns];
    sj.assign(total, 0u);
    sr.assign(total, 0.f);
    sc.assign(total, 0.f);
    for (size_t s = 0; s < ns; ++s)   // padding: the particle itself (zero length, skipped)
        for (size_t k = 0; k < swid[s]; ++k)
            for (size_t l = 0; l < 64; ++l) sj[soff[s] + k * 64 + l] = uint32_t(std::min(s * 64 + l, n - 1));
    for (size_t i = 0; i < n; ++i) {
        const size_t s = i / 64, l = i % 64;
        const uint32_t b = m.neighbor_offsets[i], d = m.neighbor_offsets[i + 1] - b;
        for (uint32_t k = 0; k < d; ++k) {
            const size_t e = soff[s] + size_t(k) * 64 + l;
            sj[e] = m.neighbor_other[b + k];
            sr[e] = m.neighbor_rest_len[b + k];
            sc[e] = m.neighbor_compliance[b + k];
        }
    }
}

// Packed adjacency: (pair << 16) | (uint16)(j - i) per entry, pair indexing
// the mesh's distinct (rest, compliance) values.  Returns false when a delta
// does not fit in 16 bits or there are more than 65535 distinct pairs.
bool pack_sell(const Mesh& m, const std::vector<uint32_t>& soff, std::vector<uint32_t>& packed,
               std::vector<float2>& pairs) {
    const size_t n = m.num_particles(), e = m.num_constraints();
    std::vector<uint64_t> key(e);
    for (size_t k = 0; k < e; ++k) {
        uint32_t rb, cb;
        std::memcpy(&rb, &m.constraint_rest[k], 4);
        std::memcpy(&cb, &m.constraint_comp[k], 4);
        key[k] = (uint64_t(rb) << 32) | cb;
    }
    std::vector<uint64_t> uniq(key);
    std::sort(uniq.begin(), uniq.end());
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    if (uniq.size() > 65535) return false;
    std::vector<uint16_t> pid(e);
    for (size_t k = 0; k < e; ++k)
        pid[k] = uint16_t(std::lower_bound(uniq.begin(), uniq.end(), key[k]) - uniq.begin());
    pairs.resize(uniq.size());
    for (size_t q = 0; q < uniq.size(); ++q) {
        const uint32_t rb = uint32_t(uniq[q] >> 32), cb = uint32_t(uniq[q]);
        std::memcpy(&pairs[q].x, &rb, 4);
        std::memcpy(&pairs[q].y, &cb, 4);
    }
    packed.assign(soff.back(), 0u);   // padding: delta 0, pair 0 -> the particle itself
    for (size_t i = 0; i < n; ++i) {
        const size_t s = i / 64, l = i % 64;
        const uint32_t b = m.neighbor_offsets[i], d = m.neighbor_offsets[i + 1] - b;
        for (uint32_t k = 0; k < d; ++k) {
            const int64_t delta = int64_t(m.neighbor_other[b + k]) - int64_t(i);
            if (delta < -32768 || delta > 32767) return false;
            const uint32_t q = pid[m.neighbor_constraint_id[b + k]];
            packed[soff[s] + size_t(k) * 64 + l] = (q << 16) | (uint32_t(delta) & 0xFFFFu);
        }
    }
    return true;
}

// Grid tiles: for one grid cloth with a column count C that is a multiple of
// 64 (so a 64-particle slice is 64 consecutive columns of one row), each
// packed entry's neighbour offset j - i = dy * C + dx (|dx|, |dy| <= 2, the
// neighbour inside the cloth) is re-encoded as dy * JT_W + dx, its offset in
// the block's LDS tile.  False (no tiled path) if any entry does not fit.
bool tile_sell(const Mesh& m, const std::vector<uint32_t>& soff, const std::vector<uint32_t>& packed,
               std::vector<uint32_t>& tiled, uint32_t& cols, uint32_t& rows) {
    const size_t n = m.num_particles();
    if (m.cloth_offsets.size() != 2 || m.cloth_cols.size() != 1) return false;
    const int64_t c = m.cloth_cols[0];
    if (c < 64 || c % 64 != 0 || n % size_t(c) != 0 || n == 0) return false;
    tiled.assign(packed.size(), 0u);
    for (size_t i = 0; i < n; ++i) {
        const size_t s = i / 64, l = i % 64;
        const uint32_t d = m.neighbor_offsets[i + 1] - m.neighbor_offsets[i];
        const int64_t col = int64_t(i) % c, row = int64_t(i) / c;
        for (uint32_t k = 0; k < d; ++k) {
            const uint32_t w = packed[soff[s] + size_t(k) * 64 + l];
            const int64_t delta = int64_t(int16_t(uint16_t(w & 0xFFFFu)));
            const int64_t j = int64_t(i) + delta, jr = j / c, jc = j % c;
            const int64_t dy = jr - row, dx = jc - col;
            if (dx < -2 || dx > 2 || dy < -2 || dy > 2) return false;
            const int64_t off = (dy * int64_t(JT_W) + dx) * int64_t(sizeof(float4));   // bytes
            const uint32_t pair = w >> 16;
            if (pair >= 4096u) return false;   // pair x 16 must fit the high half
            tiled[soff[s] + size_t(k) * 64 + l] = ((pair * uint32_t(sizeof(float4))) << 16) | (uint32_t(off) & 0xFFFFu);
        }
    }
    cols = uint32_t(c);
    rows = uint32_t(n / size_t(c));
    return true;
}

// Constraint-shared grid tiles (the SharedTiles layout in kernels.hpp), from
// the packed entries of a grid cloth: every entry (i, j) of particle i is
// the constraint anchored at i in direction j - i (forward) or at j in
// direction i - j.  False (no shared path) when the mesh needs more than
// JS_DIRS directions, two constraints of one anchor and direction differ in
// their pair, the pairs exceed JS_PAIRS_MAX, or a code does not fit 15 bits.
struct SharedLayout {
    std::vector<uint16_t> jpair;
    std::vector<uint32_t> jobs, entries, eoff;
    uint32_t njobs_main = 0, njobs_edge = 0, nslots = 0;
};

bool share_sell(const Mesh& m, const std::vector<uint32_t>& soff, const std::vector<uint32_t>& swid,
                const std::vector<uint32_t>& packed, uint32_t num_pairs, uint32_t cols, uint32_t rows,
                SharedLayout& L) {
    const size_t n = m.num_particles();
    if (num_pairs > JS_PAIRS_MAX || size_t(cols) * rows != n || n == 0) return false;
    const int64_t C = cols;
    auto forward = [](int64_t dy, int64_t dx) { return dy > 0 || (dy == 0 && dx > 0); };
    // directions, sorted by (dy, dx)
    std::vector<std::pair<int, int>> dirs;
    for (size_t i = 0; i < n; ++i) {
        const size_t s = i / 64, l = i % 64;
        const uint32_t d = m.neighbor_offsets[i + 1] - m.neighbor_offsets[i];
        for (uint32_t k = 0; k < d; ++k) {
            const uint32_t w = packed[soff[s] + size_t(k) * 64 + l];
            const int64_t j = int64_t(i) + int64_t(int16_t(uint16_t(w & 0xFFFFu)));
            int64_t dy = j / C - int64_t(i) / C, dx = j % C - int64_t(i) % C;
            if (!forward(dy, dx)) {
                dy = -dy;
                dx = -dx;
            }
            if (dx < -2 || dx > 2 || dy > 2) return false;
            const std::pair<int, int> dd{int(dy), int(dx)};
            if (std::find(dirs.begin(), dirs.end(), dd) == dirs.end()) {
                if (dirs.size() == JS_DIRS) return false;
                dirs.push_back(dd);
            }
        }
    }
    std::sort(dirs.begin(), dirs.end());
    const size_t nd = dirs.size();
    auto dir_of = [&](int64_t dy, int64_t dx) {
        return uint32_t(std::find(dirs.begin(), dirs.end(), std::pair<int, int>(int(dy), int(dx))) - dirs.begin());
    };
    // slot rows: direction t, tile rows lr in [2 - dy_t, 2 + JT_ROWS)
    std::vector<uint32_t> row_base(nd);
    uint32_t nrows = 0;
    for (size_t t = 0; t < nd; ++t) {
        row_base[t] = nrows;
        nrows += JT_ROWS + uint32_t(dirs[t].first);
    }
    auto slot_row = [&](uint32_t t, uint32_t lr) { return row_base[t] + lr - (2u - uint32_t(dirs[t].first)); };
    L.nslots = nrows * JT_W;
    // pair of every (direction, anchor)
    L.jpair.assign(nd * n, 0xFFFFu);
    for (size_t i = 0; i < n; ++i) {
        const size_t s = i / 64, l = i % 64;
        const uint32_t d = m.neighbor_offsets[i + 1] - m.neighbor_offsets[i];
        for (uint32_t k = 0; k < d; ++k) {
            const uint32_t w = packed[soff[s] + size_t(k) * 64 + l];
            const int64_t j = int64_t(i) + int64_t(int16_t(uint16_t(w & 0xFFFFu)));
            int64_t dy = j / C - int64_t(i) / C, dx = j % C - int64_t(i) % C;
            int64_t anchor = int64_t(i);
            if (!forward(dy, dx)) {
                dy = -dy;
                dx = -dx;
                anchor = j;
            }
            uint16_t& q = L.jpair[size_t(dir_of(dy, dx)) * n + size_t(anchor)];
            const uint16_t pair = uint16_t(w >> 16);
            if (q != 0xFFFFu && q != pair) return false;
            q = pair;
        }
    }
    // jobs: every slot row over the tile's 64 columns, then the anchors left
    // (dx > 0) or right (dx < 0) of the tile whose partner is inside
    auto item = [&](uint32_t t, uint32_t lr, uint32_t cl) {
        return uint32_t(dirs[t].first) | (uint32_t(dirs[t].second + 2) << 2) | (t << 5) | (lr << 8) | (cl << 12) |
               (slot_row(t, lr) << 20);
    };
    L.jobs.clear();
    for (uint32_t t = 0; t < nd; ++t)
        for (uint32_t lr = 2 - uint32_t(dirs[t].first); lr < 2 + JT_ROWS; ++lr) L.jobs.push_back(item(t, lr, 0));
    L.njobs_main = uint32_t(L.jobs.size());
    std::vector<uint32_t> edge;
    for (uint32_t t = 0; t < nd; ++t) {
        const int dy = dirs[t].first, dx = dirs[t].second;
        if (dx == 0) continue;
        for (uint32_t lr = 2 - uint32_t(dy); lr < 2 + JT_ROWS - uint32_t(dy); ++lr) {
            if (dx > 0)
                for (int cl = 2 - dx; cl < 2; ++cl) edge.push_back(item(t, lr, uint32_t(cl)));
            else
                for (int cl = 66; cl < 66 - dx; ++cl) edge.push_back(item(t, lr, uint32_t(cl)));
        }
    }
    L.njobs_edge = uint32_t((edge.size() + 63) / 64);
    edge.resize(size_t(L.njobs_edge) * 64, 0xFFFFFFFFu);
    L.jobs.insert(L.jobs.end(), edge.begin(), edge.end());
    // entry codes (slot - lane) | sign << 15, two per word
    const size_t ns = swid.size();
    L.eoff.assign(ns + 1, 0u);
    for (size_t s = 0; s < ns; ++s) L.eoff[s + 1] = L.eoff[s] + (swid[s] + 1) / 2 * 64;
    L.entries.assign(L.eoff[ns], 0u);
    for (size_t i = 0; i < n; ++i) {
        const size_t s = i / 64, l = i % 64;
        const int64_t row = int64_t(i) / C, col = int64_t(i) % C;
        const uint32_t lr_i = uint32_t(row % JT_ROWS) + 2, cl_i = uint32_t(col % 64) + 2;
        const uint32_t d = m.neighbor_offsets[i + 1] - m.neighbor_offsets[i];
        for (uint32_t k = 0; k < swid[s]; ++k) {
            uint32_t code;
            if (k < d) {
                const uint32_t w = packed[soff[s] + size_t(k) * 64 + l];
                const int64_t j = int64_t(i) + int64_t(int16_t(uint16_t(w & 0xFFFFu)));
                int64_t dy = j / C - row, dx = j % C - col;
                uint32_t lr = lr_i, cl = cl_i, sign = 0;
                if (!forward(dy, dx)) {   // anchored at j
                    lr = uint32_t(int64_t(lr_i) + dy);
                    cl = uint32_t(int64_t(cl_i) + dx);
                    dy = -dy;
                    dx = -dx;
                    sign = 1;
                }
                const uint32_t slot = slot_row(dir_of(dy, dx), lr) * JT_W + cl;
                if (slot - uint32_t(l) >= 0x8000u) return false;
                code = (slot - uint32_t(l)) | (sign << 15);
            } else {   // padding: the zero slot
                if (L.nslots - uint32_t(l) >= 0x8000u) return false;
                code = L.nslots - uint32_t(l);
            }
            uint32_t& word = L.entries[L.eoff[s] + size_t(k / 2) * 64 + l];
            word |= (k & 1u) ? (code << 16) : code;
        }
    }
    return true;
}

int validate_mesh(const Mesh& m, std::string& why) {
    const size_t n = m.num_particles();
    if (m.positions.size() != 3 * n || m.rest_positions.size() != 3 * n) return why = "position arrays", XPBD_E_INVALID;
    if (m.neighbor_offsets.size() != n + 1) return why = "neighbor_offsets size", XPBD_E_INVALID;
    if (n > (1u << 29)) return why = "too many particles (at most 2^29)", XPBD_E_INVALID;
    for (size_t k = 0; k < m.num_constraints(); ++k)
        if (m.constraint_i[k] >= n || m.constraint_j[k] >= n || m.constraint_i[k] == m.constraint_j[k])
            return why = "constraint endpoint out of range", XPBD_E_INVALID;
    for (uint32_t v : m.neighbor_other)
        if (v >= n) return why = "neighbor index out of range", XPBD_E_INVALID;
    for (size_t i = 0; i < n; ++i)
        if (m.neighbor_offsets[i] > m.neighbor_offsets[i + 1]) return why = "neighbor_offsets not monotone", XPBD_E_INVALID;
    if (m.neighbor_offsets[n] != m.neighbor_other.size()) return why = "neighbor_offsets end", XPBD_E_INVALID;
    for (uint32_t v : m.indices)
        if (v >= n) return why = "triangle index out of range", XPBD_E_INVALID;
    if (m.batch_offsets.empty() || m.batch_offsets.front() != 0 || m.batch_offsets.back() != m.num_constraints())
        return why = "batch offsets", XPBD_E_INVALID;
    for (size_t b = 0; b + 1 < m.batch_offsets.size(); ++b)
        if (m.batch_offsets[b] > m.batch_offsets[b + 1]) return why = "batch offsets not monotone", XPBD_E_INVALID;
    return XPBD_OK;
}

// Tiles of `tile` consecutive particles for the per-tile-flag GS substep
// (GsTiles): each colour's run of constraints per tile, and each tile's
// dependency list, the other tiles U whose span [first particle, highest
// endpoint of U's constraints] overlaps the tile's own.  Constraints whose
// particles two tiles share make the spans overlap, so a tile that waits for
// these tiles' previous step reads every particle after its last write.
// Fails (false) when a list would exceed GS_TILE_DEPS_MAX or a colour is not
// sorted by lower endpoint.
bool plan_gs_tiles(const std::vector<uint2>& ij, const std::vector<uint32_t>& off, uint32_t n, uint32_t tile,
                   std::vector<uint32_t>& toff, std::vector<uint32_t>& doff, std::vector<uint32_t>& dep,
                   uint32_t& max_deps) {
    const uint32_t nt = (n + tile - 1) / tile;
    const size_t nb = off.size() - 1;
    toff.assign(nb * (size_t(nt) + 1), 0u);
    std::vector<uint32_t> reach(nt);
    for (uint32_t t = 0; t < nt; ++t) reach[t] = std::min(n, (t + 1) * tile) - 1;
    for (size_t b = 0; b < nb; ++b) {
        uint32_t* to = toff.data() + b * (size_t(nt) + 1);
        size_t k = off[b];
        uint32_t prev_lo = 0;
        for (size_t q = off[b]; q < off[b + 1]; ++q) {
            const uint32_t lo = std::min(ij[q].x, ij[q].y), hi = std::max(ij[q].x, ij[q].y);
            if (lo < prev_lo || hi >= n) return false;
            prev_lo = lo;
            reach[lo / tile] = std::max(reach[lo / tile], hi);
        }
        for (uint32_t t = 0; t <= nt; ++t) {
            const uint64_t first = uint64_t(t) * tile;
            while (k < off[b + 1] && std::min(ij[k].x, ij[k].y) < first) ++k;
            to[t] = uint32_t(k);
        }
    }
    std::vector<uint32_t> cnt(nt, 0u);
    for (uint32_t u = 0; u < nt; ++u) {
        const uint32_t last = reach[u] / tile;   // tiles u+1..last overlap u
        if (last - u > GS_TILE_DEPS_MAX) return false;
        for (uint32_t t = u + 1; t <= last; ++t) {
            ++cnt[u];
            ++cnt[t];
        }
    }
    doff.assign(size_t(nt) + 1, 0u);
    max_deps = 0;
    for (uint32_t t = 0; t < nt; ++t) {
        if (cnt[t] > GS_TILE_DEPS_MAX) return false;
        max_deps = std::max(max_deps, cnt[t]);
        doff[t + 1] = doff[t] + cnt[t];
    }
    dep.assign(doff[nt], 0u);
    std::vector<uint32_t> cur(doff.begin(), doff.end() - 1);
    for (uint32_t u = 0; u < nt; ++u)
        for (uint32_t t = u + 1; t <= reach[u] / tile; ++t) {
            dep[cur[u]++] = t;
            dep[cur[t]++] = u;
        }
    return true;
}

// Plan GS flag tiles of `tile` particles and upload them into the handle's
// tile arrays; gt_ok says whether they are usable.  The grid is the tiles,
// within what can be resident at once and the bounds partial slots.
// The colour-sorted constraint rows on the host (read back from the device
// after create released them).
int gs_host_rows(xpbd_cloth* h) {
    if (h->gs_ij_host.size() == h->e && h->gs_rc_host.size() == h->e) return XPBD_OK;
    h->gs_ij_host.resize(h->e);
    h->gs_rc_host.resize(h->e);
    if (h->e) {
        HIPC(h, hipMemcpy(h->gs_ij_host.data(), h->gs_ij, size_t(h->e) * sizeof(uint2), hipMemcpyDeviceToHost));
        HIPC(h, hipMemcpy(h->gs_rc_host.data(), h->gs_rc, size_t(h->e) * sizeof(float2), hipMemcpyDeviceToHost));
    }
    return XPBD_OK;
}

void release_gs_host_rows(xpbd_cloth* h) {
    std::vector<uint2>().swap(h->gs_ij_host);
    std::vector<float2>().swap(h->gs_rc_host);
}

int build_gs_tiles(xpbd_cloth* h, uint32_t tile, uint32_t blocks = 0) {
    h->gt_ok = false;
    if (!h->gt_cap || tile < GS_TILE_MIN) return XPBD_OK;
    TRY(gs_host_rows(h));
    std::vector<uint32_t> toff, doff, dep;
    uint32_t max_deps = 0;
    if (!plan_gs_tiles(h->gs_ij_host, h->batch_offsets, h->n, tile, toff, doff, dep, max_deps)) return XPBD_OK;
    const uint32_t nt = uint32_t(doff.size() - 1);
    TRY(upload(h, const_cast<uint32_t*>(h->gt.toff), toff.data(), toff.size()));
    TRY(upload(h, const_cast<uint32_t*>(h->gt.doff), doff.data(), doff.size()));
    TRY(upload(h, const_cast<uint32_t*>(h->gt.dep), dep.data(), dep.size()));
    h->gt.ntiles = nt;
    h->gt.tile = tile;
    h->gt_deps = max_deps;
    const uint32_t cap = std::min(std::max(1u, gs_tiles_max_blocks()), h->aabb_slots);
    h->gt_blocks = std::max(1u, std::min({nt, cap, blocks ? blocks : nt}));
    h->gt_ok = true;
    return XPBD_OK;
}

// LDS grid tiles for k_gs_grid (GsGrid): tiles of 64 columns x th rows of one
// grid cloth; each constraint goes to the tile of its lower endpoint, grouped
// by (tile, colour, touches a shared position), each group in the
// colour-sorted order, with its endpoints as window positions ((th + 4) x 68,
// the tile and 2 particles around it; bit 15: the position is within 2 of the
// tile's edge or outside it, where other tiles' constraints reach).  False
// when an endpoint falls outside its tile's window.
// Group of a grid-tile constraint: (tile, colour, 1 if an endpoint is shared).
size_t grid_group(uint32_t T, uint32_t c, uint32_t nc, uint32_t pos) {
    return (size_t(T) * nc + c) * 2 + ((pos & 0x80008000u) ? 1u : 0u);
}

bool plan_gs_grid(const std::vector<uint2>& ij, const std::vector<float2>& rc, const std::vector<uint32_t>& off,
                  uint32_t cols, uint32_t rows, uint32_t th, std::vector<uint32_t>& coff, std::vector<uint32_t>& con,
                  std::vector<float2>& crc) {
    if (cols < 64 || cols % 64 || rows == 0 || th < GSG_MIN_ROWS || off.size() < 2) return false;
    const uint32_t tx = cols / 64, ty = (rows + th - 1) / th, nt = tx * ty, nc = uint32_t(off.size() - 1);
    const size_t e = ij.size();
    std::vector<uint32_t> tile(e), pos(e);
    coff.assign(size_t(nt) * nc * 2 + 1, 0u);
    for (uint32_t c = 0; c < nc; ++c)
        for (size_t k = off[c]; k < off[c + 1]; ++k) {
            const uint32_t lo = std::min(ij[k].x, ij[k].y);
            if (lo >= cols * rows || std::max(ij[k].x, ij[k].y) >= cols * rows) return false;
            const uint32_t T = (lo / cols / th) * tx + (lo % cols) / 64;
            const int r0 = int((T / tx) * th), c0 = int((T % tx) * 64);
            uint32_t w[2];
            const uint32_t q[2] = {ij[k].x, ij[k].y};
            for (int s = 0; s < 2; ++s) {
                const int lr = int(q[s] / cols) - r0 + 2, lc = int(q[s] % cols) - c0 + 2;
                if (lr < 0 || lr >= int(th) + 4 || lc < 0 || lc >= 68) return false;
                const bool shared = lr < 4 || lr >= int(th) || lc < 4 || lc >= 64;
                w[s] = uint32_t(lr * 68 + lc) | (shared ? 0x8000u : 0u);
            }
            tile[k] = T;
            pos[k] = w[0] | (w[1] << 16);
            ++coff[grid_group(T, c, nc, pos[k])];
        }
    uint32_t run = 0;   // counts -> starts; the last entry: the end
    for (size_t g = 0; g < coff.size(); ++g) {
        const uint32_t cnt = coff[g];
        coff[g] = run;
        run += cnt;
    }
    con.assign(e, 0u);
    crc.assign(e, make_float2(0.f, 0.f));
    std::vector<uint32_t> cur(coff);
    for (uint32_t c = 0; c < nc; ++c)
        for (size_t k = off[c]; k < off[c + 1]; ++k) {
            const uint32_t d = cur[grid_group(tile[k], c, nc, pos[k])]++;
            con[d] = pos[k];
            crc[d] = rc[k];
        }
    return true;
}

// Plan LDS grid tiles of th rows and upload them; gg_ok says whether they
// are usable (a plan, the LDS, and every tile resident at once).
int build_gs_grid(xpbd_cloth* h, uint32_t th) {
    h->gg_ok = false;
    if (!h->gg_cap || th < GSG_MIN_ROWS) return XPBD_OK;
    int dev = 0, lds = 0;
    HIPC(h, hipGetDevice(&dev));
    HIPC(h, hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    if (gs_grid_lds_bytes(th) > uint32_t(lds)) return XPBD_OK;
    TRY(gs_host_rows(h));
    std::vector<uint32_t> coff, con;
    std::vector<float2> crc;
    if (!plan_gs_grid(h->gs_ij_host, h->gs_rc_host, h->batch_offsets, h->gg_cols, h->gg_rows, th, coff, con, crc))
        return XPBD_OK;
    HIPC(h, gs_grid_prepare(th));
    const uint32_t nt = (h->gg_cols / 64) * ((h->gg_rows + th - 1) / th);
    if (nt > gs_grid_max_blocks(th) || nt > h->aabb_slots) return XPBD_OK;
    TRY(upload(h, const_cast<uint32_t*>(h->gg.coff), coff.data(), coff.size()));
    TRY(upload(h, const_cast<uint32_t*>(h->gg.con), con.data(), con.size()));
    TRY(upload(h, const_cast<float2*>(h->gg.crc), crc.data(), crc.size()));
    h->gg.tx = h->gg_cols / 64;
    h->gg.ty = (h->gg_rows + th - 1) / th;
    h->gg.cols = h->gg_cols;
    h->gg.rows = h->gg_rows;
    h->gg.th = th;
    h->gg.colours = h->num_batches;
    h->gg_ok = true;
    return XPBD_OK;
}

// Plan the LDS-resident Jacobi tiles of th rows (grid-tile layout only):
// jr_ok when the window pair and the staged pairs fit the LDS and every tile
// can be resident at once.
int build_jacobi_resident(xpbd_cloth* h, uint32_t th) {
    h->jr_ok = false;
    if (!h->nbtile || !h->jr.flag || !h->tile_cols || th < 16 || th > 64) return XPBD_OK;
    const uint32_t ty = (h->tile_rows + th - 1) / th, nt = (h->tile_cols / 64) * ty;
    if (nt > h->jr_cap || nt > h->aabb_slots) return XPBD_OK;
    int dev = 0, lds = 0;
    HIPC(h, hipGetDevice(&dev));
    HIPC(h, hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const uint32_t static_lds = 1024;   // the kernel's own (grid set-up, block AND)
    uint32_t npairs = h->num_pairs <= JT_PAIRS_MAX ? h->num_pairs : 0u;
    if (jacobi_resident_lds_bytes(th, npairs) + static_lds > uint32_t(lds)) npairs = 0;
    if (jacobi_resident_lds_bytes(th, npairs) + static_lds > uint32_t(lds)) return XPBD_OK;
    HIPC(h, jacobi_resident_prepare(th, npairs));
    if (nt > jacobi_resident_max_blocks(th, npairs)) return XPBD_OK;
    h->jr.th = th;
    h->jr.ty = ty;
    h->jr.npairs = npairs;
    h->jr.aabb = h->aabb;
    h->jr_ok = true;
    return XPBD_OK;
}

int create_device(xpbd_cloth* h, const Mesh& m) {
    const size_t n = h->n, e = h->e, t = h->t;
    TRY(dalloc(h, &h->X, n));
    TRY(dalloc(h, &h->V, n));
    TRY(dalloc(h, &h->A, n));
    TRY(dalloc(h, &h->B, n));
    TRY(dalloc(h, &h->V2, n));
    TRY(dalloc(h, &h->rest4, n));
    TRY(dalloc(h, &h->sP, n));
    std::vector<float> rest4(4 * n);
    h->init_pos4.assign(4 * n, 0.f);
    for (size_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) {
            rest4[4 * i + a] = m.rest_positions[3 * i + a];
            h->init_pos4[4 * i + a] = m.positions[3 * i + a];
        }
        rest4[4 * i + 3] = m.inv_mass[i];
        h->init_pos4[4 * i + 3] = m.inv_mass[i];
    }
    TRY(upload(h, reinterpret_cast<float*>(h->rest4), rest4.data(), rest4.size()));

    std::vector<uint32_t> soff, swid, sj;
    std::vector<float> sr, sc;
    build_sell(m, soff, swid, sj, sr, sc);
    TRY(dalloc(h, &h->soff, soff.size()));
    TRY(dalloc(h, &h->swid, swid.size()));
    TRY(dalloc(h, &h->nbj, sj.size()));
    TRY(dalloc(h, &h->nbr, sr.size()));
    TRY(dalloc(h, &h->nbc, sc.size()));
    TRY(upload(h, h->soff, soff.data(), soff.size()));
    TRY(upload(h, h->swid, swid.data(), swid.size()));
    TRY(upload(h, h->nbj, sj.data(), sj.size()));
    TRY(upload(h, h->nbr, sr.data(), sr.size()));
    TRY(upload(h, h->nbc, sc.data(), sc.size()));
    {   // neighbour slice range of every slice (its own slice included)
        std::vector<uint2> rg(swid.size());
        for (size_t sl = 0; sl < swid.size(); ++sl) {
            uint32_t lo = uint32_t(sl), hi = uint32_t(sl);
            for (uint32_t e = soff[sl]; e < soff[sl + 1]; ++e) {
                lo = std::min(lo, sj[e] / 64);
                hi = std::max(hi, sj[e] / 64);
            }
            rg[sl] = make_uint2(lo, hi);
        }
        TRY(dalloc(h, &h->srange, rg.size()));
        TRY(upload(h, h->srange, rg.data(), rg.size()));
    }
    if (!h->cfg.plain_adjacency) {
        std::vector<uint32_t> pk;
        std::vector<float2> pairs;
        if (pack_sell(m, soff, pk, pairs)) {
            TRY(dalloc(h, &h->nbpk, pk.size()));
            TRY(dalloc(h, &h->pairs, pairs.size()));
            TRY(upload(h, h->nbpk, pk.data(), pk.size()));
            TRY(upload(h, h->pairs, pairs.data(), pairs.size()));
            h->num_pairs = uint32_t(pairs.size());
            std::vector<uint32_t> tk;
            uint32_t tc = 0, tr = 0;
            if (h->cfg.grid_tiles && tile_sell(m, soff, pk, tk, tc, tr)) {
                TRY(dalloc(h, &h->nbtile, tk.size()));
                TRY(upload(h, h->nbtile, tk.data(), tk.size()));
                h->tile_cols = tc;
                h->tile_rows = tr;
                SharedLayout L;
                int lds_max = 0;
                (void)hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device);
                if (h->cfg.grid_tiles != 2 && share_sell(m, soff, swid, pk, h->num_pairs, tc, tr, L) &&
                    jacobi_shared_lds_bytes(L.nslots, h->num_pairs) <= size_t(lds_max)) {
                    uint16_t* jp = nullptr;
                    uint32_t *jobs = nullptr, *ent = nullptr, *eoff = nullptr;
                    TRY(dalloc(h, &jp, L.jpair.size()));
                    TRY(dalloc(h, &jobs, L.jobs.size()));
                    TRY(dalloc(h, &ent, L.entries.size()));
                    TRY(dalloc(h, &eoff, L.eoff.size()));
                    TRY(upload(h, jp, L.jpair.data(), L.jpair.size()));
                    TRY(upload(h, jobs, L.jobs.data(), L.jobs.size()));
                    TRY(upload(h, ent, L.entries.data(), L.entries.size()));
                    TRY(upload(h, eoff, L.eoff.data(), L.eoff.size()));
                    h->js = SharedTiles{jp, jobs, ent, eoff, L.njobs_main, L.njobs_edge, L.nslots};
                    h->js_ok = true;
                }
            }
        }
    }

    h->has_ground = std::any_of(m.ground.begin(), m.ground.end(), [](uint32_t g) { return g != 0; });
    TRY(dalloc(h, &h->ground, n));
    {
        std::vector<uint8_t> g(n, 0);
        for (size_t i = 0; i < n && i < m.ground.size(); ++i) g[i] = m.ground[i] ? 1 : 0;
        TRY(upload(h, h->ground, g.data(), n));
    }

    std::vector<uint2> ij(e);
    std::vector<float2> rc(e);
    h->rest_sq_sum = 0.0;
    for (size_t k = 0; k < e; ++k) {
        ij[k] = make_uint2(m.constraint_i[k], m.constraint_j[k]);
        rc[k] = make_float2(m.constraint_rest[k], m.constraint_comp[k]);
        h->rest_sq_sum += double(m.constraint_rest[k]) * double(m.constraint_rest[k]);
    }
    {
        // Coloured GS reads the rows colour by colour.  Within a colour the
        // constraints share no vertex, so their order is free: sorted by
        // (lower, higher) endpoint, neighbouring threads touch neighbouring
        // particles instead of the colouring's degree-first order.
        std::vector<uint32_t> ord(e);
        for (size_t k = 0; k < e; ++k) ord[k] = uint32_t(k);
        auto key = [&](uint32_t k) {
            const uint32_t a = ij[k].x, b = ij[k].y;
            return (uint64_t(std::min(a, b)) << 32) | std::max(a, b);
        };
        const std::vector<uint32_t>& off = m.batch_offsets;
        for (size_t c = 0; c + 1 < off.size(); ++c)
            std::sort(ord.begin() + off[c], ord.begin() + off[c + 1],
                      [&](uint32_t a, uint32_t b) { return key(a) < key(b); });
        std::vector<uint2> gij(e);
        std::vector<float2> grc(e);
        for (size_t k = 0; k < e; ++k) {
            gij[k] = ij[ord[k]];
            grc[k] = rc[ord[k]];
        }
        TRY(dalloc(h, &h->gs_ij, e));
        TRY(dalloc(h, &h->gs_rc, e));
        TRY(upload(h, h->gs_ij, gij.data(), e));
        TRY(upload(h, h->gs_rc, grc.data(), e));
        h->gs_ij_host = std::move(gij);
        h->gs_rc_host = std::move(grc);
    }
    {
        std::vector<uint32_t> ord(e);
        for (size_t k = 0; k < e; ++k) ord[k] = uint32_t(k);
        auto lo = [&](uint32_t k) { return std::min(ij[k].x, ij[k].y); };
        std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return lo(a) < lo(b); });
        std::vector<uint2> nij(e);
        std::vector<float> nrest(e);
        for (size_t k = 0; k < e; ++k) {
            nij[k] = ij[ord[k]];
            nrest[k] = rc[ord[k]].x;
        }
        TRY(dalloc(h, &h->near_ij, e));
        TRY(dalloc(h, &h->near_rest, e));
        TRY(upload(h, h->near_ij, nij.data(), e));
        TRY(upload(h, h->near_rest, nrest.data(), e));
    }
    h->batch_offsets = m.batch_offsets;
    h->num_batches = uint32_t(m.batch_offsets.size() - 1);
    TRY(dalloc(h, &h->gs_off, h->batch_offsets.size()));
    TRY(upload(h, h->gs_off, h->batch_offsets.data(), h->batch_offsets.size()));

    TRY(dalloc(h, &h->keys, n));
    TRY(dalloc(h, &h->ids, n));
    TRY(dalloc(h, &h->keys2, n));
    TRY(dalloc(h, &h->ids2, n));
    TRY(dalloc(h, &h->hist, radix_hist_words(uint32_t(n))));
    // the cloth count as settled below (offsets that do not cover [0, n) in
    // two or more pieces make one cloth)
    const bool one_cloth = !(m.cloth_offsets.size() >= 3 && m.cloth_offsets.front() == 0 && m.cloth_offsets.back() == n);
    if (one_cloth) {
        // direct group table: entry g for 8-cell group g (keys < 2^27), valid
        // when its tag is the pass's; tag 0 is never current, so zeros are
        // empty entries and nothing is cleared between passes (256 MB)
        h->table_slots = ONE_CLOTH_GROUPS;
        TRY(dalloc(h, &h->table, h->table_slots));
        HIPC(h, hipMemset(h->table, 0, size_t(h->table_slots) * sizeof(uint4)));
        TRY(dalloc(h, &h->epoch, 1));
        HIPC(h, hipMemset(h->epoch, 0, sizeof(uint32_t)));
    } else {
        // hashed: one entry per occupied 8-cell group: at most n groups (a
        // stretched sheet puts nearly every particle in its own group), so
        // load <= 1/4; >= 4 slots per particle keeps the sweep's probe chains
        // short (the hash table notes in kernels.hip).  next_pow2 caps at
        // 2^31 slots, which still leaves load <= 1/4 up to the particle limit
        // checked at creation.
        h->table_slots = next_pow2(uint32_t(std::min<size_t>(std::max<size_t>(4 * size_t(n), 64), 0x80000000u)));
        TRY(dalloc(h, &h->table, h->table_slots));
        HIPC(h, hipMemset(h->table, 0xFF, size_t(h->table_slots) * sizeof(uint4)));   // every slot empty
        TRY(dalloc(h, &h->slot_rec, n));
        HIPC(h, hipMemset(h->slot_rec, 0xFF, size_t(n) * sizeof(uint32_t)));   // no slot recorded
    }
    TRY(dalloc(h, &h->ovf, size_t(n) + 2));
    HIPC(h, hipMemset(h->ovf, 0, (size_t(n) + 2) * sizeof(uint32_t)));
    h->aabb_slots = std::max<uint32_t>({h->nslices, (h->n + 255) / 256, 4096u, uint32_t(m.cloth_offsets.size()), 1u});
    {   // bounds pieces of a batch: each cloth cut into ranges of at most cloth_span_max() particles
        std::vector<uint2> sp;
        std::vector<uint32_t> pf;
        const size_t nc = m.cloth_offsets.size() >= 2 ? m.cloth_offsets.size() - 1 : 0;
        for (size_t c = 0; c < nc; ++c) {
            pf.push_back(uint32_t(sp.size()));
            for (uint32_t b = m.cloth_offsets[c]; b < m.cloth_offsets[c + 1]; b += cloth_span_max())
                sp.push_back(make_uint2(b, std::min(b + cloth_span_max(), m.cloth_offsets[c + 1])));
        }
        pf.push_back(uint32_t(sp.size()));
        h->aabb_slots = std::max<uint32_t>(h->aabb_slots, uint32_t(sp.size()));
        h->ncspans = uint32_t(sp.size());
        if (!sp.empty()) {
            TRY(dalloc(h, &h->cspan, sp.size()));
            TRY(upload(h, h->cspan, sp.data(), sp.size()));
        }
        TRY(dalloc(h, &h->cpfirst, pf.size()));
        TRY(upload(h, h->cpfirst, pf.data(), pf.size()));
    }
    TRY(dalloc(h, &h->aabb, size_t(h->aabb_slots) * 8));
    // GS per-tile flags: device arrays for tiles of GS_TILE_MIN particles or
    // more, and the first tile size whose dependency lists fit
    if (n < (1u << 28)) {
        const uint32_t cap = (n + GS_TILE_MIN - 1) / GS_TILE_MIN;
        h->gt_cap = cap;
        uint32_t *toff = nullptr, *doff = nullptr, *dep = nullptr;
        TRY(dalloc(h, &toff, size_t(h->num_batches) * (cap + 1)));
        TRY(dalloc(h, &doff, size_t(cap) + 1));
        TRY(dalloc(h, &dep, size_t(cap) * GS_TILE_DEPS_MAX));
        TRY(dalloc(h, &h->gt.flag, cap));
        h->gt.toff = toff;
        h->gt.doff = doff;
        h->gt.dep = dep;
        for (uint32_t tile = std::max(GS_TILE_MIN, next_pow2((n + 1023) / 1024)); !h->gt_ok; tile *= 2) {
            TRY(build_gs_tiles(h, tile));
            if (tile >= n) break;
        }
    }
    // GS on LDS grid tiles: one grid cloth of 64k columns; 64-row tiles
    // where the LDS allows, else 32
    if (n < (1u << 28) && h->num_batches && m.cloth_offsets.size() == 2 && m.cloth_cols.size() == 1 &&
        m.cloth_cols[0] >= 64 && m.cloth_cols[0] % 64 == 0 && n % m.cloth_cols[0] == 0) {
        h->gg_cols = m.cloth_cols[0];
        h->gg_rows = n / h->gg_cols;
        h->gg_cap = (h->gg_cols / 64) * ((h->gg_rows + GSG_MIN_ROWS - 1) / GSG_MIN_ROWS);
        uint32_t *coff = nullptr, *con = nullptr;
        float2* crc = nullptr;
        TRY(dalloc(h, &coff, size_t(h->gg_cap) * h->num_batches * 2 + 1));
        TRY(dalloc(h, &con, e));
        TRY(dalloc(h, &crc, e));
        TRY(dalloc(h, &h->gg.flag, h->gg_cap));
        h->gg.coff = coff;
        h->gg.con = con;
        h->gg.crc = crc;
        TRY(build_gs_grid(h, 64));
        if (!h->gg_ok) TRY(build_gs_grid(h, 32));
        h->gg_on = h->gg_ok;
    }
    // Jacobi on LDS-resident grid tiles: 64 rows where the LDS and residency
    // allow, else 32, else 16
    if (h->nbtile && h->tile_cols >= 64) {
        h->jr_cap = (h->tile_cols / 64) * ((h->tile_rows + 15) / 16);
        TRY(dalloc(h, &h->jr.flag, h->jr_cap));
        for (uint32_t th : {64u, 32u, 16u})
            if (!h->jr_ok) TRY(build_jacobi_resident(h, th));
    }
    {
        std::vector<uint32_t> offs = m.cloth_offsets;
        if (offs.size() < 2 || offs.front() != 0 || offs.back() != n) offs = {0u, uint32_t(n)};
        h->nclothes = uint32_t(offs.size() - 1);
        uint32_t bits = 0;
        while ((1u << bits) < h->nclothes) ++bits;
        h->local_bits = h->nclothes <= 1 ? 32u : 32u - bits;
        h->coll_digit_bits = h->nclothes <= 1 ? 9u : 8u;
        TRY(dalloc(h, &h->grids, h->nclothes));
        if (h->nclothes > 1) {
            std::vector<uint32_t> cof(n);
            for (uint32_t c = 0; c < h->nclothes; ++c)
                for (uint32_t i = offs[c]; i < offs[c + 1]; ++i) cof[i] = c;
            TRY(dalloc(h, &h->cloth_of, n));
            TRY(dalloc(h, &h->cloth_off, offs.size()));
            TRY(upload(h, h->cloth_of, cof.data(), n));
            TRY(upload(h, h->cloth_off, offs.data(), offs.size()));
        }
    }

    // vertex -> incident triangles (ascending) for the normals gather
    std::vector<uint32_t> vt_off(n + 1, 0), vt_tri(3 * t);
    for (size_t k = 0; k < 3 * t; ++k) vt_off[m.indices[k] + 1]++;
    for (size_t i = 0; i < n; ++i) vt_off[i + 1] += vt_off[i];
    {
        std::vector<uint32_t> cur(vt_off.begin(), vt_off.end() - 1);
        for (size_t k = 0; k < t; ++k)
            for (int c = 0; c < 3; ++c) vt_tri[cur[m.indices[3 * k + c]]++] = uint32_t(k);
    }
    TRY(dalloc(h, &h->tris, 3 * t));
    TRY(dalloc(h, &h->vt_off, n + 1));
    TRY(dalloc(h, &h->vt_tri, 3 * t));
    TRY(upload(h, h->tris, m.indices.data(), 3 * t));
    TRY(upload(h, h->vt_off, vt_off.data(), n + 1));
    TRY(upload(h, h->vt_tri, vt_tri.data(), 3 * t));

    TRY(dalloc(h, &h->barrier, GRID_SYNC_WORDS));
    TRY(dalloc(h, &h->err, 4));
    TRY(dalloc(h, &h->tameA, h->nslices));
    TRY(dalloc(h, &h->tameB, h->nslices));
    HIPC(h, hipMemset(h->tameA, 0, h->nslices));
    HIPC(h, hipMemset(h->tameB, 0, h->nslices));
    // static ranges of the fast-division proof
    h->max_inv_mass = 0.f;
    for (float w : m.inv_mass) h->max_inv_mass = std::isfinite(w) ? std::max(h->max_inv_mass, w) : INFINITY;
    h->mass_floor = true;   // every inverse mass 0 or >= 2^-26
    for (float w : m.inv_mass)
        if (!(w == 0.f || w >= 1.4901161193847656e-08f)) h->mass_floor = false;
    h->max_comp = 0.f;
    h->rest_tame = true;
    h->comp_nonneg = true;
    for (size_t k = 0; k < m.constraint_comp.size(); ++k) {
        const float c = m.constraint_comp[k], r = m.constraint_rest[k];
        if (!(c >= 0.f)) h->comp_nonneg = false;
        h->max_comp = std::isfinite(c) ? std::max(h->max_comp, std::fabs(c)) : INFINITY;
        if (!(r == 0.f || (r >= 9.313225746154785e-10f && r <= 1073741824.0f))) h->rest_tame = false;   // [2^-30, 2^30]
    }
    TRY(dalloc(h, &h->resid, residual_max_blocks()));
    HIPC(h, hipMemset(h->err, 0, 16));
    HIPC(h, hipMemset(h->barrier, 0, GRID_SYNC_WORDS * sizeof(uint32_t)));
    return XPBD_OK;
}

int upload_state(xpbd_cloth* h, const float* pos4, const float* vel4) {
    std::vector<float> x(4 * size_t(h->n)), v(4 * size_t(h->n), 0.f);
    for (size_t i = 0; i < h->n; ++i) {
        for (int a = 0; a < 3; ++a) {
            x[4 * i + a] = pos4[4 * i + a];
            if (vel4) v[4 * i + a] = vel4[4 * i + a];
        }
        x[4 * i + 3] = h->init_pos4[4 * i + 3];
    }
    HIPC(h, hipStreamSynchronize(h->stream));
    TRY(upload(h, reinterpret_cast<float*>(h->X), x.data(), x.size()));
    TRY(upload(h, reinterpret_cast<float*>(h->V), v.data(), v.size()));
    return XPBD_OK;
}

void apply_tuning(xpbd_cloth* h) {
    if (h->cfg.jacobi_block) h->jacobi_block = h->cfg.jacobi_block;
    if (h->cfg.xcd_remap) h->remap = h->cfg.xcd_remap == 2 ? 1u : 0u;
    // 0: automatic, on where the grid-tiled kernel runs (it stages the
    // predicted positions and proves their range itself); 1: on; 2: off
    h->fused_predict = (h->cfg.fused_predict == 1 || (h->cfg.fused_predict == 0 && h->nbtile)) ? 1u : 0u;
    // persistent Jacobi: every block must be resident at once (grid barrier)
    const uint32_t groups = std::max(1u, (h->nslices + 3) / 4);
    h->pj_blocks = std::max(1u, std::min(persistent_max_blocks(h->nbpk != nullptr, h->cfg.fast_division != 0), groups));
    // coloured GS: 1 barriers, 2 one launch per colour, 3 per-tile flags,
    // 0 the form the autotuner measured fastest (barriers without autotuning)
    h->gs_form = h->cfg.gs_launch == 1   ? GS_FORM_BARRIERS
                 : h->cfg.gs_launch == 2 ? GS_FORM_COLOURS
                 : h->cfg.gs_launch == 3 ? GS_FORM_TILES
                                         : h->gs_form_tuned;
    if (h->gs_form == GS_FORM_TILES && !h->gt_ok && !h->gg_on) h->gs_form = GS_FORM_BARRIERS;
    // grid tiles: 1 = the form the autotuner found faster, 3 = constraint-
    // shared, 2 = per entry
    h->js_on = h->js_ok && (h->cfg.grid_tiles == 3 || (h->cfg.grid_tiles == 1 && h->js_tuned)) ? 1u : 0u;
    // Jacobi substeps on LDS-resident tiles: 2 = on, 0 = where the autotuner
    // measured it faster (per entry tiles only)
    h->jr_on = h->jr_ok && !h->js_on &&
                       (h->cfg.persistent_jacobi == 2 || (h->cfg.persistent_jacobi == 0 && h->jr_tuned))
                   ? 1u
                   : 0u;
}

// Grid neighbours and feature constants of the warm-start network.  Needs a
// grid mesh (cloth_cols): vertex i of cloth c sits at (i - off_c) % cols,
// (i - off_c) / cols.  Neighbour slots: +x, -x, +y, -y, (+x,+y), (-x,-y),
// (+x,-y), (-x,+y).  Per vertex: scaled log10 compliance of its stretch,
// shear and bend constraints (classified by grid offset of the partner) and
// 1 / spacing (rest length of a stretch partner).
float scaled_log_compliance(float c) { return (std::log10(std::max(c, 1e-10f)) + 5.5f) / 2.5f; }

int build_warm_topology(xpbd_cloth* h, const Mesh& m) {
    h->warm_ok = false;
    const size_t C = m.cloth_cols.size();
    if (C == 0 || C + 1 != m.cloth_offsets.size()) return XPBD_OK;
    const size_t n = h->n;
    std::vector<int32_t> nbr(8 * n, -1);
    std::vector<float> aux(4 * n, 0.f);
    static const int DX[8] = {1, -1, 0, 0, 1, -1, 1, -1};
    static const int DY[8] = {0, 0, 1, -1, 1, -1, -1, 1};
    for (size_t c = 0; c < C; ++c) {
        const uint32_t off = m.cloth_offsets[c], end = m.cloth_offsets[c + 1], cols = m.cloth_cols[c];
        if (cols == 0 || (end - off) % cols != 0) return XPBD_OK;
        const int rows = int((end - off) / cols), ic = int(cols);
        for (uint32_t i = off; i < end; ++i) {
            const int col = int((i - off) % cols), row = int((i - off) / cols);
            for (int k = 0; k < 8; ++k) {
                const int nc = col + DX[k], nr = row + DY[k];
                if (nc >= 0 && nc < ic && nr >= 0 && nr < rows) nbr[8 * i + k] = int32_t(off + uint32_t(nr * ic + nc));
            }
            float cs = 0.f, csh = 0.f, cb = 0.f, hs = 0.f;
            bool fs = false, fsh = false, fb = false;
            for (uint32_t q = m.neighbor_offsets[i]; q < m.neighbor_offsets[i + 1]; ++q) {
                const uint32_t j = m.neighbor_other[q];
                if (j < off || j >= end) continue;
                const int dc = std::abs(int((j - off) % cols) - col), dr = std::abs(int((j - off) / cols) - row);
                if (!fs && dc + dr == 1) {
                    fs = true;
                    cs = m.neighbor_compliance[q];
                    hs = m.neighbor_rest_len[q];
                } else if (!fsh && dc == 1 && dr == 1) {
                    fsh = true;
                    csh = m.neighbor_compliance[q];
                } else if (!fb && ((dc == 2 && dr == 0) || (dc == 0 && dr == 2))) {
                    fb = true;
                    cb = m.neighbor_compliance[q];
                }
            }
            aux[4 * i] = scaled_log_compliance(cs);
            aux[4 * i + 1] = scaled_log_compliance(csh);
            aux[4 * i + 2] = scaled_log_compliance(cb);
            aux[4 * i + 3] = hs > 0.f ? 1.0f / hs : 0.f;
        }
    }
    TRY(dalloc(h, &h->nbr8, 8 * n));
    TRY(dalloc(h, &h->vaux, n));
    TRY(upload(h, h->nbr8, nbr.data(), nbr.size()));
    TRY(upload(h, reinterpret_cast<float*>(h->vaux), aux.data(), aux.size()));
    h->warm_ok = true;
    return XPBD_OK;
}

int ensure_warm_buffers(xpbd_cloth* h) {
    if (!h->W0) TRY(dalloc(h, &h->W0, h->n));
    if (!h->tameW) {
        TRY(dalloc(h, &h->tameW, h->nslices));
        TRY(dalloc(h, &h->tameS, h->nslices));
        HIPC(h, hipMemset(h->tameW, 0, h->nslices));
        HIPC(h, hipMemset(h->tameS, 0, h->nslices));
    }
    if (!h->warm_flag) {
        TRY(dalloc(h, &h->warm_flag, 4));
        HIPC(h, hipMemsetAsync(h->warm_flag, 0, 4 * sizeof(uint32_t), h->stream));
    }
    if (!h->warm_acc) {
        TRY(dalloc(h, &h->warm_acc, warm_gate_words()));
        HIPC(h, hipMemsetAsync(h->warm_acc, 0, warm_gate_words() * sizeof(uint32_t), h->stream));
    }
    return XPBD_OK;
}

int autotune(xpbd_cloth* h) {
    if (h->n == 0 || h->e == 0) return XPBD_OK;
    const DevParams p = dev_params(h);
    HIPC(h, hipMemcpyAsync(h->A, h->X, size_t(h->n) * sizeof(float4), hipMemcpyDeviceToDevice, h->stream));
    hipEvent_t e0, e1;
    HIPC(h, hipEventCreate(&e0));
    HIPC(h, hipEventCreate(&e1));
    float best = 1e30f;
    uint32_t best_block = h->jacobi_block, best_remap = h->remap;
    const uint32_t reps = h->n > (1u << 18) ? 10 : 30;
    for (uint32_t block : {64u, 128u, 256u, 512u})
        for (uint32_t remap : {0u, 1u}) {
            JacobiArgs a = jacobi_args(h, p);
            a.remap = remap;
            a.src = h->A;
            a.dst = h->B;
            if (a.fastdiv) {   // the production path: slice flags of both iterates
                launch_tame_scan(h->A, h->n, tame_slot(h, h->A), h->stream);
                a.tame_in = tame_slot(h, h->A);
                a.tame_out = tame_slot(h, h->B);
            }
            launch_jacobi(0, a, block, h->stream);   // warm-up
            (void)hipEventRecord(e0, h->stream);
            for (uint32_t r = 0; r < reps; ++r) {
                a.src = a.dst;
                a.dst = (a.dst == h->A) ? h->B : h->A;
                if (a.fastdiv) {
                    a.tame_in = tame_slot(h, a.src);
                    a.tame_out = tame_slot(h, a.dst);
                }
                launch_jacobi(0, a, block, h->stream);
            }
            (void)hipEventRecord(e1, h->stream);
            (void)hipEventSynchronize(e1);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms < best) {
                best = ms;
                best_block = block;
                best_remap = remap;
            }
        }
    h->jacobi_block = best_block;
    h->remap = best_remap;
    // grid tiles: per-entry or constraint-shared, whichever runs faster here
    h->js_tuned = false;
    if (h->js_ok && h->cfg.grid_tiles == 1) {
        float t_form[2] = {0.f, 0.f};
        for (uint32_t form = 0; form < 2; ++form) {
            JacobiArgs a = jacobi_args(h, p);
            a.sh = form ? h->js : SharedTiles{};
            a.src = h->A;
            a.dst = h->B;
            if (a.fastdiv) {
                launch_tame_scan(h->A, h->n, tame_slot(h, h->A), h->stream);
                a.tame_in = tame_slot(h, h->A);
                a.tame_out = tame_slot(h, h->B);
            }
            launch_jacobi(0, a, h->jacobi_block, h->stream);   // warm-up
            (void)hipEventRecord(e0, h->stream);
            for (uint32_t r = 0; r < reps; ++r) {
                a.src = a.dst;
                a.dst = (a.dst == h->A) ? h->B : h->A;
                if (a.fastdiv) {
                    a.tame_in = tame_slot(h, a.src);
                    a.tame_out = tame_slot(h, a.dst);
                }
                launch_jacobi(0, a, h->jacobi_block, h->stream);
            }
            (void)hipEventRecord(e1, h->stream);
            (void)hipEventSynchronize(e1);
            (void)hipEventElapsedTime(&t_form[form], e0, e1);
        }
        h->js_tuned = t_form[1] < t_form[0];
    }
    // Jacobi substeps: per-iteration launches (predict fused into the first,
    // bounds into the last) against one launch on LDS-resident tiles, each
    // timed as a replayed graph of a collision substep's Jacobi part (the
    // state is only read: the iterates go to A/B).
    h->jr_tuned = false;
    const bool js_would = h->js_ok && (h->cfg.grid_tiles == 3 || (h->cfg.grid_tiles == 1 && h->js_tuned));
    if (h->jr_ok && h->cfg.persistent_jacobi == 0 && h->cfg.solver == XPBD_SOLVER_JACOBI && !js_would) {
        const uint32_t iters = std::max(h->cfg.solve_inner_iterations, 1u);
        auto time_graph = [&](auto body) -> float {
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            float ms = 1e30f;
            if (hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return ms;
            body();
            if (hipStreamEndCapture(h->stream, &graph) != hipSuccess || !graph) return ms;
            if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess) {
                (void)hipGraphLaunch(exec, h->stream);   // warm-up
                (void)hipEventRecord(e0, h->stream);
                for (uint32_t r = 0; r < 5; ++r) (void)hipGraphLaunch(exec, h->stream);
                (void)hipEventRecord(e1, h->stream);
                (void)hipEventSynchronize(e1);
                (void)hipEventElapsedTime(&ms, e0, e1);
                (void)hipGraphExecDestroy(exec);
            }
            (void)hipGraphDestroy(graph);
            return ms;
        };
        const float t_iter = time_graph([&]() {
            JacobiArgs a = jacobi_args(h, p);
            a.sh = SharedTiles{};
            a.src = h->X;
            a.vel = h->V;
            a.dst = h->A;
            if (iters == 1) {
                launch_jacobi(JM_PREDICT | JM_AABB, a, h->jacobi_block, h->stream);
                return;
            }
            launch_jacobi(JM_PREDICT, a, h->jacobi_block, h->stream);
            float4* cur = h->A;
            float4* nxt = h->B;
            for (uint32_t k = 1; k < iters; ++k) {
                a.src = cur;
                a.dst = nxt;
                launch_jacobi(k + 1 < iters ? 0u : JM_AABB, a, h->jacobi_block, h->stream);
                std::swap(cur, nxt);
            }
        });
        const float t_res = time_graph([&]() {
            JacobiArgs a = jacobi_args(h, p);
            PersistArgs q{};
            q.X = h->X;
            q.V = h->V;
            q.A = h->A;
            q.B = h->B;
            q.iterations = iters;
            q.final_write = 0;
            q.barrier = h->barrier;
            q.error = h->err;
            launch_jacobi_resident(a, q, h->jr, h->stream);
        });
        uint32_t flag = 0;
        HIPC(h, hipMemcpy(&flag, h->err, sizeof(flag), hipMemcpyDeviceToHost));
        if (flag) {   // a wait expired: these tiles are not resident at once here
            const uint32_t zero = 0;
            HIPC(h, hipMemcpy(h->err, &zero, sizeof(zero), hipMemcpyHostToDevice));
            h->jr_ok = false;
        } else {
            h->jr_tuned = t_res < t_iter;
        }
    }
    // Coloured GS: the grid size of the persistent substep.  Every block takes
    // part in each of the I x colours grid barriers, whose cost grows with the
    // grid, while fewer blocks leave more constraints per thread per colour;
    // the best point is measured (same results for every grid: each colour's
    // constraints are vertex-disjoint).
    if (h->cfg.solver == XPBD_SOLVER_GAUSS_SEIDEL && h->e && h->gs_blocks > 64) {
        GsArgs g{};
        g.X = h->X;
        g.V = h->V;
        g.W = h->A;
        g.xout = h->B;
        g.vout = h->V2;
        g.ij = h->gs_ij;
        g.rc = h->gs_rc;
        g.batch_offsets = h->gs_off;
        g.num_batches = h->num_batches;
        g.iterations = std::max(h->cfg.solve_inner_iterations, 1u);
        g.barrier = h->barrier;
        g.error = h->err;
        g.aabb = h->aabb;
        g.final_write = 1;
        g.p = p;
        float gbest = 1e30f;
        uint32_t gbest_blocks = h->gs_blocks;
        for (uint32_t b = h->gs_blocks; b >= 64; b /= 2) {
            launch_zero_u32(h->barrier, GRID_SYNC_WORDS, h->stream);
            launch_gs(g, b, h->gs_block, h->stream);   // warm-up
            (void)hipEventRecord(e0, h->stream);
            for (uint32_t r = 0; r < 3; ++r) {
                launch_zero_u32(h->barrier, GRID_SYNC_WORDS, h->stream);
                launch_gs(g, b, h->gs_block, h->stream);
            }
            (void)hipEventRecord(e1, h->stream);
            (void)hipEventSynchronize(e1);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms < gbest) {
                gbest = ms;
                gbest_blocks = b;
            }
        }
        h->gs_blocks = gbest_blocks;
    }
    // Coloured GS: one persistent launch (grid barriers between colours)
    // against one launch per colour (launch boundaries instead) and one
    // persistent launch ordered by per-tile flags (at each tile size from the
    // create-time one to 8 times it), each timed as a replayed graph of one
    // substep's GS part, the form frames replay.  Barriers cost more than
    // graph kernel boundaries on large grids, while small colours leave
    // launches mostly latency.
    h->gs_form_tuned = GS_FORM_BARRIERS;
    if (h->cfg.solver == XPBD_SOLVER_GAUSS_SEIDEL && h->e) {
        GsArgs g{};
        g.X = h->X;
        g.V = h->V;
        g.W = h->A;
        g.xout = h->B;
        g.vout = h->V2;
        g.ij = h->gs_ij;
        g.rc = h->gs_rc;
        g.batch_offsets = h->gs_off;
        g.num_batches = h->num_batches;
        g.iterations = std::max(h->cfg.solve_inner_iterations, 1u);
        g.barrier = h->barrier;
        g.error = h->err;
        g.aabb = h->aabb;
        g.final_write = 1;
        g.p = p;
        auto time_graph = [&](auto body) -> float {
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            float ms = 1e30f;
            if (hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return ms;
            body();
            if (hipStreamEndCapture(h->stream, &graph) != hipSuccess || !graph) return ms;
            if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess) {
                (void)hipGraphLaunch(exec, h->stream);   // warm-up
                (void)hipEventRecord(e0, h->stream);
                for (uint32_t r = 0; r < 3; ++r) (void)hipGraphLaunch(exec, h->stream);
                (void)hipEventRecord(e1, h->stream);
                (void)hipEventSynchronize(e1);
                (void)hipEventElapsedTime(&ms, e0, e1);
                (void)hipGraphExecDestroy(exec);
            }
            (void)hipGraphDestroy(graph);
            return ms;
        };
        const float tp = time_graph([&]() {
            launch_zero_u32(h->barrier, GRID_SYNC_WORDS, h->stream);
            launch_gs(g, h->gs_blocks, h->gs_block, h->stream);
        });
        const float tc = time_graph([&]() {
            uint32_t np = 0;
            enqueue_gs_colours(h, g, false, np, h->stream);
        });
        float tt = 1e30f;
        if (h->gt_ok) {
            const uint32_t base = h->gt.tile;
            uint32_t best_tile = base;
            for (uint32_t tile = base; tile <= 8 * base && tile < 2 * h->n; tile *= 2) {
                TRY(build_gs_tiles(h, tile));
                if (!h->gt_ok) continue;
                const float t = time_graph([&]() { launch_gs_tiles(g, h->gt, h->gt_blocks, h->stream); });
                uint32_t flag = 0;
                HIPC(h, hipMemcpy(&flag, h->err, sizeof(flag), hipMemcpyDeviceToHost));
                if (flag) {   // a wait expired: this grid is not resident at once here
                    const uint32_t zero = 0;
                    HIPC(h, hipMemcpy(h->err, &zero, sizeof(zero), hipMemcpyHostToDevice));
                    continue;
                }
                if (t < tt) {
                    tt = t;
                    best_tile = tile;
                }
            }
            TRY(build_gs_tiles(h, best_tile));
            if (tt >= 1e30f) h->gt_ok = false;
        }
        // the flag form on LDS grid tiles, at 64, 32 and 16 rows per tile
        float tg = 1e30f;
        if (h->gg_cap) {
            uint32_t best_th = 0;
            for (uint32_t th : {64u, 32u, 16u}) {
                TRY(build_gs_grid(h, th));
                if (!h->gg_ok) continue;
                const float t = time_graph([&]() { launch_gs_grid(g, h->gg, h->stream); });
                uint32_t flag = 0;
                HIPC(h, hipMemcpy(&flag, h->err, sizeof(flag), hipMemcpyDeviceToHost));
                if (flag) {   // a wait expired: these tiles are not resident at once here
                    const uint32_t zero = 0;
                    HIPC(h, hipMemcpy(h->err, &zero, sizeof(zero), hipMemcpyHostToDevice));
                    continue;
                }
                if (t < tg) {
                    tg = t;
                    best_th = th;
                }
            }
            if (best_th) TRY(build_gs_grid(h, best_th));
            else h->gg_ok = false;
            h->gg_on = h->gg_ok && (tg < tt || !h->gt_ok);
            tt = std::min(tt, tg);
        }
        h->gs_form_tuned = tc < tp && tc <= tt ? GS_FORM_COLOURS : tt < tp ? GS_FORM_TILES : GS_FORM_BARRIERS;
    }
    apply_tuning(h);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPC(h, hipGetLastError());
    return check_device_flag(h);
}

// Co-residency of the persistent grids, checked on this device.  A grid
// barrier needs every block of the launch resident at once; the starting grid
// is the occupancy figure capped at 7 of 8 wavefront slots per SIMD
// (coresident_per_cu).  Each grid the handle will launch is run once in probe
// mode -- the kernel itself, so the same registers and LDS, doing its barrier
// set-up and one barrier -- and where the bounded wait expires (a block did
// not become resident in time) the grid shrinks by one block per CU and the
// probe repeats.  Work of other processes on the same GPU can still take
// slots at run time; the bounded wait stays the guard there.
int verify_coresident(xpbd_cloth* h, bool gs, bool persistent) {
    TRY(check_device_flag(h));
    int dev = 0, cus = 0;
    HIPC(h, hipGetDevice(&dev));
    HIPC(h, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const DevParams p = dev_params(h);
    hipStream_t s = h->stream;
    auto check = [&](auto launch, uint32_t& blocks, const char* what) -> int {
        for (int tries = 0; tries < 8 && blocks > 0; ++tries) {
            launch_zero_u32(h->barrier, GRID_SYNC_WORDS, s);
            launch(blocks);
            HIPC(h, hipGetLastError());
            uint32_t flag = 0;
            HIPC(h, hipMemcpyAsync(&flag, h->err, sizeof(flag), hipMemcpyDeviceToHost, s));
            HIPC(h, hipStreamSynchronize(s));
            if (!flag) return XPBD_OK;
            const uint32_t zero = 0;
            HIPC(h, hipMemcpy(h->err, &zero, sizeof(zero), hipMemcpyHostToDevice));
            blocks = blocks > uint32_t(cus) ? blocks - uint32_t(cus) : blocks / 2;
        }
        return fail(h, XPBD_E_STATE, std::string("no co-resident grid found for ") + what);
    };
    if (gs && h->e) {
        TRY(check(
            [&](uint32_t b) {
                GsArgs g{};
                g.barrier = h->barrier;
                g.error = h->err;
                g.probe = 1;
                g.p = p;
                launch_gs(g, b, h->gs_block, s);
            },
            h->gs_blocks, "the coloured Gauss-Seidel kernel"));
    }
    if (gs && h->gs_form == GS_FORM_TILES && h->gg_on) {
        // every grid tile needs its own resident block: probed once; where
        // the probe's wait expires the index-range tiles run instead
        launch_zero_u32(h->barrier, GRID_SYNC_WORDS, s);
        GsArgs g{};
        g.barrier = h->barrier;
        g.error = h->err;
        g.probe = 1;
        g.p = p;
        launch_gs_grid(g, h->gg, s);
        HIPC(h, hipGetLastError());
        uint32_t flag = 0;
        HIPC(h, hipMemcpyAsync(&flag, h->err, sizeof(flag), hipMemcpyDeviceToHost, s));
        HIPC(h, hipStreamSynchronize(s));
        if (flag) {
            const uint32_t zero = 0;
            HIPC(h, hipMemcpy(h->err, &zero, sizeof(zero), hipMemcpyHostToDevice));
            h->gg_on = false;
            h->gg_ok = false;
            drop_graph(h);
        }
    }
    if (gs && h->gs_form == GS_FORM_TILES && !h->gg_on) {
        TRY(check(
            [&](uint32_t b) {
                GsArgs g{};
                g.barrier = h->barrier;
                g.error = h->err;
                g.probe = 1;
                g.p = p;
                launch_gs_tiles(g, h->gt, b, s);
            },
            h->gt_blocks, "the per-tile Gauss-Seidel kernel"));
    }
    if (h->cfg.solver == XPBD_SOLVER_JACOBI && h->jr_on) {
        // every resident tile needs its own resident block: probed once; where
        // the probe's wait expires the per-iteration launches run instead
        launch_zero_u32(h->barrier, GRID_SYNC_WORDS, s);
        JacobiArgs a = jacobi_args(h, p);
        PersistArgs q{};
        q.barrier = h->barrier;
        q.error = h->err;
        q.probe = 1;
        launch_jacobi_resident(a, q, h->jr, s);
        HIPC(h, hipGetLastError());
        uint32_t flag = 0;
        HIPC(h, hipMemcpyAsync(&flag, h->err, sizeof(flag), hipMemcpyDeviceToHost, s));
        HIPC(h, hipStreamSynchronize(s));
        if (flag) {
            const uint32_t zero = 0;
            HIPC(h, hipMemcpy(h->err, &zero, sizeof(zero), hipMemcpyHostToDevice));
            h->jr_on = 0;
            h->jr_ok = false;
            drop_graph(h);
        }
    }
    if (persistent) {
        TRY(check(
            [&](uint32_t b) {
                JacobiArgs a = jacobi_args(h, p);
                PersistArgs q{};
                q.barrier = h->barrier;
                q.error = h->err;
                q.probe = 1;
                launch_jacobi_persistent(a, q, b, s);
            },
            h->pj_blocks, "the persistent Jacobi kernel"));
    }
    return XPBD_OK;
}

bool uses_persistent_jacobi(const xpbd_cloth* h) {
    return h->cfg.solver == XPBD_SOLVER_JACOBI &&
           (h->cfg.persistent_jacobi == 1 || (h->warm && std::isfinite(h->warm_threshold)));
}

}  // namespace

// Reason for the last failed call without a handle on this thread: a
// static string, or the text kept in g_thread_text.
static thread_local const char* g_thread_error = "";
static thread_local std::string g_thread_text;

extern "C" {

XPBD_API int xpbd_abi_version(void) { return XPBD_ABI_VERSION; }

// ---------------------------------------------------------------- meshes
static int wrap_mesh(Mesh* m, xpbd_mesh** out) {
    if (!m) return XPBD_E_INVALID;
    auto* w = new (std::nothrow) xpbd_mesh();
    if (!w) {
        delete m;
        return XPBD_E_NOMEM;
    }
    w->m = std::move(*m);
    delete m;
    *out = w;
    return XPBD_OK;
}

XPBD_API int xpbd_mesh_grid_cloth_hanging(uint32_t qc, uint32_t qr, float cell, xpbd_mesh** out) {
    if (!out || qc < 1 || qr < 1 || !(cell > 0.f) || !std::isfinite(cell)) return XPBD_E_INVALID;
    try {
        return wrap_mesh(grid_cloth_hanging(qc, qr, cell), out);
    } catch (const std::bad_alloc&) {
        return XPBD_E_NOMEM;
    }
}

XPBD_API int xpbd_mesh_parse_welded_obj(const char* text, size_t len, xpbd_mesh** out) {
    if (!out || (!text && len)) return XPBD_E_INVALID;
    try {
        return wrap_mesh(parse_welded_obj(std::string(text ? text : "", len)), out);
    } catch (const std::bad_alloc&) {
        return XPBD_E_NOMEM;
    }
}

XPBD_API int xpbd_mesh_sheet(const xpbd_sheet_desc* d, xpbd_mesh** out) { return xpbd_mesh_batch(d, 1, out); }


XPBD_API const char* xpbd_thread_last_error(void) { return g_thread_error; }

XPBD_API int xpbd_mesh_from_arrays(const xpbd_mesh_desc* d, xpbd_mesh** out) {
    g_thread_error = "";
    if (!d || !out) {
        g_thread_error = "null argument";
        return XPBD_E_INVALID;
    }
    try {
        std::string why;
        int code = XPBD_E_INVALID;
        Mesh* m = mesh_from_arrays(*d, why, code);
        if (!m) {
            g_thread_text = why;
            g_thread_error = g_thread_text.c_str();
            return code;
        }
        return wrap_mesh(m, out);
    } catch (const std::bad_alloc&) {
        g_thread_error = "out of host memory";
        return XPBD_E_NOMEM;
    }
}

XPBD_API int xpbd_mesh_batch(const xpbd_sheet_desc* d, uint32_t count, xpbd_mesh** out) {
    if (!out || !d || count == 0) return XPBD_E_INVALID;
    uint64_t total = 0;
    for (uint32_t c = 0; c < count; ++c) {
        if (d[c].cols < 2 || d[c].rows < 2 || !(d[c].spacing > 0.f) || d[c].layout > 1 || d[c].pins > 3)
            return XPBD_E_INVALID;
        total += uint64_t(d[c].cols) * d[c].rows;
    }
    if (total >= 0x40000000ull) return XPBD_E_INVALID;
    try {
        return wrap_mesh(sheet_batch(d, count), out);
    } catch (const std::bad_alloc&) {
        return XPBD_E_NOMEM;
    }
}

static const void* mesh_field(const Mesh& m, int f, uint64_t& count) {
    switch (f) {
        case XPBD_MESH_POSITIONS: count = m.positions.size(); return m.positions.data();
        case XPBD_MESH_REST_POSITIONS: count = m.rest_positions.size(); return m.rest_positions.data();
        case XPBD_MESH_NORMALS: count = m.normals.size(); return m.normals.data();
        case XPBD_MESH_UVS: count = m.uvs.size(); return m.uvs.data();
        case XPBD_MESH_INDICES: count = m.indices.size(); return m.indices.data();
        case XPBD_MESH_TRI_EDGE_NEIGHBORS: count = m.tri_edge_neighbors.size(); return m.tri_edge_neighbors.data();
        case XPBD_MESH_INV_MASS: count = m.inv_mass.size(); return m.inv_mass.data();
        case XPBD_MESH_NEIGHBOR_OTHER: count = m.neighbor_other.size(); return m.neighbor_other.data();
        case XPBD_MESH_NEIGHBOR_REST_LEN: count = m.neighbor_rest_len.size(); return m.neighbor_rest_len.data();
        case XPBD_MESH_NEIGHBOR_COMPLIANCE: count = m.neighbor_compliance.size(); return m.neighbor_compliance.data();
        case XPBD_MESH_NEIGHBOR_CONSTRAINT_ID: count = m.neighbor_constraint_id.size(); return m.neighbor_constraint_id.data();
        case XPBD_MESH_NEIGHBOR_OFFSETS: count = m.neighbor_offsets.size(); return m.neighbor_offsets.data();
        case XPBD_MESH_CONSTRAINT_I: count = m.constraint_i.size(); return m.constraint_i.data();
        case XPBD_MESH_CONSTRAINT_J: count = m.constraint_j.size(); return m.constraint_j.data();
        case XPBD_MESH_CONSTRAINT_REST: count = m.constraint_rest.size(); return m.constraint_rest.data();
        case XPBD_MESH_CONSTRAINT_COMPLIANCE: count = m.constraint_comp.size(); return m.constraint_comp.data();
        case XPBD_MESH_BATCH_OFFSETS: count = m.batch_offsets.size(); return m.batch_offsets.data();
        case XPBD_MESH_GROUND: count = m.ground.size(); return m.ground.data();
        case XPBD_MESH_CLOTH_OFFSETS: count = m.cloth_offsets.size(); return m.cloth_offsets.data();
        case XPBD_MESH_CLOTH_COLS: count = m.cloth_cols.size(); return m.cloth_cols.data();
        default: count = 0; return nullptr;
    }
}

XPBD_API uint64_t xpbd_mesh_count(const xpbd_mesh* mesh, int field) {
    uint64_t c = 0;
    if (mesh) mesh_field(mesh->m, field, c);
    return c;
}

XPBD_API int xpbd_mesh_read(const xpbd_mesh* mesh, int field, void* dst, uint64_t bytes) {
    if (!mesh || !dst) return XPBD_E_INVALID;
    uint64_t c = 0;
    const void* p = mesh_field(mesh->m, field, c);
    if (!p && c) return XPBD_E_INVALID;
    if (field < 0 || field >= XPBD_MESH_NUM_FIELDS) return XPBD_E_INVALID;
    if (bytes < c * 4) return XPBD_E_INVALID;
    if (c) std::memcpy(dst, p, c * 4);
    return XPBD_OK;
}

XPBD_API void xpbd_mesh_destroy(xpbd_mesh* mesh) { delete mesh; }

XPBD_API int xpbd_derive_collision_grid(const float* rest_vec4, uint32_t n, float thickness, xpbd_coll_grid* out) {
    if (!out || (!rest_vec4 && n)) return XPBD_E_INVALID;
    derive_collision_grid(rest_vec4, n, thickness, out);
    return XPBD_OK;
}

// ---------------------------------------------------------------- config
XPBD_API void xpbd_sim_uniforms_default(uint32_t solver, uint32_t substeps, xpbd_sim_uniforms* out) {
    if (!out) return;
    const uint32_t s = std::max(substeps, 1u);
    xpbd_sim_uniforms u{};
    u.dt = REFERENCE_FRAME_DELTA_SECS / float(s);
    u.inv_dt = 1.0f / u.dt;
    u.jacobi_omega = solver == XPBD_SOLVER_GAUSS_SEIDEL ? 1.0f : 0.32f;   // cloth_jacobi.rs:946-948
    u.inner_iterations = INNER_ITERS;
    u.thickness = THICKNESS;
    u.coll_scale = DEFAULT_COLL_SCALE;
    u.gravity[1] = -9.81f;
    u.grab_idx = -1;
    u.grab_active = 0;
    u.grab_stiffness = solver == XPBD_SOLVER_GAUSS_SEIDEL ? 0.45f : 0.28f;
    u.linear_drag_per_sec = DEFAULT_LINEAR_AIR_DRAG_PER_SEC;
    *out = u;
}

XPBD_API void xpbd_sim_config_default(xpbd_sim_config* out) {
    if (!out) return;
    xpbd_sim_config c{};
    c.solve_substeps = SUBSTEPS;
    c.solve_inner_iterations = INNER_ITERS;
    c.collision_every_n_substeps = 1;
    c.solver = XPBD_SOLVER_JACOBI;
    c.collision_cell = 0.f;
    c.use_graph = 1;
    c.autotune = 1;
    c.jacobi_block = 0;
    c.fused_predict = 0;
    c.fast_division = 1;
    c.grid_tiles = 1;
    c.gs_launch = 0;
    *out = c;
}

// ---------------------------------------------------------------- cloth
XPBD_API int xpbd_create(const xpbd_mesh* mesh, const xpbd_sim_config* cfg, const xpbd_sim_uniforms* u,
                         void* stream, xpbd_cloth** out) {
    if (!mesh || !cfg || !u || !out) return XPBD_E_INVALID;
    *out = nullptr;
    if (cfg->solver > XPBD_SOLVER_GAUSS_SEIDEL || cfg->gs_launch > 3 || cfg->persistent_jacobi > 2)
        return XPBD_E_INVALID;
    if (cfg->jacobi_block && (cfg->jacobi_block % 64 || cfg->jacobi_block > 512)) return XPBD_E_INVALID;
    std::string why;
    if (validate_mesh(mesh->m, why) != XPBD_OK) return XPBD_E_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return XPBD_E_NO_DEVICE;
    auto* h = new (std::nothrow) xpbd_cloth();
    if (!h) return XPBD_E_NOMEM;
    std::unique_ptr<xpbd_cloth, void (*)(xpbd_cloth*)> guard(h, xpbd_destroy);
    h->cfg = *cfg;
    h->u = *u;
    const Mesh& m = mesh->m;
    h->n = uint32_t(m.num_particles());
    h->e = uint32_t(m.num_constraints());
    h->t = uint32_t(m.num_tris());
    h->nslices = (h->n + 63) / 64;
    h->u.num_particles = h->n;
    h->u.num_tris = h->t;
    if (hipGetDevice(&h->device) != hipSuccess) return XPBD_E_HIP;
    if (stream) {
        h->stream = static_cast<hipStream_t>(stream);
    } else {
        if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return XPBD_E_HIP;
        h->own_stream = true;
    }
    int r = create_device(h, m);
    if (r != XPBD_OK) return r;
    if (cfg->gs_launch == 3 && !h->gt_ok && !h->gg_on)
        return fail(h, XPBD_E_INVALID, "gs_launch 3: no per-tile flag plan for this mesh");
    if (cfg->persistent_jacobi == 2 && !h->jr_ok)
        return fail(h, XPBD_E_INVALID, "persistent_jacobi 2: no LDS-resident tiles for this mesh");
    r = build_warm_topology(h, m);
    if (r != XPBD_OK) return r;
    r = upload_state(h, h->init_pos4.data(), nullptr);
    if (r != XPBD_OK) return r;
    // GS grid: enough threads for the largest colour (one constraint each) and
    // a quarter of the particles, within what can be resident at once (every
    // block takes part in each grid barrier, whose cost grows with the grid).
    uint32_t widest = 0;
    for (size_t b = 0; b + 1 < h->batch_offsets.size(); ++b)
        widest = std::max(widest, h->batch_offsets[b + 1] - h->batch_offsets[b]);
    const uint32_t gs_work = std::max(widest, (h->n + 3) / 4);
    h->gs_blocks = std::min<uint32_t>(gs_max_blocks(h->gs_block), std::max(1u, (gs_work + h->gs_block - 1) / h->gs_block));
    h->gs_blocks = std::max<uint32_t>(h->gs_blocks, 1u);
    apply_tuning(h);
    if (cfg->autotune) {
        r = autotune(h);
        if (r != XPBD_OK) return r;
    }
    r = verify_coresident(h, h->cfg.solver == XPBD_SOLVER_GAUSS_SEIDEL, uses_persistent_jacobi(h));
    if (r != XPBD_OK) return r;
    if (hipStreamSynchronize(h->stream) != hipSuccess) return XPBD_E_HIP;
    release_gs_host_rows(h);
    *out = guard.release();
    return XPBD_OK;
}

XPBD_API void xpbd_destroy(xpbd_cloth* h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    drop_graph(h);
    release_imports(h);
    for (void* p : h->allocs) (void)hipFree(p);
    if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

XPBD_API const char* xpbd_last_error(const xpbd_cloth* h) { return h ? h->msg.c_str() : "null handle"; }

XPBD_API int xpbd_set_uniforms(xpbd_cloth* h, const xpbd_sim_uniforms* u) {
    if (!h || !u) return XPBD_E_INVALID;
    h->u = *u;
    h->u.num_particles = h->n;
    h->u.num_tris = h->t;
    drop_graph(h);
    return XPBD_OK;
}

XPBD_API int xpbd_set_config(xpbd_cloth* h, const xpbd_sim_config* cfg) {
    if (!h || !cfg || cfg->solver > XPBD_SOLVER_GAUSS_SEIDEL || cfg->gs_launch > 3 || cfg->persistent_jacobi > 2)
        return XPBD_E_INVALID;
    if (cfg->jacobi_block && (cfg->jacobi_block % 64 || cfg->jacobi_block > 512)) return XPBD_E_INVALID;
    const bool retune = cfg->autotune && !h->cfg.autotune;
    if (cfg->gs_launch == 3 && !h->gt_ok && !h->gg_on)
        return fail(h, XPBD_E_INVALID, "gs_launch 3: no per-tile flag plan for this mesh");
    if (cfg->persistent_jacobi == 2 && !h->jr_ok)
        return fail(h, XPBD_E_INVALID, "persistent_jacobi 2: no LDS-resident tiles for this mesh");
    h->cfg = *cfg;
    apply_tuning(h);
    drop_graph(h);
    if (retune) TRY(autotune(h));
    release_gs_host_rows(h);
    return verify_coresident(h, h->cfg.solver == XPBD_SOLVER_GAUSS_SEIDEL, uses_persistent_jacobi(h));
}

XPBD_API int xpbd_step_frames(xpbd_cloth* h, uint32_t frames) {
    if (!h) return XPBD_E_INVALID;
    if (h->n == 0 || frames == 0) return XPBD_OK;
    if (h->cfg.use_graph) {
        if (!h->graph_valid) {
            drop_graph(h);
            HIPC(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            enqueue_frame(h);
            hipGraph_t g = nullptr;
            hipError_t ce = hipStreamEndCapture(h->stream, &g);
            if (ce != hipSuccess) return fail(h, XPBD_E_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
            h->graph = g;
            HIPC(h, hipGraphInstantiate(&h->gexec, h->graph, nullptr, nullptr, 0));
            h->graph_valid = true;
        }
        for (uint32_t f = 0; f < frames; ++f) HIPC(h, hipGraphLaunch(h->gexec, h->stream));
    } else {
        for (uint32_t f = 0; f < frames; ++f) enqueue_frame(h);
    }
    HIPC(h, hipGetLastError());
    return XPBD_OK;
}

XPBD_API int xpbd_step_substeps(xpbd_cloth* h, uint32_t count, uint32_t collision) {
    if (!h) return XPBD_E_INVALID;
    if (h->n == 0) return XPBD_OK;
    const bool coll = collision != 0 && h->u.coll_scale > 1e-12f;
    for (uint32_t k = 0; k < count; ++k) enqueue_substep(h, coll);
    HIPC(h, hipGetLastError());
    return XPBD_OK;
}

XPBD_API int xpbd_sync(xpbd_cloth* h) {
    if (!h) return XPBD_E_INVALID;
    HIPC(h, hipStreamSynchronize(h->stream));
    return check_device_flag(h);
}

XPBD_API void* xpbd_stream(xpbd_cloth* h) { return h ? static_cast<void*>(h->stream) : nullptr; }

XPBD_API int xpbd_reset(xpbd_cloth* h) {
    if (!h) return XPBD_E_INVALID;
    return upload_state(h, h->init_pos4.data(), nullptr);
}

XPBD_API int xpbd_read_state(xpbd_cloth* h, float* pos4, float* vel4) {
    if (!h) return XPBD_E_INVALID;
    HIPC(h, hipStreamSynchronize(h->stream));
    TRY(check_device_flag(h));
    const size_t bytes = size_t(h->n) * sizeof(float4);
    if (pos4) {
        HIPC(h, hipMemcpy(pos4, h->X, bytes, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < h->n; ++i) pos4[4 * i + 3] = 0.f;
    }
    if (vel4) {
        HIPC(h, hipMemcpy(vel4, h->V, bytes, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < h->n; ++i) vel4[4 * i + 3] = 0.f;
    }
    return XPBD_OK;
}

XPBD_API int xpbd_write_state(xpbd_cloth* h, const float* pos4, const float* vel4) {
    if (!h || !pos4) return XPBD_E_INVALID;
    return upload_state(h, pos4, vel4);
}

XPBD_API int xpbd_publish(xpbd_cloth* h, void* dpos, void* dnrm) {
    if (!h) return XPBD_E_INVALID;
    launch_publish(h->X, h->n, h->tris, h->vt_off, h->vt_tri, static_cast<float4*>(dpos), static_cast<float4*>(dnrm),
                   h->stream);
    HIPC(h, hipGetLastError());
    return XPBD_OK;
}

XPBD_API int xpbd_bind_render_buffers(xpbd_cloth* h, void* dpos, void* dnrm) {
    if (!h) return XPBD_E_INVALID;
    if (h->stream) HIPC(h, hipStreamSynchronize(h->stream));
    release_imports(h);
    h->render_pos = static_cast<float4*>(dpos);
    h->render_nrm = static_cast<float4*>(dnrm);
    drop_graph(h);
    return XPBD_OK;
}

XPBD_API int xpbd_import_render_buffers_fd(xpbd_cloth* h, int pos_fd, uint64_t pos_size, uint64_t pos_offset,
                                          int nrm_fd, uint64_t nrm_size, uint64_t nrm_offset) {
    if (!h) return XPBD_E_INVALID;
    const uint64_t need = uint64_t(h->n) * sizeof(float4);
    const int fds[2] = {pos_fd, nrm_fd};
    const uint64_t sizes[2] = {pos_size, nrm_size}, offs[2] = {pos_offset, nrm_offset};
    for (int k = 0; k < 2; ++k)
        if (fds[k] >= 0 && (offs[k] > sizes[k] || sizes[k] - offs[k] < need || offs[k] % 16))
            return fail(h, XPBD_E_INVALID, "render buffer smaller than 16 N bytes past its offset, or misaligned");
    if (pos_fd < 0 && nrm_fd < 0) return XPBD_E_INVALID;
    if (h->stream) HIPC(h, hipStreamSynchronize(h->stream));
    release_imports(h);
    h->render_pos = h->render_nrm = nullptr;
    drop_graph(h);
    for (int k = 0; k < 2; ++k) {
        if (fds[k] < 0) continue;
        hipExternalMemoryHandleDesc d{};
        d.type = hipExternalMemoryHandleTypeOpaqueFd;
        d.handle.fd = fds[k];
        d.size = sizes[k];
        hipError_t e = hipImportExternalMemory(&h->ext_mem[k], &d);
        if (e == hipSuccess) {
            hipExternalMemoryBufferDesc b{};
            b.offset = offs[k];
            b.size = need;
            e = hipExternalMemoryGetMappedBuffer(&h->ext_ptr[k], h->ext_mem[k], &b);
        }
        if (e != hipSuccess) {
            release_imports(h);
            return fail(h, XPBD_E_HIP, std::string("importing a render buffer: ") + hipGetErrorString(e));
        }
    }
    h->render_pos = static_cast<float4*>(h->ext_ptr[0]);
    h->render_nrm = static_cast<float4*>(h->ext_ptr[1]);
    return XPBD_OK;
}

XPBD_API int xpbd_residual(xpbd_cloth* h, double* out) {
    if (!h || !out) return XPBD_E_INVALID;
    const uint32_t blocks = launch_residual(h->X, h->near_ij, h->near_rest, h->e, h->resid, h->stream);
    std::vector<double> part(blocks);
    if (blocks)
        HIPC(h, hipMemcpyAsync(part.data(), h->resid, blocks * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPC(h, hipStreamSynchronize(h->stream));
    double c2 = 0.0;
    for (double v : part) c2 += v;
    *out = h->rest_sq_sum > 0.0 ? std::sqrt(c2) / std::sqrt(h->rest_sq_sum) : 0.0;
    return XPBD_OK;
}

XPBD_API int xpbd_residual_noalloc(const void* d_pos, uint32_t n, const void* d_ci, const void* d_cj,
                                   const void* d_rest, uint32_t e, void* stream, void* d_out) {
    g_thread_error = "";
    if (!d_out || (e && (!d_pos || !d_ci || !d_cj || !d_rest))) {
        g_thread_error = "null device pointer";
        return XPBD_E_INVALID;
    }
    if (e > XPBD_RESIDUAL_NOALLOC_MAX_CONSTRAINTS) {
        g_thread_error = "num_constraints above XPBD_RESIDUAL_NOALLOC_MAX_CONSTRAINTS (2^31)";
        return XPBD_E_INVALID;
    }
    const hipError_t err = launch_residual_noalloc(static_cast<const float4*>(d_pos), n, static_cast<const uint32_t*>(d_ci),
                                                   static_cast<const uint32_t*>(d_cj), static_cast<const float*>(d_rest),
                                                   e, static_cast<hipStream_t>(stream), static_cast<double*>(d_out));
    if (err != hipSuccess) {
        g_thread_error = hipGetErrorString(err);
        return XPBD_E_HIP;
    }
    return XPBD_OK;
}

XPBD_API int xpbd_time_kernel(xpbd_cloth* h, int which, uint32_t reps, float* avg_ms) {
    if (!h || !avg_ms || reps == 0) return XPBD_E_INVALID;
    if (h->n == 0) return XPBD_E_UNSUPPORTED;
    hipStream_t s = h->stream;
    const DevParams p = dev_params(h);
    HIPC(h, hipMemcpyAsync(h->A, h->X, size_t(h->n) * sizeof(float4), hipMemcpyDeviceToDevice, s));
    HIPC(h, hipMemcpyAsync(h->B, h->X, size_t(h->n) * sizeof(float4), hipMemcpyDeviceToDevice, s));
    hipEvent_t e0, e1;
    HIPC(h, hipEventCreate(&e0));
    HIPC(h, hipEventCreate(&e1));
    JacobiArgs a = jacobi_args(h, p);
    a.src = h->A;
    a.dst = h->B;
    if (a.fastdiv) {   // as in a substep: slice flags of the input and output iterates
        launch_tame_scan(h->A, h->n, tame_slot(h, h->A), s);
        launch_tame_scan(h->B, h->n, tame_slot(h, h->B), s);
        a.tame_in = tame_slot(h, h->A);
        a.tame_out = tame_slot(h, h->B);
    }
    bool alt = false;   // the sorted (key, id) pairs are in keys2 / ids2
    auto prep_collision = [&]() {
        const uint32_t np = h->nclothes == 1 ? launch_aabb_of(h->A, h->n, h->aabb, s) : 0u;
        enqueue_keys(h, h->A, np, coll_cell(h), s);
        alt = launch_radix_sort(h->keys, h->ids, h->keys2, h->ids2, h->n, h->hist, s, true, h->coll_digit_bits, h->A,
                                h->sP);
        launch_coll_table(alt ? h->keys2 : h->keys, h->n, h->ovf, h->table, h->table_slots - 1, h->slot_rec, h->epoch,
                          s);
    };
    CollArgs c{};
    if (which == XPBD_KERNEL_COLLISION_SWEEP || which == XPBD_KERNEL_RADIX_PASS) {
        prep_collision();
        c.F = h->A;
        c.sP = h->sP;
        c.rest4 = h->rest4;
        c.prev = h->X;
        c.xout = h->B;
        c.vout = h->V2;
        c.ground = h->has_ground ? h->ground : nullptr;
        c.grids = h->grids;
        c.cloth_of = h->cloth_of;
        c.skeys = alt ? h->keys2 : h->keys;
        c.sids = alt ? h->ids2 : h->ids;
        c.table = h->table;
        c.table_mask = h->table_slots - 1;
        c.epoch = h->epoch;
        c.p = p;
        c.ovf = h->ovf;
    c.xcd_seg = SWEEP_XCD_SEG;
    }
    GsArgs g{};
    if (which == XPBD_KERNEL_GS_SUBSTEP) {
        g.X = h->X;
        g.V = h->V;
        g.W = h->A;
        g.xout = h->B;
        g.vout = h->V2;
        g.ground = nullptr;
        g.aabb = h->aabb;
        g.ij = h->gs_ij;
        g.rc = h->gs_rc;
        g.batch_offsets = h->gs_off;
        g.num_batches = h->e ? h->num_batches : 0;
        g.iterations = std::max(h->cfg.solve_inner_iterations, 1u);
        g.barrier = h->barrier;
        g.error = h->err;
        g.final_write = 1;
        g.p = p;
    }
    auto one = [&]() {
        switch (which) {
            case XPBD_KERNEL_JACOBI_ITER:
                launch_jacobi(0, a, h->jacobi_block, s);
                a.src = a.dst;
                a.dst = (a.dst == h->A) ? h->B : h->A;
                if (a.fastdiv) {
                    a.tame_in = tame_slot(h, a.src);
                    a.tame_out = tame_slot(h, a.dst);
                }
                break;
            case XPBD_KERNEL_COLLISION_SWEEP:
                launch_zero_u32(h->ovf, 1, s);   // the gather clears it before a substep's sweep
                launch_coll_sweep(c, s);
                break;
            case XPBD_KERNEL_RADIX_PASS:
                launch_radix_pass(h->keys, h->ids, h->keys2, h->ids2, h->n, 0, h->hist, s, false, h->coll_digit_bits);
                break;
            case XPBD_KERNEL_GS_SUBSTEP:
            {
                uint32_t np = 0;
                enqueue_gs(h, g, false, np, s);
                break;
            }
            case XPBD_KERNEL_COPY: launch_copy4(h->X, h->B, h->n, s); break;
            case XPBD_KERNEL_JACOBI_PERSISTENT: {
                launch_zero_u32(h->barrier, GRID_SYNC_WORDS, s);
                JacobiArgs pa = jacobi_args(h, p);
                PersistArgs q{};
                q.X = h->X;   // read only: the result stays in A/B
                q.V = h->V;
                q.A = h->A;
                q.B = h->B;
                q.iterations = std::max(h->cfg.solve_inner_iterations, 1u);
                q.final_write = 0;
                q.barrier = h->barrier;
                q.error = h->err;
                if (pa.fastdiv) {
                    q.tameA = h->tameA;
                    q.tameB = h->tameB;
                }
                launch_jacobi_persistent(pa, q, h->pj_blocks, s);
                break;
            }
            default: break;
        }
    };
    if (which < XPBD_KERNEL_JACOBI_ITER || which > XPBD_KERNEL_JACOBI_PERSISTENT) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return XPBD_E_INVALID;
    }
    one();   // warm-up
    hipError_t er = hipEventRecord(e0, s);
    for (uint32_t r = 0; r < reps; ++r) one();
    if (er == hipSuccess) er = hipEventRecord(e1, s);
    HIPC(h, er);
    HIPC(h, hipEventSynchronize(e1));
    float ms = 0.f;
    HIPC(h, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPC(h, hipGetLastError());
    TRY(check_device_flag(h));
    *avg_ms = ms / float(reps);
    return XPBD_OK;
}

XPBD_API int xpbd_profile_frames(xpbd_cloth* h, uint32_t frames, int which, float* avg_ms, uint32_t* launches) {
    if (!h || !avg_ms || frames == 0) return XPBD_E_INVALID;
    if (which != XPBD_KERNEL_JACOBI_ITER && which != XPBD_KERNEL_COLLISION_SWEEP && which != XPBD_KERNEL_GS_SUBSTEP &&
        which != XPBD_KERNEL_JACOBI_PERSISTENT)
        return XPBD_E_INVALID;
    if (h->n == 0) return XPBD_E_UNSUPPORTED;
    const uint32_t S = std::max(h->cfg.solve_substeps, 1u);
    const uint32_t I = std::max(h->cfg.solve_inner_iterations, 1u);
    const size_t pairs = size_t(S) * I + 1;   // upper bound of launches of one kind per frame
    auto& pr = h->probe;
    pr.ev.assign(2 * pairs, nullptr);
    int rc = XPBD_OK;
    for (auto& e : pr.ev)
        if (hipEventCreate(&e) != hipSuccess) rc = XPBD_E_HIP;
    double total = 0.0;
    uint32_t count = 0;
    for (uint32_t f = 0; f < frames && rc == XPBD_OK; ++f) {
        pr.which = which;
        pr.used = 0;
        enqueue_frame(h);
        pr.which = -1;
        if (hipStreamSynchronize(h->stream) != hipSuccess) {
            rc = fail(h, XPBD_E_HIP, "profile_frames: stream synchronize failed");
            break;
        }
        for (size_t k = 0; k + 1 < pr.used; k += 2) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, pr.ev[k], pr.ev[k + 1]) != hipSuccess) {
                rc = fail(h, XPBD_E_HIP, "profile_frames: event timing failed");
                break;
            }
            total += ms;
            ++count;
        }
    }
    pr.which = -1;
    for (auto& e : pr.ev)
        if (e) (void)hipEventDestroy(e);
    pr.ev.clear();
    if (rc != XPBD_OK) return rc;
    HIPC(h, hipGetLastError());
    TRY(check_device_flag(h));
    *avg_ms = count ? float(total / count) : 0.f;
    if (launches) *launches = count;
    return XPBD_OK;
}

// ---------------------------------------------------------------- warm start
}  // extern "C"

namespace {

int mfail(xpbd_mlp* m, int code, const std::string& msg) {
    if (m) m->msg = msg;
    return code;
}

#define MHIPC(m, expr)                                                                                       \
    do {                                                                                                     \
        hipError_t _e = (expr);                                                                              \
        if (_e != hipSuccess)                                                                                \
            return mfail((m), _e == hipErrorOutOfMemory ? XPBD_E_NOMEM : XPBD_E_HIP,                       \
                         std::string(#expr) + ": " + hipGetErrorString(_e));                                \
    } while (0)

template <class T>
int malloc_mlp(xpbd_mlp* m, T** p, size_t count) {
    void* q = nullptr;
    MHIPC(m, hipMalloc(&q, std::max<size_t>(count * sizeof(T), 256)));
    m->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return XPBD_OK;
}

// He-style initial weights from a seeded SplitMix64 + Box-Muller stream;
// biases zero; padded rows / columns zero.
void mlp_init_params(uint64_t seed, std::vector<float>& P) {
    P.assign(MLP_PARAMS, 0.f);
    uint64_t st = seed;
    auto next = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    auto normal = [&]() {
        const double u1 = (double(next() >> 11) + 1.0) * (1.0 / 9007199254740993.0);
        const double u2 = double(next() >> 11) * (1.0 / 9007199254740992.0);
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    for (int f = 0; f < MLP_FEAT; ++f)
        for (int j = 0; j < MLP_HID; ++j) P[OFF_W1 + f * MLP_HID + j] = float(normal() * std::sqrt(2.0 / MLP_FEAT));
    for (int i = 0; i < MLP_HID; ++i)
        for (int j = 0; j < MLP_HID; ++j) P[OFF_W2 + i * MLP_HID + j] = float(normal() * std::sqrt(2.0 / MLP_HID));
    for (int i = 0; i < MLP_HID; ++i)
        for (int o = 0; o < MLP_OUT; ++o) P[OFF_W3 + i * MLP_OUTP + o] = float(normal() * std::sqrt(1.0 / MLP_HID));
}

bool mlp_param_fixed_zero(int p) {
    if (p >= OFF_W1 + MLP_FEAT * MLP_HID && p < OFF_B1) return true;
    if (p >= OFF_W3 && p < OFF_B3) return (p - OFF_W3) % MLP_OUTP >= MLP_OUT;
    if (p >= OFF_B3) return p - OFF_B3 >= MLP_OUT;
    return false;
}

}  // namespace

extern "C" {

XPBD_API int xpbd_mlp_create(uint64_t seed, xpbd_mlp** out) {
    if (!out) return XPBD_E_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return XPBD_E_NO_DEVICE;
    std::unique_ptr<xpbd_mlp, void (*)(xpbd_mlp*)> guard(new (std::nothrow) xpbd_mlp(), xpbd_mlp_destroy);
    xpbd_mlp* m = guard.get();
    if (!m) return XPBD_E_NOMEM;
    if (hipGetDevice(&m->device) != hipSuccess) return XPBD_E_HIP;
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) return XPBD_E_HIP;
    TRY(malloc_mlp(m, &m->P, MLP_PARAMS));
    TRY(malloc_mlp(m, &m->m, MLP_PARAMS));
    TRY(malloc_mlp(m, &m->v, MLP_PARAMS));
    TRY(malloc_mlp(m, &m->grad, MLP_PARAMS));
    TRY(malloc_mlp(m, &m->bf, MLP_BF));
    TRY(malloc_mlp(m, &m->partial, size_t(MLP_TRAIN_WAVES) * MLP_PSTRIDE));
    m->loss_cap = 1;
    TRY(malloc_mlp(m, &m->loss, m->loss_cap));
    std::vector<float> P;
    mlp_init_params(seed, P);
    MHIPC(m, hipMemcpy(m->P, P.data(), P.size() * sizeof(float), hipMemcpyHostToDevice));
    MHIPC(m, hipMemset(m->m, 0, MLP_PARAMS * sizeof(float)));
    MHIPC(m, hipMemset(m->v, 0, MLP_PARAMS * sizeof(float)));
    MHIPC(m, hipMemset(m->bf, 0, MLP_BF * sizeof(uint16_t)));
    launch_mlp_pack(m->P, m->bf, m->stream);
    MHIPC(m, hipStreamSynchronize(m->stream));
    *out = guard.release();
    return XPBD_OK;
}

XPBD_API void xpbd_mlp_destroy(xpbd_mlp* m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void* p : m->allocs) (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

XPBD_API const char* xpbd_mlp_last_error(const xpbd_mlp* m) { return m ? m->msg.c_str() : "null handle"; }

XPBD_API int xpbd_mlp_get_params(xpbd_mlp* m, float* params, float* out_scale) {
    if (!m || !params) return XPBD_E_INVALID;
    MHIPC(m, hipStreamSynchronize(m->stream));
    MHIPC(m, hipMemcpy(params, m->P, MLP_PARAMS * sizeof(float), hipMemcpyDeviceToHost));
    if (out_scale) *out_scale = m->out_scale;
    return XPBD_OK;
}

XPBD_API int xpbd_mlp_set_params(xpbd_mlp* m, const float* params, float out_scale) {
    if (!m || !params || !(out_scale > 0.f)) return XPBD_E_INVALID;
    std::vector<float> P(params, params + MLP_PARAMS);
    for (int p = 0; p < MLP_PARAMS; ++p)
        if (mlp_param_fixed_zero(p)) P[p] = 0.f;
    MHIPC(m, hipStreamSynchronize(m->stream));
    MHIPC(m, hipMemcpy(m->P, P.data(), MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice));
    MHIPC(m, hipMemset(m->m, 0, MLP_PARAMS * sizeof(float)));
    MHIPC(m, hipMemset(m->v, 0, MLP_PARAMS * sizeof(float)));
    m->step = 0;
    m->out_scale = out_scale;
    launch_mlp_pack(m->P, m->bf, m->stream);
    MHIPC(m, hipStreamSynchronize(m->stream));
    return XPBD_OK;
}

XPBD_API int xpbd_mlp_forward(xpbd_mlp* m, const void* d_features, uint32_t rows, void* d_out) {
    if (!m || (rows && (!d_features || !d_out))) return XPBD_E_INVALID;
    launch_mlp_forward(static_cast<const uint16_t*>(d_features), rows, m->P, m->bf, static_cast<float*>(d_out),
                       m->stream);
    MHIPC(m, hipGetLastError());
    MHIPC(m, hipStreamSynchronize(m->stream));
    return XPBD_OK;
}

XPBD_API int xpbd_mlp_gradient(xpbd_mlp* m, const void* d_features, const void* d_targets, uint32_t rows,
                               float* grad, float* loss) {
    if (!m || !grad || rows == 0 || !d_features || !d_targets) return XPBD_E_INVALID;
    launch_mlp_train(static_cast<const uint16_t*>(d_features), static_cast<const float*>(d_targets), rows,
                     m->out_scale, m->P, m->bf, m->partial, m->stream);
    launch_mlp_reduce_adam(m->partial, m->P, m->m, m->v, m->bf, m->grad, m->loss, 0.f, 1.f, 1.f, m->stream);
    MHIPC(m, hipGetLastError());
    MHIPC(m, hipMemcpyAsync(grad, m->grad, MLP_PARAMS * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    float sq = 0.f;
    MHIPC(m, hipMemcpyAsync(&sq, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MHIPC(m, hipStreamSynchronize(m->stream));
    if (loss) *loss = sq / (3.0f * float(rows));
    return XPBD_OK;
}

XPBD_API int xpbd_mlp_train(xpbd_mlp* m, const void* d_features, const void* d_targets, uint32_t rows,
                            uint32_t steps, uint32_t batch, float lr, float* loss_curve) {
    if (!m || !d_features || !d_targets || rows == 0 || batch == 0 || !(lr > 0.f)) return XPBD_E_INVALID;
    batch = std::min(batch, rows);
    if (steps > m->loss_cap) {
        float* q = nullptr;
        TRY(malloc_mlp(m, &q, steps));
        m->loss = q;
        m->loss_cap = steps;
    }
    const uint16_t* X = static_cast<const uint16_t*>(d_features);
    const float* T = static_cast<const float*>(d_targets);
    const uint32_t nb = rows / batch;   // whole minibatches per epoch (rows in their given order)
    double b1t = std::pow(0.9, double(m->step)), b2t = std::pow(0.999, double(m->step));
    for (uint32_t k = 0; k < steps; ++k) {
        const uint32_t start = (k % nb) * batch;
        ++m->step;
        b1t *= 0.9;
        b2t *= 0.999;
        // cosine decay from lr to lr / 10 over the call
        const float lrk = lr * (0.1f + 0.9f * 0.5f * (1.0f + std::cos(3.14159265f * float(k) / float(steps))));
        launch_mlp_train(X + size_t(start) * MLP_IN, T + size_t(start) * 3, batch, m->out_scale, m->P, m->bf,
                         m->partial, m->stream);
        launch_mlp_reduce_adam(m->partial, m->P, m->m, m->v, m->bf, nullptr, m->loss + k, lrk, float(b1t),
                               float(b2t), m->stream);
    }
    MHIPC(m, hipGetLastError());
    if (loss_curve) {
        std::vector<float> sq(steps);
        MHIPC(m, hipMemcpyAsync(sq.data(), m->loss, steps * sizeof(float), hipMemcpyDeviceToHost, m->stream));
        MHIPC(m, hipStreamSynchronize(m->stream));
        for (uint32_t k = 0; k < steps; ++k) loss_curve[k] = sq[k] / (3.0f * float(batch));
    }
    MHIPC(m, hipStreamSynchronize(m->stream));
    return XPBD_OK;
}

XPBD_API int xpbd_warm_sample(xpbd_cloth* h, uint32_t iterations, void* d_features, void* d_targets) {
    if (!h || !d_features || !d_targets) return XPBD_E_INVALID;
    if (!h->warm_ok) return fail(h, XPBD_E_UNSUPPORTED, "warm start needs a grid mesh (cloth columns)");
    if (h->n == 0) return XPBD_OK;
    hipStream_t s = h->stream;
    const DevParams p = dev_params(h);
    launch_predict(h->X, h->V, h->sP, p, s);   // predicted positions, kept for the targets
    launch_predict(h->X, h->V, h->A, p, s);
    WarmFeatArgs fa{};
    fa.xp = h->sP;
    fa.vel = h->V;
    fa.rest4 = h->rest4;
    fa.nbr8 = h->nbr8;
    fa.vaux = h->vaux;
    fa.dt = h->u.dt;
    fa.n = h->n;
    fa.feat = static_cast<uint16_t*>(d_features);
    launch_warm_features(fa, s);
    JacobiArgs a = jacobi_args(h, p);
    float4* cur = h->A;
    float4* nxt = h->B;
    for (uint32_t k = 0; k < iterations; ++k) {
        a.src = cur;
        a.dst = nxt;
        launch_jacobi(0, a, h->jacobi_block, s);
        std::swap(cur, nxt);
    }
    launch_warm_targets(h->sP, cur, h->vaux, h->n, static_cast<float*>(d_targets), s);
    HIPC(h, hipGetLastError());
    HIPC(h, hipStreamSynchronize(s));
    return XPBD_OK;
}

XPBD_API int xpbd_warm_features(xpbd_cloth* h, void* d_features) {
    if (!h || !d_features) return XPBD_E_INVALID;
    if (!h->warm_ok) return fail(h, XPBD_E_UNSUPPORTED, "warm start needs a grid mesh (cloth columns)");
    if (h->n == 0) return XPBD_OK;
    const DevParams p = dev_params(h);
    launch_predict(h->X, h->V, h->sP, p, h->stream);
    WarmFeatArgs fa{};
    fa.xp = h->sP;
    fa.vel = h->V;
    fa.rest4 = h->rest4;
    fa.nbr8 = h->nbr8;
    fa.vaux = h->vaux;
    fa.dt = h->u.dt;
    fa.n = h->n;
    fa.feat = static_cast<uint16_t*>(d_features);
    launch_warm_features(fa, h->stream);
    HIPC(h, hipGetLastError());
    HIPC(h, hipStreamSynchronize(h->stream));
    return XPBD_OK;
}

XPBD_API int xpbd_set_warm_start(xpbd_cloth* h, xpbd_mlp* m, uint32_t warm_iterations, float residual_threshold) {
    if (!h) return XPBD_E_INVALID;
    if (m) {
        if (!h->warm_ok) return fail(h, XPBD_E_UNSUPPORTED, "warm start needs a grid mesh (cloth columns)");
        if (h->cfg.solver != XPBD_SOLVER_JACOBI) return fail(h, XPBD_E_UNSUPPORTED, "warm start is for the Jacobi solver");
        if (!(residual_threshold >= 0.f)) return XPBD_E_INVALID;
        TRY(ensure_warm_buffers(h));
        if (m->device != h->device) return fail(h, XPBD_E_INVALID, "network and cloth on different devices");
        HIPC(h, hipStreamSynchronize(m->stream));
    }
    h->warm = m;
    h->warm_iters = warm_iterations;
    h->warm_threshold = residual_threshold;
    drop_graph(h);
    return m && std::isfinite(residual_threshold) ? verify_coresident(h, false, true) : XPBD_OK;
}

XPBD_API int xpbd_warm_stats(xpbd_cloth* h, uint32_t* substeps, uint32_t* fallbacks, uint32_t reset) {
    if (!h) return XPBD_E_INVALID;
    uint32_t f[4] = {0, 0, 0, 0};
    if (h->warm_flag) {
        HIPC(h, hipMemcpyAsync(f, h->warm_flag, sizeof(f), hipMemcpyDeviceToHost, h->stream));
        HIPC(h, hipStreamSynchronize(h->stream));
        if (reset) HIPC(h, hipMemset(h->warm_flag, 0, sizeof(f)));
    }
    if (substeps) *substeps = f[1];
    if (fallbacks) *fallbacks = f[2];
    return XPBD_OK;
}

XPBD_API int xpbd_selftest_division(int exp_lo, int exp_hi, uint64_t pairs, uint64_t seed, uint32_t* out5) {
    if (!out5 || exp_lo > exp_hi || exp_lo < -126 || exp_hi > 127) return XPBD_E_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return XPBD_E_NO_DEVICE;
    uint32_t* d = nullptr;
    if (hipMalloc(&d, 8 * sizeof(uint32_t)) != hipSuccess) return XPBD_E_NOMEM;
    hipStream_t s = nullptr;
    int rc = XPBD_OK;
    if (hipMemset(d, 0, 8 * sizeof(uint32_t)) != hipSuccess || hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        rc = XPBD_E_HIP;
    } else {
        launch_selftest_rcp(exp_lo, exp_hi, d, d + 1, s);
        if (pairs) launch_selftest_div(pairs, seed, d + 2, d + 3, s);
        uint32_t h[8] = {0};
        if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess ||
            hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess)
            rc = XPBD_E_HIP;
        std::memcpy(out5, h, 5 * sizeof(uint32_t));
    }
    if (s) (void)hipStreamDestroy(s);
    (void)hipFree(d);
    return rc;
}

XPBD_API int xpbd_selftest_sqrt(int exp_lo, int exp_hi, uint32_t* out2) {
    if (!out2 || exp_lo > exp_hi || exp_lo < -126 || exp_hi > 127) return XPBD_E_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return XPBD_E_NO_DEVICE;
    uint32_t* d = nullptr;
    if (hipMalloc(&d, 4 * sizeof(uint32_t)) != hipSuccess) return XPBD_E_NOMEM;
    hipStream_t s = nullptr;
    int rc = XPBD_OK;
    if (hipMemset(d, 0, 4 * sizeof(uint32_t)) != hipSuccess ||
        hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        rc = XPBD_E_HIP;
    } else {
        launch_selftest_sqrt(exp_lo, exp_hi, d, d + 1, s);
        uint32_t h[4] = {0};
        if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess ||
            hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess)
            rc = XPBD_E_HIP;
        out2[0] = h[0];
        out2[1] = h[1];
    }
    if (s) (void)hipStreamDestroy(s);
    (void)hipFree(d);
    return rc;
}

XPBD_API int xpbd_get_layout(xpbd_cloth* h, uint32_t* out4) {
    if (!h || !out4) return XPBD_E_INVALID;
    out4[0] = h->js_on ? XPBD_ADJ_SHARED_TILES
                       : h->nbtile ? XPBD_ADJ_GRID_TILES : (h->nbpk ? XPBD_ADJ_PACKED : XPBD_ADJ_PLAIN);
    out4[1] = h->tile_cols;
    out4[2] = h->tile_rows;
    out4[3] = fast_division_active(h) ? 1u : 0u;
    return XPBD_OK;
}

XPBD_API int xpbd_debug_raise_device_error(xpbd_cloth* h) {
    if (!h) return XPBD_E_INVALID;
    const uint32_t one = 1;
    HIPC(h, hipMemcpyAsync(h->err, &one, sizeof(one), hipMemcpyHostToDevice, h->stream));
    HIPC(h, hipStreamSynchronize(h->stream));
    return XPBD_OK;
}

XPBD_API int xpbd_debug_gs_tiles(xpbd_cloth* h, uint32_t tile, uint32_t blocks, uint32_t* out4) {
    if (!h || !out4) return XPBD_E_INVALID;
    if (tile) {
        const uint32_t old_tile = h->gt.tile, old_blocks = h->gt_blocks;
        TRY(build_gs_tiles(h, tile, blocks));
        if (!h->gt_ok) {
            if (old_tile) TRY(build_gs_tiles(h, old_tile, old_blocks));
            release_gs_host_rows(h);
            return fail(h, XPBD_E_INVALID, "no per-tile flag plan at this tile size");
        }
        release_gs_host_rows(h);
        h->gg_on = false;   // the flag form runs these index-range tiles
        drop_graph(h);
        apply_tuning(h);
        TRY(verify_coresident(h, h->cfg.solver == XPBD_SOLVER_GAUSS_SEIDEL, uses_persistent_jacobi(h)));
    }
    out4[0] = h->gt_ok ? h->gt.tile : 0u;
    out4[1] = h->gt_ok ? h->gt.ntiles : 0u;
    out4[2] = h->gt_ok ? h->gt_blocks : 0u;
    out4[3] = h->gt_ok ? h->gt_deps : 0u;
    return XPBD_OK;
}

XPBD_API int xpbd_debug_gs_grid(xpbd_cloth* h, uint32_t rows, uint32_t* out4) {
    if (!h || !out4) return XPBD_E_INVALID;
    if (rows) {
        const uint32_t old = h->gg_ok ? h->gg.th : 0u;
        TRY(build_gs_grid(h, rows));
        if (!h->gg_ok) {
            if (old) TRY(build_gs_grid(h, old));
            release_gs_host_rows(h);
            return fail(h, XPBD_E_INVALID, "no LDS grid-tile plan at this tile height");
        }
        release_gs_host_rows(h);
        h->gg_on = true;
        drop_graph(h);
        apply_tuning(h);
        TRY(verify_coresident(h, h->cfg.solver == XPBD_SOLVER_GAUSS_SEIDEL, uses_persistent_jacobi(h)));
    }
    out4[0] = h->gg_ok ? h->gg.th : 0u;
    out4[1] = h->gg_ok ? h->gg.tx * h->gg.ty : 0u;
    out4[2] = h->gg_ok ? gs_grid_block(h->gg.th) : 0u;
    out4[3] = h->gg_on ? 1u : 0u;
    return XPBD_OK;
}

XPBD_API int xpbd_debug_jacobi_resident(xpbd_cloth* h, uint32_t rows, uint32_t* out4) {
    if (!h || !out4) return XPBD_E_INVALID;
    if (rows) {
        const uint32_t old = h->jr_ok ? h->jr.th : 0u;
        TRY(build_jacobi_resident(h, rows));
        if (!h->jr_ok) {
            if (old) TRY(build_jacobi_resident(h, old));
            return fail(h, XPBD_E_INVALID, "no LDS-resident Jacobi tiles at this tile height");
        }
        drop_graph(h);
        apply_tuning(h);
        TRY(verify_coresident(h, h->cfg.solver == XPBD_SOLVER_GAUSS_SEIDEL, uses_persistent_jacobi(h)));
    }
    out4[0] = h->jr_ok ? h->jr.th : 0u;
    out4[1] = h->jr_ok ? (h->tile_cols / 64) * h->jr.ty : 0u;
    out4[2] = h->jr_ok ? h->jr.npairs : 0u;
    out4[3] = h->jr_on;
    return XPBD_OK;
}

XPBD_API int xpbd_debug_collision_keys(xpbd_cloth* h, float cell_size, uint32_t* grid_out, uint32_t* sorted_keys,
                                       uint32_t* sorted_ids) {
    if (!h || !grid_out || !sorted_keys || !sorted_ids || !(cell_size > 0.f)) return XPBD_E_INVALID;
    hipStream_t s = h->stream;
    const uint32_t np = h->nclothes == 1 ? launch_aabb_of(h->X, h->n, h->aabb, s) : 0u;
    enqueue_keys(h, h->X, np, cell_size, s);
    const bool alt = launch_radix_sort(h->keys, h->ids, h->keys2, h->ids2, h->n, h->hist, s, true, h->coll_digit_bits);
    HIPC(h, hipGetLastError());
    HIPC(h, hipStreamSynchronize(s));
    HIPC(h, hipMemcpy(grid_out, h->grids, size_t(h->nclothes) * sizeof(CollGrid), hipMemcpyDeviceToHost));
    HIPC(h, hipMemcpy(sorted_keys, alt ? h->keys2 : h->keys, size_t(h->n) * 4, hipMemcpyDeviceToHost));
    HIPC(h, hipMemcpy(sorted_ids, alt ? h->ids2 : h->ids, size_t(h->n) * 4, hipMemcpyDeviceToHost));
    return XPBD_OK;
}

}  // extern "C"

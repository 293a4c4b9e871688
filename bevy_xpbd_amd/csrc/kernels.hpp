// Launch interface between the host runtime (xpbd_api.cpp) and the HIP
// kernels (kernels.hip).  Plain structs passed by value as kernel arguments.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace xpbd {

// Per-substep scalars (ClothSimParamsGpu, cloth_compute.rs:246-273, minus the
// padding; damp = exp(-k dt) is evaluated once on the host with libm expf,
// the function Rust's f32::exp calls, xpbd_cpu.rs:402).
struct DevParams {
    float dt, inv_dt, inv_dt_sq, omega;
    float gx, gy, gz, damp;
    float grab_x, grab_y, grab_z, grab_k;
    int32_t grab_idx;
    uint32_t grab_active;
    float thickness, coll_scale;
    uint32_t n;
    uint32_t nslices;
};

// Sliced-ELL adjacency: 64 consecutive particles per slice (one wavefront),
// entry k of lane l at soff[s] + k*64 + l.  Each particle's entries keep the
// reference CSR order (the gather's summation order); a particle's unused
// tail entries point at the particle itself (j = i, rest = compliance = 0;
// packed: 0), which the gather skips as a zero-length constraint.
struct Sell {
    const uint32_t* soff;
    const uint32_t* swid;
    const uint32_t* j;
    const float* rest;
    const float* comp;
    // Packed form (used when every entry fits): one u32 per entry,
    // (pair index << 16) | (uint16)(j - i), with pair = (rest, compliance)
    // looked up in a table of the mesh's distinct pairs; padding is 0.
    const uint32_t* packed;
    const float2* pairs;
    // Grid tiles (one grid cloth, tcols a multiple of 64): the packed entries
    // re-encoded as (pair index x 16) << 16 | (int16) byte offset of the
    // neighbour in the block's LDS tile ((dy * JT_W + dx) x 16), so both
    // addresses come from one operation each; nullptr: no tiled path
    const uint32_t* tpacked;
    uint32_t tcols, trows;
};

// Constraint-shared grid tiles (k_jacobi_shared): the tiles of the grid-tiled
// kernel, with each constraint touching a tile evaluated once per tile.
// Directions: the forward offsets (dy > 0, or dy == 0 and dx > 0) of the
// mesh's constraints, at most JS_DIRS; the constraint of particle a in
// direction t joins a and a + (dy_t, dx_t), and a is its anchor.
//   jpair[t * n + a]: its pair index (0xFFFF: none).
//   Result slots in LDS: slot row r = (direction t, anchor tile row lr) for
//   lr in [2 - dy_t, 2 + JT_ROWS) (tile rows as in the position tile), 68
//   columns each (tile columns), then one zero slot at index nslots.
//   jobs[0 .. njobs_main): one slot row each, anchors at the tile's 64
//   columns; then njobs_edge wavefronts of per-lane items for anchors left or
//   right of the tile whose partner is inside (0xFFFFFFFF: none).  An item
//   packs dy | (dx + 2) << 2 | t << 5 | lr << 8 | cl << 12 | row << 20.
//   entries: per particle, its CSR entries in order as 16-bit codes (slot -
//   lane) | sign << 15 -- the slot of the entry's constraint relative to the
//   lane's column, sign 1 when the particle is the constraint's second
//   endpoint; padding codes point at the zero slot.  Two codes per word,
//   slice s at words [eoff[s], eoff[s] + ceil(width / 2) * 64).
struct SharedTiles {
    const uint16_t* jpair;
    const uint32_t* jobs;
    const uint32_t* entries;
    const uint32_t* eoff;
    uint32_t njobs_main, njobs_edge, nslots;   // nslots = slot rows x 68
};
constexpr uint32_t JS_DIRS = 8;
constexpr uint32_t JS_PAIRS_MAX = 1024;
constexpr uint32_t JT_PAIRS_MAX = 1024;   // the tiled kernels' staged pair table (float4)
// LDS of one k_jacobi_shared block (position tile, result slots, pairs).
size_t jacobi_shared_lds_bytes(uint32_t nslots, uint32_t num_pairs);

// LDS tile of the grid-tiled Jacobi: JT_ROWS rows of 64 particles (one
// wavefront each) and a 2-particle halo on every side
constexpr uint32_t JT_ROWS = 8;
constexpr uint32_t JT_W = 64 + 4;
constexpr uint32_t JT_H = JT_ROWS + 4;

// A tame coordinate is +0 or of magnitude in [2^-50, 2^20]: on slices whose
// neighbourhood is tame the Jacobi kernels use the short correctly rounded
// division and square root (exact on that range).
constexpr float TAME_LO = 8.881784197001252e-16f;   // 2^-50
constexpr float TAME_HI = 1048576.0f;               // 2^20
__device__ __forceinline__ bool tame_coord(float c) {
    const float a = fabsf(c);
    return __float_as_uint(c) == 0u || (a >= TAME_LO && a <= TAME_HI);
}
__device__ __forceinline__ bool tame3(float4 v) { return tame_coord(v.x) && tame_coord(v.y) && tame_coord(v.z); }

enum JacobiMode : uint32_t {
    JM_PREDICT = 1,   // inputs are the substep-start state; predict on the fly
    JM_FINAL = 2,     // write the new state in place + velocity (no collision)
    JM_AABB = 4,      // accumulate the bounds of the output (collision follows)
    JM_KEEP = 8,      // the caller stores the result (LDS-resident tiles)
};

struct JacobiArgs {
    const float4* src;   // JM_PREDICT: X; otherwise the previous iterate
    const float4* vel;   // JM_PREDICT: V
    float4* dst;         // iterate out (JM_FINAL: X, in place)
    float4* vout;        // JM_FINAL: V
    const uint8_t* ground;
    float* aabb;         // JM_AABB: one (lo xyz, hi xyz) partial per block, stride 8
    Sell sell;
    SharedTiles sh;      // sh.jobs == nullptr: no constraint-shared path
    DevParams p;
    uint32_t remap;      // 1: XCD-aware block order
    uint32_t fastdiv;    // 1: shared-reciprocal correctly rounded divisions
    uint32_t num_pairs;  // packed: size of sell.pairs
    const uint32_t* gate;   // non-null: the launch does nothing unless *gate != 0
    // fast-division proof (FASTDIV kernels), per slice of 64 particles:
    // tame_in[s] != 0 when every coordinate of slice s of the input iterate is
    // 0 or of magnitude in [2^-50, 2^20]; srange[s] = first and last slice
    // holding a neighbour of slice s (x > y: too wide, never fast).  The launch
    // writes tame_out[s] for the slices it produces.
    const uint8_t* tame_in;
    uint8_t* tame_out;
    const uint2* srange;
};

// Persistent Jacobi substep: predict X,V -> A, iterate A/B; the last iterate
// becomes the state X,V (final_write) or is left for a collision pass in A
// (even iteration count) or B (odd).
struct PersistArgs {
    float4* X;
    float4* V;
    float4* A;
    float4* B;
    uint32_t iterations;
    uint32_t final_write;
    uint32_t* barrier;
    uint32_t* error;
    const uint32_t* gate;   // non-null: the launch does nothing unless *gate != 0
    uint32_t probe;         // 1: only the barrier set-up and one barrier (co-residency check)
    // per-slice tame bytes of A and B (both non-null and fast division on:
    // the short exact operations on slices whose neighbourhood is tame)
    uint8_t* tameA;
    uint8_t* tameB;
};

// Cell grid of one cloth in the collision broad phase (device memory, 12 words).
struct CollGrid {
    float lo[3], hi[3];
    float cell, inv_cell;
    uint32_t dims[3];
    uint32_t key_base;   // cloth << (32 - b): first key of the cloth's range
};

struct CollArgs {
    const float4* F;     // positions after the constraint iterations
    const float4* sP;    // sorted order: position xyz, w = cell key bits
    const float4* rest4; // by particle id: rest position xyz, w = inverse mass
    const float4* prev;  // X (substep start) for the velocity update
    float4* xout;
    float4* vout;
    const uint8_t* ground;
    const CollGrid* grids;      // one per cloth
    const uint32_t* cloth_of;   // particle -> cloth (nullptr: one cloth)
    const uint32_t* skeys;
    const uint32_t* sids;
    const uint4* table;
    uint32_t table_mask;
    // direct group table (one cloth): the current pass tag; nullptr: the
    // hashed table (batches)
    const uint32_t* epoch;
    DevParams p;
    // particles with more than LMAX candidates, resolved by a second kernel:
    // [0] count (cleared by launch_coll_table), [2..] sorted indices; n + 2 words
    uint32_t* ovf;
    uint32_t xcd_seg;   // sweep blocks per XCD segment (0: block b sweeps chunk b)
};

// Groups of the direct group table of one cloth: its keys are below 2^27
// (the grid's cell cap), so its 8-cell groups are below 2^24.
constexpr uint32_t ONE_CLOTH_GROUPS = 1u << 24;

// Sweep segment length in blocks of 256 particles.  On the 1024^2 sheet 16
// cut the sweep's fetched bytes from 223 to 133 MB per launch at the same
// time (4..64 all within 1% of each other).
constexpr uint32_t SWEEP_XCD_SEG = 16;

// Words of the grid-barrier state of a persistent launch (zeroed before each
// launch with launch_zero_u32).
constexpr uint32_t GRID_SYNC_WORDS = 2048;

struct GsArgs {
    const float4* X;
    const float4* V;
    float4* W;           // work buffer (final iterate)
    float4* xout;        // FINAL: X
    float4* vout;        // FINAL: V
    const uint8_t* ground;
    float* aabb;         // one bounds partial per block (collision follows)
    const uint2* ij;     // colour-sorted (i, j)
    const float2* rc;    // colour-sorted (rest, compliance)
    const uint32_t* batch_offsets;
    uint32_t num_batches;
    uint32_t iterations;
    uint32_t* barrier;
    uint32_t* error;
    uint32_t final_write;
    uint32_t probe;      // 1: only the barrier set-up (k_gs_substep: and one barrier), the co-residency check
    DevParams p;
};

// Coloured GS ordered by per-tile flags instead of grid barriers.  The
// particles are cut into tiles of `tile` consecutive indices; a constraint
// belongs to the tile of its lower endpoint, so within a colour (sorted by
// lower endpoint) each tile's constraints are one run.  The steps of a substep
// are the predict, the I x colours constraint steps and the final write (or
// bounds); a tile's step s may start once each tile in its dependency list
// (every other tile whose particle range [first index, highest endpoint of its
// constraints] overlaps its own) has finished step s - 1.
struct GsTiles {
    const uint32_t* toff;   // colours x (ntiles + 1): first constraint of each tile in each colour
    const uint32_t* doff;   // ntiles + 1: dependency list offsets into dep
    const uint32_t* dep;    // dependency lists (tile ids)
    uint32_t* flag;         // ntiles: steps finished (zeroed by launch_gs_tiles)
    uint32_t ntiles;
    uint32_t tile;          // particles per tile
};
// A tile's dependencies are checked by the lanes of one wavefront.
constexpr uint32_t GS_TILE_DEPS_MAX = 64;

// Coloured GS on LDS-resident grid tiles (k_gs_grid): one grid cloth, tiles
// of 64 columns x th rows (th >= 8), one block each.  A constraint belongs to
// the tile of its lower endpoint; its endpoints are stored as positions in
// the tile's window ((th + 4) x 68 particles: the tile and 2 around it), with
// bit 15 set on a position that other tiles can touch.
struct GsGrid {
    // tiles x colours x 2 + 1 offsets: each tile's constraints of each
    // colour, first those touching only its interior, then the others
    const uint32_t* coff;
    const uint32_t* con;    // window positions (i | j << 16) in that order
    const float2* crc;      // (rest, compliance) in that order
    uint32_t* flag;         // tiles: steps finished (zeroed by launch_gs_grid)
    uint32_t tx, ty;        // tiles per row of tiles, rows of tiles
    uint32_t cols, rows;    // the cloth's grid
    uint32_t th;            // tile rows
    uint32_t colours;
};

// Jacobi substep on LDS-resident grid tiles (k_jacobi_resident): one grid
// cloth (the grid-tile layout: sell.tcols a multiple of 64), tiles of 64
// columns x th rows (th in {16, 32, 64}; blocks of th x 16 threads), all
// resident at once.  With PersistArgs: predict X,V, I iterations, the state
// (final_write) or the last iterate in A (even I) / B (odd I) with one bounds
// partial per tile in aabb.  A and B carry the rings between tiles.
struct JResident {
    uint32_t* flag;      // tiles (zeroed by launch_jacobi_resident)
    float* aabb;         // bounds partials (tiles), collision follows
    uint32_t th;         // tile rows
    uint32_t ty;         // rows of tiles
    uint32_t npairs;     // pairs staged in LDS (0: the pair table is read from memory)
};

// ---- launchers (kernels.hip) ----
uint32_t jacobi_blocks(uint32_t nslices, uint32_t block);
uint32_t launch_jacobi(uint32_t mode, const JacobiArgs& a, uint32_t block, hipStream_t s);
void launch_predict(const float4* X, const float4* V, float4* P, DevParams p, hipStream_t s,
                    uint8_t* tame_out = nullptr);
void launch_finalize(const float4* F, float4* X, float4* V, const uint8_t* ground, DevParams p,
                     hipStream_t s);
// per-cloth bounds of P: piece b reduces particles [span[b].x, span[b].y)
// of one cloth into part[8b..]; cloth c owns pieces [pfirst[c], pfirst[c+1]);
// then the grids
void launch_coll_grids(const float4* P, const uint2* span, uint32_t nspans, const uint32_t* pfirst,
                       uint32_t nclothes, float* part, float cell_size, uint32_t local_bits, CollGrid* grids,
                       hipStream_t s);
uint32_t cloth_span_max();   // particles per piece at most
// cell keys (and ids = identity) plus the sort's pass-0 tile counts in hist;
// also empties the group-table slots the previous pass recorded in slot_rec.
// part != nullptr (one cloth): the grid is derived from the nparts bounds
// partials and stored to grids[0]; otherwise grids[] is read (batches).
void launch_coll_keys(const float4* F, uint32_t n, CollGrid* grids, const uint32_t* cloth_of, uint32_t* keys,
                      uint32_t* ids, uint4* table, const uint32_t* slot_rec, uint32_t* hist, uint32_t digit_bits,
                      const float* part, uint32_t nparts, float cell_size, uint32_t* epoch, hipStream_t s);
// Stable sort of (keys, ids) by key: 4 passes of 8-bit digits (32-bit keys) or
// 3 passes of 9-bit digits (keys below 2^27: one cloth's grid).  Returns true
// when the result is in (keys_alt, ids_alt) (an odd pass count), false when it
// is in (keys, ids).  first_hist_done: hist already holds pass 0's tile counts
// (launch_coll_keys with the same digit_bits).
// F, sP (a collision sort): the last pass also writes the sweep's sorted
// copy sP[s] = (F[id] xyz, key bits).
bool launch_radix_sort(uint32_t* keys, uint32_t* ids, uint32_t* keys_alt, uint32_t* ids_alt, uint32_t n,
                       uint32_t* hist, hipStream_t s, bool first_hist_done, uint32_t digit_bits,
                       const float4* F = nullptr, float4* sP = nullptr);
uint32_t radix_hist_words(uint32_t n);
void launch_radix_pass(const uint32_t* kin, const uint32_t* iin, uint32_t* kout, uint32_t* iout, uint32_t n,
                       uint32_t shift, uint32_t* hist, hipStream_t s, bool hist_done = false,
                       uint32_t digit_bits = 8, const float4* F = nullptr, float4* sP = nullptr);
// group-table inserts over the sorted keys (slots recorded in slot_rec, n
// words; the table must be empty on entry), overflow count clear
void launch_coll_table(const uint32_t* skeys, uint32_t n, uint32_t* ovf, uint4* table, uint32_t mask,
                       uint32_t* slot_rec, const uint32_t* epoch, hipStream_t s);
void launch_coll_sweep(const CollArgs& a, hipStream_t s);
void launch_gs(const GsArgs& a, uint32_t blocks, uint32_t block, hipStream_t s);
// The constraint iterations of a GS substep as one launch per colour and
// iteration (iterations x colours launches; off = the colour offsets on the
// host): the launch boundaries order the colours.  W must hold the predicted
// positions; a.W, a.ij, a.rc and a.p are read.
void launch_gs_colours(const GsArgs& a, const std::vector<uint32_t>& off, uint32_t iterations, hipStream_t s);
uint32_t gs_max_blocks(uint32_t block);
// One GS substep ordered by per-tile flags (zeroes t.flag first).  blocks
// must be resident at once (block b takes tiles b, b + blocks, ...); every
// wait is bounded and sets a.error on expiry.  a.W must hold n * 16 < 2^32
// bytes.  With a.final_write == 0 one bounds partial per block goes to a.aabb.
void launch_gs_tiles(const GsArgs& a, const GsTiles& t, uint32_t blocks, hipStream_t s);
uint32_t gs_tiles_max_blocks();
// The grid-tile GS substep (k_gs_grid): tiles x blocks of gs_grid_block(th)
// threads and gs_grid_lds_bytes(th) of LDS, all resident at once (zeroes
// t.flag first).  gs_grid_prepare raises the kernel's LDS limit (once per
// tile height, before launching or capturing).
uint32_t gs_grid_lds_bytes(uint32_t th);
uint32_t gs_grid_block(uint32_t th);
hipError_t gs_grid_prepare(uint32_t th);
uint32_t gs_grid_max_blocks(uint32_t th);
void launch_gs_grid(const GsArgs& a, const GsGrid& t, hipStream_t s);
uint32_t jacobi_resident_lds_bytes(uint32_t th, uint32_t npairs);
hipError_t jacobi_resident_prepare(uint32_t th, uint32_t npairs);
uint32_t jacobi_resident_max_blocks(uint32_t th, uint32_t npairs);
void launch_jacobi_resident(const JacobiArgs& a, const PersistArgs& q, const JResident& t, hipStream_t s);
void launch_copy4(const float4* a, float4* b, uint32_t n, hipStream_t s);
// per-slice tame bytes of P (1: every coordinate 0 or of magnitude in [2^-50, 2^20])
void launch_tame_scan(const float4* P, uint32_t n, uint8_t* slice_tame, hipStream_t s);
// gated variants for the warm-start fallback: run only if *gate != 0
void launch_predict_gated(const float4* X, const float4* V, float4* P, DevParams p, const uint32_t* gate,
                          hipStream_t s);
// copy a -> b only if *gate != 0
void launch_copy4_if(const float4* a, float4* b, uint32_t n, const uint32_t* gate, hipStream_t s);
// warm-start gate: flag[0] = relative residual of X (f32 per constraint) >
// threshold or NaN; flag[1] += 1; flag[2] += flag[0]; clears barrier
// (GRID_SYNC_WORDS).  acc: warm_gate_words() words (256-byte aligned):
// arrival counters, zero before the first launch and left zero by each, and
// one partial-sum slot per block, summed in block order (deterministic)
uint32_t warm_gate_words();
void launch_warm_gate(const float4* X, const uint2* ij, const float* rest, uint32_t e, double rest_sq,
                      float threshold, uint32_t* acc, uint32_t* flag, uint32_t* barrier, hipStream_t s);
// finalize from F, or from alt if *gate != 0
void launch_finalize_sel(const float4* F, const float4* alt, const uint32_t* gate, float4* X, float4* V,
                         const uint8_t* ground, DevParams p, hipStream_t s);
uint32_t persistent_max_blocks(bool packed, bool fastdiv);
void launch_jacobi_persistent(const JacobiArgs& a, const PersistArgs& q, uint32_t blocks, hipStream_t s);
void launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t s);
void launch_publish(const float4* X, uint32_t n, const uint32_t* tris, const uint32_t* vt_off,
                    const uint32_t* vt_tri, float4* dpos, float4* dnrm, hipStream_t s);
// sum of C^2 over the constraints (ij, rest) in double: one partial per
// block into part[0, blocks), blocks <= residual_max_blocks(); returns blocks
uint32_t residual_max_blocks();
// The same relative residual over caller-owned device arrays (separate
// endpoint arrays, as ClothSimConfig holds them), allocating nothing: block
// partials live in a fixed device array of 2 x residual_noalloc_max_blocks()
// doubles; out = device double.  Calls that overlap on one device race.
uint32_t residual_noalloc_max_blocks();
hipError_t launch_residual_noalloc(const float4* X, uint32_t n, const uint32_t* ci, const uint32_t* cj,
                                   const float* rest, uint32_t e, hipStream_t s, double* out);
uint32_t launch_residual(const float4* X, const uint2* ij, const float* rest, uint32_t e, double* part,
                         hipStream_t s);
uint32_t launch_aabb_of(const float4* P, uint32_t n, float* part, hipStream_t s);
void launch_selftest_rcp(int exp_lo, int exp_hi, uint32_t* bad, uint32_t* first, hipStream_t s);
void launch_selftest_div(uint64_t n, uint64_t seed, uint32_t* bad, uint32_t* first, hipStream_t s);
void launch_selftest_sqrt(int exp_lo, int exp_hi, uint32_t* bad, uint32_t* first, hipStream_t s);

}  // namespace xpbd

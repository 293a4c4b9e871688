/*
 * xpbd.h — C ABI of libxpbd_hip.so, the MI355X-native XPBD cloth compute path.
 *
 * Drop-in boundary for the reference crate's compute path (mate-h/bevy_xpbd,
 * crate bevy_softbody).  Each entry point names the reference interface it
 * replaces as file:line.  Plain pointers and sizes only; no torch or HIP types
 * appear in the signatures (streams and device pointers are void*).
 *
 * Threading: calls on one xpbd_cloth are externally synchronised (the
 * reference runs its render-world system on one thread).  Different handles
 * are independent.  All device work of a handle is ordered on its stream.
 * One exception: the persistent launches (coloured Gauss-Seidel, persistent
 * Jacobi, the learned warm start's gated fallback) need every block of their
 * grid resident on the GPU at once, so while one runs, no other stream may
 * run work on the same GPU: order handles that step on one GPU with
 * xpbd_sync.  Otherwise the persistent launch's bounded wait can expire, and
 * the next synchronising call reports XPBD_E_STATE.
 */
#ifndef XPBD_H
#define XPBD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XPBD_API __attribute__((visibility("default")))
#define XPBD_ABI_VERSION 1

/* Status codes.  The reference early-returns silently on missing resources
 * (cloth_compute.rs:1222-1239) and panics on shader errors
 * (cloth_jacobi.rs:668); this ABI reports both as codes instead. */
enum {
    XPBD_OK = 0,
    XPBD_E_INVALID = 1,      /* bad argument or inconsistent topology       */
    XPBD_E_HIP = 2,          /* HIP runtime error (message in last_error)   */
    XPBD_E_NOMEM = 3,        /* device or host allocation failed            */
    XPBD_E_UNSUPPORTED = 4,  /* feature not available for this mesh/config */
    XPBD_E_STATE = 5,        /* wrong call order / device-side guard fired  */
    XPBD_E_NO_DEVICE = 6     /* no HIP device visible                       */
};

/* Solver switch: replaces the compile-time cargo features solver-jacobi /
 * solver-gauss-seidel (Cargo.toml:27-30, lib.rs:3-4) with a runtime enum. */
enum { XPBD_SOLVER_JACOBI = 0, XPBD_SOLVER_GAUSS_SEIDEL = 1 };

/* ---------------------------------------------------------------------------
 * L0 — topology preparation (mesh_prep.rs)
 * ------------------------------------------------------------------------- */

/* ClothMeshData (mesh_prep.rs:103-131).  Opaque; read fields with
 * xpbd_mesh_count / xpbd_mesh_read.  Constraint rows are in colour-sorted
 * order and the CSR slices are cursor-scattered in that order, as
 * finalize_cloth_mesh builds them (mesh_prep.rs:584-635). */
typedef struct xpbd_mesh xpbd_mesh;

enum {
    XPBD_MESH_POSITIONS = 0,               /* f32 x 3N  initial state           */
    XPBD_MESH_REST_POSITIONS = 1,          /* f32 x 3N                          */
    XPBD_MESH_NORMALS = 2,                 /* f32 x 3N                          */
    XPBD_MESH_UVS = 3,                     /* f32 x 2N                          */
    XPBD_MESH_INDICES = 4,                 /* u32 x 3T                          */
    XPBD_MESH_TRI_EDGE_NEIGHBORS = 5,      /* u32 x 3T (reference builders)     */
    XPBD_MESH_INV_MASS = 6,                /* f32 x N                           */
    XPBD_MESH_NEIGHBOR_OTHER = 7,          /* u32 x 2E                          */
    XPBD_MESH_NEIGHBOR_REST_LEN = 8,       /* f32 x 2E                          */
    XPBD_MESH_NEIGHBOR_COMPLIANCE = 9,     /* f32 x 2E                          */
    XPBD_MESH_NEIGHBOR_CONSTRAINT_ID = 10, /* u32 x 2E                          */
    XPBD_MESH_NEIGHBOR_OFFSETS = 11,       /* u32 x (N+1)                       */
    XPBD_MESH_CONSTRAINT_I = 12,           /* u32 x E                           */
    XPBD_MESH_CONSTRAINT_J = 13,           /* u32 x E                           */
    XPBD_MESH_CONSTRAINT_REST = 14,        /* f32 x E                           */
    XPBD_MESH_CONSTRAINT_COMPLIANCE = 15,  /* f32 x E                           */
    XPBD_MESH_BATCH_OFFSETS = 16,          /* u32 x (B+1) colour batch starts   */
    XPBD_MESH_GROUND = 17,                 /* u32 x N  1 = ground plane y=0     */
    XPBD_MESH_CLOTH_OFFSETS = 18,          /* u32 x (C+1) first vertex of cloth */
    XPBD_MESH_CLOTH_COLS = 19,             /* u32 x C grid columns per cloth
                                              (empty for meshes that are not grids) */
    XPBD_MESH_NUM_FIELDS = 20
};

/* grid_cloth_hanging (mesh_prep.rs:363-406). */
XPBD_API int xpbd_mesh_grid_cloth_hanging(uint32_t quad_cols, uint32_t quad_rows, float cell_size,
                                          xpbd_mesh** out);
/* parse_welded_obj (mesh_prep.rs:223-357).  text need not be NUL-terminated. */
XPBD_API int xpbd_mesh_parse_welded_obj(const char* text, size_t len, xpbd_mesh** out);

/* Benchmark sheets (no reference counterpart: the reference has no seeded
 * sheet builder).  Axial stretch, both quad diagonals as shear, distance-2
 * axial bend; uniform inverse mass; explicit pins; seeded jitter. */
enum { XPBD_LAYOUT_HANGING_XY = 0, XPBD_LAYOUT_HORIZONTAL_XZ = 1 };
enum { XPBD_PIN_NONE = 0, XPBD_PIN_TWO_CORNERS = 1, XPBD_PIN_TOP_ROW = 2, XPBD_PIN_CENTER = 3 };
typedef struct xpbd_sheet_desc {
    uint32_t cols, rows;      /* vertices per row / column (>= 2)          */
    float spacing;            /* rest grid spacing [m]                     */
    uint32_t layout;          /* XPBD_LAYOUT_*                             */
    float height;             /* base y of the sheet [m]                   */
    float angle;              /* rotation about +Y, horizontal layout      */
    uint32_t pins;            /* XPBD_PIN_*                                */
    uint32_t ground;          /* 1: ground plane y = 0 for this cloth      */
    float inv_mass;           /* inverse mass of free vertices [1/kg]      */
    float stretch_compliance, shear_compliance, bend_compliance;
    float jitter;             /* uniform out-of-plane jitter amplitude [m] */
    uint64_t seed;            /* SplitMix64 seed for the jitter            */
} xpbd_sheet_desc;
XPBD_API int xpbd_mesh_sheet(const xpbd_sheet_desc* desc, xpbd_mesh** out);
/* Many independent cloths in one mesh (block-diagonal constraint graph). */
XPBD_API int xpbd_mesh_batch(const xpbd_sheet_desc* descs, uint32_t count, xpbd_mesh** out);

/* A caller's own topology: the arrays of ClothSimConfig (cloth_compute.rs:
 * 142-179) as to_sim_config copies them out of a ClothMeshData
 * (cloth_compute.rs:1599-1700; mesh_prep.rs:103-131).  Host arrays; the
 * library copies them and validates every one.
 *
 * Constraint rows: with constraint_batch_offsets, the rows are taken in the
 * order given, and each batch must be vertex-disjoint (the coloured
 * Gauss-Seidel sweep relies on it; the reference's colouring guarantees it,
 * mesh_prep.rs:153-221).  With constraint_batch_offsets NULL, the rows are
 * coloured and colour-sorted here exactly as finalize_cloth_mesh does
 * (greedy colouring, mesh_prep.rs:153-221), so any foreign constraint list
 * can be given.
 *
 * Adjacency: neighbor_offsets + neighbor_packed are the Jacobi build's
 * fields (pack_neighbor_gpu, cloth_jacobi.rs:95-111: vec4 per directed entry
 * = (other particle, rest length, compliance, constraint row), indices as
 * exact f32 integers).  Given, they are checked against the rows (every row
 * once at each endpoint, equal rest and compliance) and kept in their order,
 * which is the Jacobi gather's summation order (xpbd_cpu.rs:187-226); they
 * then need constraint_batch_offsets.  NULL, they are built by cursor
 * scatter over the rows (mesh_prep.rs:588-635).
 *
 * A mesh from arrays carries no grid columns: the grid-tile Jacobi layout
 * and the learned warm start (both grid-only) are not available for it.
 * Errors: XPBD_E_INVALID, with the reason in xpbd_thread_last_error(). */
typedef struct xpbd_mesh_desc {
    uint32_t num_particles;                   /* N, 1 .. 2^29                          */
    uint32_t num_tris;                        /* T                                     */
    uint32_t num_distance_constraints;        /* E                                     */
    uint32_t constraint_batch_count;          /* B (with constraint_batch_offsets)     */
    const float* initial_pos;                 /* vec4 x N (w ignored), finite; required */
    const float* rest_pos;                    /* vec4 x N, finite; NULL = initial_pos  */
    const float* initial_normals;             /* vec4 x N; NULL = +Y                   */
    const float* inv_mass;                    /* f32 x N, finite, >= 0 (0 = pinned); required */
    const uint32_t* constraint_i;             /* u32 x E, < N                          */
    const uint32_t* constraint_j;             /* u32 x E, < N, != constraint_i         */
    const float* constraint_rest_len;         /* f32 x E, finite, >= 0                 */
    const float* constraint_compliance;       /* f32 x E, finite, >= 0                 */
    const uint32_t* constraint_batch_offsets; /* u32 x (B+1) or NULL                   */
    const uint32_t* tri_indices;              /* u32 x 3T, < N (render normals)        */
    const uint32_t* neighbor_offsets;         /* u32 x (N+1) or NULL                   */
    const float* neighbor_packed;             /* vec4 x 2E or NULL                     */
    const uint32_t* ground;                   /* u32 x N or NULL: nonzero = ground
                                                 plane y = 0 for that particle (no
                                                 reference counterpart)               */
    uint32_t device_prepare;                  /* 1: the colouring, colour sort and
                                                 adjacency scatter run on the GPU
                                                 (constraint_batch_offsets and the
                                                 adjacency NULL; the same results as
                                                 the host; at most 64 colours, else
                                                 XPBD_E_UNSUPPORTED; no device:
                                                 XPBD_E_NO_DEVICE).  Rows of different
                                                 cloths are coloured in parallel.     */
    uint32_t num_cloths;                      /* C; 0 or 1 = one cloth                 */
    const uint32_t* cloth_offsets;            /* u32 x (C+1) first particle of each
                                                 cloth (no reference counterpart: many
                                                 cloths in one handle collide only
                                                 within themselves); rows may not join
                                                 two cloths; NULL = one cloth          */
} xpbd_mesh_desc;
XPBD_API int xpbd_mesh_from_arrays(const xpbd_mesh_desc* desc, xpbd_mesh** out);
/* Why the last failed call without a handle on this thread
 * (xpbd_mesh_from_arrays, xpbd_residual_noalloc) failed; "" if it succeeded. */
XPBD_API const char* xpbd_thread_last_error(void);

XPBD_API uint64_t xpbd_mesh_count(const xpbd_mesh* mesh, int field);
XPBD_API int xpbd_mesh_read(const xpbd_mesh* mesh, int field, void* dst, uint64_t dst_bytes);
XPBD_API void xpbd_mesh_destroy(xpbd_mesh* mesh);

/* derive_collision_grid (mesh_prep.rs:16-85); rest is vec4 x n. */
typedef struct xpbd_coll_grid {
    float origin[3];
    float inv_cell;
    uint32_t dims[3];
    uint32_t num_cells;
    uint32_t radix_digits;
} xpbd_coll_grid;
XPBD_API int xpbd_derive_collision_grid(const float* rest_vec4, uint32_t n, float thickness,
                                        xpbd_coll_grid* out);

/* ---------------------------------------------------------------------------
 * L3/L4 — ClothSimConfig / ClothSimUniforms / ClothComputePlugin
 * ------------------------------------------------------------------------- */

/* ClothSimConfig scalar fields (cloth_compute.rs:142-179).  The topology
 * arrays of the reference config come from the xpbd_mesh. */
typedef struct xpbd_sim_config {
    uint32_t solve_substeps;              /* S  (default SUBSTEPS = 36)      */
    uint32_t solve_inner_iterations;      /* I  (default INNER_ITERS = 22)   */
    uint32_t collision_every_n_substeps;  /* collision when si % N == N-1    */
    uint32_t solver;                      /* XPBD_SOLVER_*                   */
    float collision_cell;                 /* hash cell side [m]; 0 = thickness */
    uint32_t use_graph;                   /* 1: replay frames from a hipGraph */
    uint32_t autotune;                    /* 1 (default): time launch shapes at create
                                             (Jacobi block size and order, coloured GS
                                             grid size and launch form); same results
                                             either way */
    uint32_t jacobi_block;                /* Jacobi workgroup size 64..512; 0 = 256 or autotuned */
    uint32_t fused_predict;               /* predict inside the first iteration:
                                             0 (default) where grid tiles run,
                                             1 always, 2 never (same results) */
    uint32_t plain_adjacency;             /* 1: 12-byte adjacency entries even when the
                                             4-byte packed form fits (read at create) */
    uint32_t fast_division;               /* 1 (default): in the Jacobi gather, the
                                             shorter correctly rounded division
                                             (Markstein step) and square root
                                             (Tuckerman test) wherever their range
                                             conditions are proven (static inputs on
                                             the host, iterate coordinates on the
                                             device); the compiler's general
                                             operations elsewhere.  Same results. */
    uint32_t xcd_remap;                   /* Jacobi block order: 0 = default/autotuned,
                                             1 = hardware order, 2 = XCD-contiguous tiles */
    uint32_t persistent_jacobi;           /* 1: predict + all I Jacobi iterations of a
                                             substep in one co-resident launch with grid
                                             barriers; 2: the same in one launch on
                                             LDS-resident grid tiles (one grid cloth of
                                             64k columns: each tile stays in LDS and only
                                             a 2-particle ring goes between tiles); 0
                                             (default): one launch per iteration, or 2
                                             where the autotuner measures it faster.
                                             Same results */
    uint32_t grid_tiles;                  /* 1 (default): a single grid cloth whose
                                             column count is a multiple of 64 runs the
                                             Jacobi gather on 64x8 tiles staged in LDS
                                             with a 2-particle halo, per entry or with
                                             each constraint evaluated once per tile,
                                             whichever the autotuner measures faster
                                             (per entry without autotuning); 2: per
                                             entry; 3: once per tile; 0: no tiles.
                                             Tiles are chosen at create; all forms
                                             give the same results */
    uint32_t gs_launch;                   /* coloured GS launch form: 0 (default) =
                                             autotuned (else 1), 1 = one persistent
                                             launch per substep with grid barriers
                                             between colours, 2 = one launch per colour
                                             and iteration, 3 = one persistent launch
                                             per substep in which tiles of particles
                                             wait only for their neighbour tiles'
                                             previous colour (per-tile flags).  Same
                                             results */
} xpbd_sim_config;

/* ClothSimUniforms (cloth_compute.rs:182-202). */
typedef struct xpbd_sim_uniforms {
    float dt, inv_dt;
    uint32_t num_particles, num_tris;
    float jacobi_omega;
    uint32_t inner_iterations;
    float thickness, coll_scale;
    float gravity[4];
    float grab_target[4];
    int32_t grab_idx;
    uint32_t grab_active;
    float grab_stiffness;
    float linear_drag_per_sec;
    uint32_t constraint_batch_idx;
} xpbd_sim_uniforms;

/* ClothSimUniforms::default (cloth_compute.rs:204-243) for `solver`, with
 * dt = (1/60)/substeps as sync_cloth_solve_budget_to_uniforms sets it
 * (cloth_compute.rs:419-441). */
XPBD_API void xpbd_sim_uniforms_default(uint32_t solver, uint32_t substeps, xpbd_sim_uniforms* out);
/* ClothMeshData::to_sim_config defaults (cloth_compute.rs:1675-1678). */
XPBD_API void xpbd_sim_config_default(xpbd_sim_config* out);

/* One simulated cloth (or batch): the device state the reference keeps in
 * ClothSimBuffersJacobi / ClothSimBuffers. */
typedef struct xpbd_cloth xpbd_cloth;

/* ClothMeshData::to_sim_config + init_cloth_sim (cloth_compute.rs:1599,571;
 * cloth_jacobi.rs:213): copies the topology to the device and uploads the
 * initial state.  hip_stream: a hipStream_t to order work on, or NULL for a
 * library-owned stream. */
XPBD_API int xpbd_create(const xpbd_mesh* mesh, const xpbd_sim_config* cfg, const xpbd_sim_uniforms* u,
                         void* hip_stream, xpbd_cloth** out);
XPBD_API void xpbd_destroy(xpbd_cloth* h);
XPBD_API const char* xpbd_last_error(const xpbd_cloth* h);
/* Extracted ClothSimUniforms / ClothSimConfig changes (ExtractResourcePlugin,
 * cloth_compute.rs:458-462). */
XPBD_API int xpbd_set_uniforms(xpbd_cloth* h, const xpbd_sim_uniforms* u);
XPBD_API int xpbd_set_config(xpbd_cloth* h, const xpbd_sim_config* cfg);

/* run_cloth_sim (cloth_compute.rs:1197-1269; cloth_jacobi.rs:675-916):
 * enqueue `frames` frames of S substeps (asynchronous). */
XPBD_API int xpbd_step_frames(xpbd_cloth* h, uint32_t frames);
/* `count` substeps with an explicit collision flag (test granularity:
 * gpu_cpu_parity_jacobi.rs:635-764 runs one substep). */
XPBD_API int xpbd_step_substeps(xpbd_cloth* h, uint32_t count, uint32_t collision);
XPBD_API int xpbd_sync(xpbd_cloth* h);
XPBD_API void* xpbd_stream(xpbd_cloth* h);

/* ClothSimControl::reinit_serial (cloth_compute.rs:509-569,
 * cloth_jacobi.rs:918-944): state back to the initial pose, zero velocity. */
XPBD_API int xpbd_reset(xpbd_cloth* h);
/* State read-back / upload in the reference buffer layout: vec4<f32> per
 * particle, w = 0 (cloth_jacobi.rs:238-273).  Host pointers; synchronous. */
XPBD_API int xpbd_read_state(xpbd_cloth* h, float* pos_vec4, float* vel_vec4);
XPBD_API int xpbd_write_state(xpbd_cloth* h, const float* pos_vec4, const float* vel_vec4);

/* Render publish: finalize_normals (cloth_sim_jacobi.wgsl:527-547) writes
 * render_positions / render_normals as array<vec4<f32>>, w = 0, indexed by
 * particle id (cloth_material.rs:16-19, cloth_vertex.wgsl:21-22).  dst_* are
 * DEVICE pointers (the ClothMatExt storage buffers); either may be NULL. */
XPBD_API int xpbd_publish(xpbd_cloth* h, void* dst_pos_vec4, void* dst_nrm_vec4);
/* Bind the render buffers so every frame ends with a publish, as the
 * reference's normals pass does once per frame (cloth_jacobi.rs:885-913). */
XPBD_API int xpbd_bind_render_buffers(xpbd_cloth* h, void* dst_pos_vec4, void* dst_nrm_vec4);

/* Zero-copy render buffers: import the renderer's own position and normal
 * buffers and bind them as xpbd_bind_render_buffers does, so each frame's
 * publish writes straight into them.  Each buffer comes as a POSIX file
 * descriptor of exported device memory (Vulkan VK_KHR_external_memory_fd,
 * VK_EXTERNAL_MEMORY_HANDLE_TYPE_OPAQUE_FD_BIT), the exported allocation's
 * size, and the buffer's offset in it (16-byte aligned, 16 N bytes needed
 * past it); fd -1 leaves that buffer unbound.  On success the library owns
 * the descriptors; the imports are released by the next bind or import and
 * by xpbd_destroy.  Replaces the wgpu bind group of ClothMatExt
 * {sim_positions, sim_normals} (cloth_material.rs:14-48) that the
 * reference's finalize_normals writes (cloth_sim_jacobi.wgsl:527-547). */
XPBD_API int xpbd_import_render_buffers_fd(xpbd_cloth* h, int pos_fd, uint64_t pos_size, uint64_t pos_offset,
                                          int nrm_fd, uint64_t nrm_size, uint64_t nrm_offset);

/* Relative constraint residual sqrt(sum C^2)/sqrt(sum rest^2) of the current
 * state (no reference counterpart; the metric the learned warm start is
 * measured and gated by).  Synchronous. */
XPBD_API int xpbd_residual(xpbd_cloth* h, double* out);

/* No counterpart in the reference.  xpbd_residual's relative residual,
 * sqrt(sum C^2) / sqrt(sum rest^2) in f64, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid hipMalloc): it takes caller-owned device arrays -- positions
 * vec4 x num_particles, constraint_i / constraint_j u32 and rest lengths f32
 * x num_constraints, as ClothSimConfig holds them (cloth_compute.rs:142-179)
 * -- and writes the result to d_out (one device double), enqueued on
 * hip_stream (NULL: the default stream) of the current device, without
 * synchronising.  It allocates nothing: its block partials live in a fixed
 * 32 KiB device array that each device holds for the life of the process, so
 * calls that overlap on one device (other streams, other threads) race and
 * must be ordered by the caller.  num_constraints above
 * XPBD_RESIDUAL_NOALLOC_MAX_CONSTRAINTS is refused with XPBD_E_INVALID; an
 * endpoint index >= num_particles is not read and makes the result NaN. */
#define XPBD_RESIDUAL_NOALLOC_MAX_CONSTRAINTS 0x80000000u
XPBD_API int xpbd_residual_noalloc(const void* d_pos_vec4, uint32_t num_particles, const void* d_constraint_i,
                                   const void* d_constraint_j, const void* d_constraint_rest_len,
                                   uint32_t num_constraints, void* hip_stream, void* d_out);

/* Measurement: average duration [ms] of one launch of kernel `which`
 * (XPBD_KERNEL_*) on the handle's stream, over `reps` launches bracketed by
 * HIP events.  The kernel runs on scratch buffers; the state is untouched.
 * COPY is a plain 16N-byte read + 16N-byte write of the position buffer (the
 * calibration for hardware byte counters); JACOBI_PERSISTENT is one substep's
 * predict + I iterations in the persistent kernel. */
enum { XPBD_KERNEL_JACOBI_ITER = 0, XPBD_KERNEL_COLLISION_SWEEP = 1, XPBD_KERNEL_RADIX_PASS = 2,
       XPBD_KERNEL_GS_SUBSTEP = 3, XPBD_KERNEL_COPY = 4, XPBD_KERNEL_JACOBI_PERSISTENT = 5 };
XPBD_API int xpbd_time_kernel(xpbd_cloth* h, int which, uint32_t reps, float* avg_ms);

/* Measurement on the live state: advances the simulation by `frames` frames,
 * enqueued eagerly (no graph) with a pair of HIP events around every launch
 * of kind `which` (JACOBI_ITER, COLLISION_SWEEP, GS_SUBSTEP or
 * JACOBI_PERSISTENT, whichever the configuration runs), and returns their
 * mean duration [ms] and their number.  Synchronous. */
XPBD_API int xpbd_profile_frames(xpbd_cloth* h, uint32_t frames, int which, float* avg_ms, uint32_t* launches);

/* Debug/test hook for the collision broad phase: computes the cell grids and
 * the stable (key, id) radix sort of the CURRENT positions on the device and
 * copies them back.  grid_out: 12 words per cloth — lo[3], hi[3], cell,
 * inv_cell (floats), dims[3], key base (u32). */
XPBD_API int xpbd_debug_collision_keys(xpbd_cloth* h, float cell_size, uint32_t* grid_out,
                                       uint32_t* sorted_keys, uint32_t* sorted_ids);

/* Test hook (no reference counterpart): sets the handle's device error word,
 * as an expired grid-barrier wait does.  Every later barrier wait of a
 * persistent launch (coloured Gauss-Seidel, persistent Jacobi) then returns at
 * once, and the next xpbd_sync reports XPBD_E_STATE and clears the word. */
XPBD_API int xpbd_debug_raise_device_error(xpbd_cloth* h);

/* Test/tuning hook (no reference counterpart): the per-tile-flag Gauss-Seidel
 * plan (gs_launch = 3).  With tile > 0, re-plans with tiles of `tile`
 * consecutive particles (at least 64) on a grid of `blocks` blocks (0: as many
 * as there are tiles, within what can be resident at once); all plans give
 * the same results.  out4 = {particles per tile, tiles, blocks, most
 * dependencies of one tile}, zeros when the mesh has no plan.
 * XPBD_E_INVALID (plan kept) when this tile size has none.  Makes the flag
 * form run index-range tiles. */
XPBD_API int xpbd_debug_gs_tiles(xpbd_cloth* h, uint32_t tile, uint32_t blocks, uint32_t* out4);

/* Test/tuning hook (no reference counterpart): the flag form on LDS grid
 * tiles (gs_launch = 3 on one grid cloth whose column count is a multiple of
 * 64: tiles of 64 columns x `rows` rows, one block each, kept in LDS for the
 * whole substep).  With rows > 0, re-plans at that height (at least 8) and
 * makes the flag form use these tiles (xpbd_debug_gs_tiles switches it back
 * to index-range tiles).  out4 = {rows per tile, tiles, threads per block,
 * 1 if the flag form uses them}, zeros when there is no plan.
 * XPBD_E_INVALID (plan kept) when this height has none. */
XPBD_API int xpbd_debug_gs_grid(xpbd_cloth* h, uint32_t rows, uint32_t* out4);

/* Test/tuning hook (no reference counterpart): the LDS-resident Jacobi tiles
 * (persistent_jacobi = 2: tiles of 64 columns x `rows` rows of one grid cloth,
 * one block each).  With rows in {16, 32, 64}, re-plans at that height.
 * out4 = {rows per tile, tiles, pairs staged in LDS, 1 if substeps use them},
 * zeros when there is no plan.  XPBD_E_INVALID (plan kept) when this height
 * has none. */
XPBD_API int xpbd_debug_jacobi_resident(xpbd_cloth* h, uint32_t rows, uint32_t* out4);

/* Layout the handle uses (no reference counterpart; for tests and tools):
 * out4 = {adjacency: 0 plain 12-byte entries, 1 packed 4-byte entries,
 * 2 packed entries on LDS grid tiles, 3 grid tiles with each constraint
 * evaluated once per tile; grid-tile columns; grid-tile rows; 1 if the fast
 * correctly rounded division is active}. */
#define XPBD_ADJ_PLAIN 0
#define XPBD_ADJ_PACKED 1
#define XPBD_ADJ_GRID_TILES 2
#define XPBD_ADJ_SHARED_TILES 3
XPBD_API int xpbd_get_layout(xpbd_cloth* h, uint32_t* out4);

/* Device self-test of the shared-reciprocal division used by the Jacobi
 * gather (no reference counterpart; a numerics check of this library):
 * out[0] = significands b (every one, signs +-, exponents exp_lo..exp_hi)
 * whose refined reciprocal differs from the correctly rounded 1/b,
 * out[1] = the first such b (bits); out[2] = mismatches of the quotient
 * against '/' over `pairs` pseudo-random pairs, out[3..4] = first (a, b). */
XPBD_API int xpbd_selftest_division(int exp_lo, int exp_hi, uint64_t pairs, uint64_t seed, uint32_t* out5);
/* Device self-test of the fast square root: every positive normal float with
 * exponent in [exp_lo, exp_hi] (and 0) against the correctly rounded sqrtf.
 * out2 = {mismatches, first mismatching input (bits)}. */
XPBD_API int xpbd_selftest_sqrt(int exp_lo, int exp_hi, uint32_t* out2);

/* ---------------------------------------------------------------------------
 * Learned warm start (no reference counterpart).  A per-vertex MLP
 * (30 -> 64 -> 64 -> 3, GELU; BF16 weights and activations on MFMA, FP32
 * accumulate) predicts the converged constraint correction of a substep from
 * the predicted positions; the substep then runs `warm_iterations` Jacobi
 * iterations from that start, and falls back to the plain substep (predict +
 * I iterations) when the relative residual of the result exceeds the
 * threshold.  Features (per vertex, 30): the 8 grid neighbours' rest-relative
 * displacements / spacing (24), velocity * dt / spacing (3), scaled log10
 * stretch / shear / bend compliance (3).  Grid meshes only (cloth columns).
 * ------------------------------------------------------------------------- */
typedef struct xpbd_mlp xpbd_mlp;
#define XPBD_MLP_PARAMS 7312   /* f32 master parameters, layout in mlp.hpp */

/* Seeded He-style initialisation; the network lives on the current device. */
XPBD_API int xpbd_mlp_create(uint64_t seed, xpbd_mlp** out);
XPBD_API void xpbd_mlp_destroy(xpbd_mlp* m);
XPBD_API const char* xpbd_mlp_last_error(const xpbd_mlp* m);
/* Master parameters (host, XPBD_MLP_PARAMS floats) and the output scale
 * (model output = correction / spacing * out_scale).  set resets Adam. */
XPBD_API int xpbd_mlp_get_params(xpbd_mlp* m, float* params, float* out_scale);
XPBD_API int xpbd_mlp_set_params(xpbd_mlp* m, const float* params, float out_scale);
/* Forward of `rows` feature rows (device, rows x 32 bf16) -> device rows x 3 f32. */
XPBD_API int xpbd_mlp_forward(xpbd_mlp* m, const void* d_features, uint32_t rows, void* d_out);
/* Gradient of the mean squared error against device targets (rows x 3 f32,
 * correction / spacing; scaled by out_scale inside) -> host grad
 * (XPBD_MLP_PARAMS floats) and the loss. */
XPBD_API int xpbd_mlp_gradient(xpbd_mlp* m, const void* d_features, const void* d_targets, uint32_t rows,
                               float* grad, float* loss);
/* `steps` Adam steps on consecutive minibatches of `batch` rows (rows in the
 * order given; cosine decay lr -> lr/10); loss_curve: host, `steps` floats or NULL. */
XPBD_API int xpbd_mlp_train(xpbd_mlp* m, const void* d_features, const void* d_targets, uint32_t rows,
                            uint32_t steps, uint32_t batch, float lr, float* loss_curve);
/* Training pairs from the cloth's current state (state unchanged): features
 * of the predicted positions (device, N x 32 bf16) and the targets (device,
 * N x 3 f32) = (result of `iterations` plain Jacobi iterations - predicted) / spacing. */
XPBD_API int xpbd_warm_sample(xpbd_cloth* h, uint32_t iterations, void* d_features, void* d_targets);
/* Features of the predicted positions only. */
XPBD_API int xpbd_warm_features(xpbd_cloth* h, void* d_features);
/* Attach (m != NULL) or detach the network.  The network must outlive the
 * attachment; attach again after changing its output scale.  A finite
 * threshold computes the relative residual of every warm substep's result
 * (f32 per constraint; within 0.1% of xpbd_residual's figure) and runs the
 * plain substep (one persistent launch) when it exceeds the threshold or is
 * NaN; an infinite threshold skips the check. */
XPBD_API int xpbd_set_warm_start(xpbd_cloth* h, xpbd_mlp* m, uint32_t warm_iterations, float residual_threshold);
/* Substeps run with the warm start and how many fell back; reset != 0 zeroes them. */
XPBD_API int xpbd_warm_stats(xpbd_cloth* h, uint32_t* substeps, uint32_t* fallbacks, uint32_t reset);

/* Library info. */
XPBD_API int xpbd_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif /* XPBD_H */

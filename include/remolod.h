/*
 * remolod.h -- C ABI of the MI355X-native RemoLOD progressive-LOD path.
 *
 * One context = one octree on one GPU, externally synchronised (one host thread
 * at a time), every call stream-ordered on the context's own HIP stream. No C++
 * types, no HIP types and no exceptions cross this boundary: device buffers are
 * plain pointers obtained from remo_dev_alloc() (or any hipMalloc'd memory).
 *
 * Each entry point names the reference interface it replaces (paths relative to
 * the RemoBench repository).
 */
#ifndef REMOLOD_H
#define REMOLOD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
/* The reference reports init/allocate failure as bool + std::string*
 * (include/remo/ILodPipeline.h:124-127) and calls reportDeadContextAndExit(70)
 * on a sticky device fault (src/cuda/CudaCheck.cpp:8-35). Here every call returns
 * one of these; a sticky fault is REMO_ERR_DEVICE and the context is dead. */
#define REMO_OK              0
#define REMO_DONE            1   /* build_step: every batch of the cloud consumed  */
#define REMO_ERR_INVALID    -1   /* bad argument or shape                          */
#define REMO_ERR_OOM        -2   /* device allocation failed / budget too small    */
#define REMO_ERR_DEVICE     -3   /* sticky device fault; the context is dead       */
#define REMO_ERR_RING       -4   /* ring addressing does not match the context     */
#define REMO_ERR_STATE      -5   /* call out of lifecycle order                    */
#define REMO_ERR_CAPACITY   -6   /* a fixed capacity (spill, backlog, dir) was hit */

/* ---- constants of the path (kernels/remolod/remolod_structures.cuh:12-25) -- */
#define REMO_MAX_POINTS_PER_NODE   50000u
#define REMO_POINTS_PER_CHUNK      1000u
#define REMO_GRID_SIZE             128u
#define REMO_GRID_WORDS            65536u     /* 128^3 bits / 32 */
#define REMO_MAX_DEPTH             20u
#define REMO_DEFAULT_MAX_NODES     200000u
#define REMO_DEFAULT_RING_SLOTS    50u
#define REMO_DEFAULT_SLOT_CAPACITY 1000000u

/* ---- plain data -------------------------------------------------------- */

/* == remo::Point (include/remo/HostDeviceCommon.h:15-18): 16 B AoS,
 *    color = r | g<<8 | b<<16 | a<<24. */
typedef struct remo_point {
    float x, y, z;
    uint32_t color;
} remo_point;

/* == remo::mat4 (include/remo/HostDeviceCommon.h:20-22): row-major rows[r][c]. */
typedef struct remo_mat4 {
    float rows[4][4];
} remo_mat4;

typedef struct remo_vec3f {
    float x, y, z;
} remo_vec3f;

/* == remo::SharedUniforms (include/remo/HostDeviceCommon.h:28-60), same layout. */
typedef struct remo_shared_uniforms {
    float width, height, fovy_rad, time;
    remo_mat4 transform, view, proj;
    remo_mat4 transform_frozen, transform_frozen_inv;
    remo_vec3f box_min, box_max;
    uint64_t frame_counter;
    float lod_pixel_budget, min_node_size, lod_scale;
    int32_t point_size, color_mode;
    float edl_strength;
    int32_t do_update_visibility, use_high_quality_shading, enable_edl;
    int32_t show_bounding_box, show_points, pad0;
} remo_shared_uniforms;

/* CloudMeta subset the path uses (include/remo/PointSource.h:16-32). Points are
 * already translated so the box minimum is 0 (src/io/PointSource.cpp:360-379). */
typedef struct remo_cloud_meta {
    uint64_t num_points;
    double translation[3];
    float box_size[3];
    uint32_t pad0;
} remo_cloud_meta;

/* == remo::BatchView (include/remo/PointSource.h:43-67). Device pointers.
 * Batch B lives in slot B % num_slots at slots + slot * slot_capacity points,
 * with batch_sizes[slot] valid points; *num_batches_uploaded is the producer
 * cursor. */
typedef struct remo_batch_view {
    const remo_point* slots;
    const uint32_t* batch_sizes;
    const uint32_t* num_batches_uploaded;
    uint32_t slot_capacity;
    uint32_t num_slots;
    uint32_t num_batches_total;
    uint32_t pad0;
} remo_batch_view;

/* Sizing and policy of one context: what RemolodPipeline::allocate derives from
 * CloudMeta + DeviceBudget (src/pipelines/RemolodPipeline.cpp:235-325). Zero
 * fields take the reference defaults. */
typedef struct remo_config {
    uint64_t persistent_bytes;      /* octree store (grids + chunks); 0: 48 B/pt, >= 512 MiB,
                                       plus 800 MB with fixed_depth (its store margin is
                                       1 GB instead of 200 MB, remolod_octree.cu:1191-1195) */
    uint32_t max_nodes;             /* node pool, MAX_NODES_CAPACITY (200,000)            */
    uint32_t max_batches_per_step;  /* batches folded per build_step (20)                 */
    float    step_budget_ms;        /* device-time budget per build_step (10); < 0: none   */
    uint32_t slot_capacity;         /* ring slot stride in points (1,000,000)             */
    uint32_t ring_slots;            /* ring depth the batch index wraps at (50)           */
    uint32_t spill_capacity;        /* spilled points per batch (10,000,000)              */
    uint32_t backlog_capacity;      /* new voxels per batch (10,000,000), rounded up to a
                                       multiple of 64: 64 equal segments, one per group of
                                       workgroups; a full segment sets backlog_overflow    */
    uint32_t flags;                 /* REMO_FLAG_*                                        */
    uint32_t closure_batches;       /* Analysis: a leaf untouched for this many batches is
                                       closed (8); the reference's named alternative to a
                                       Morton watermark, plans/03_AccumulatorHook.md:507-509 */
    uint32_t fixed_depth;           /* 0: the iterative expand; 1..8: the analytic-bucketing
                                       arm that pre-splits to this depth
                                       (--remolod-fixed-depth, remolod_octree.cu:464-675) */
} remo_config;

#define REMO_MAX_FIXED_DEPTH 8u

#define REMO_FLAG_NO_ACCUM   1u     /* --remolod-no-accum                       */

/* == Stats (kernels/simlod/HostDeviceInterface.h:51-75) + PipelineStats
 * (include/remo/ILodPipeline.h:88-111) + the DeviceTimeline counter block
 * (include/remo/HostDeviceCommon.h:102-166) + AccumGlobals
 * (include/remo/RemoAccum.h:33-40) + DeviceDiagnostics. */
typedef struct remo_stats {
    uint32_t num_nodes;             /* nodes in the tree: the pool's high-water mark less the
                                       slots of blocks pooled by remo_collapse            */
    uint32_t num_inner, num_leaves, num_nonempty_leaves;
    uint64_t num_points;            /* sum of leaf numPoints      */
    uint64_t num_voxels;            /* sum of inner-node numVoxels */
    uint64_t num_points_processed;
    uint32_t batchlet_index;
    uint32_t num_chunks_points, num_chunks_voxels;
    uint32_t max_points_per_node;
    uint32_t counter_underflows;    /* leaves with counter < numPoints */
    uint32_t max_depth;
    uint64_t num_allocated_chunks, chunk_pool_size;
    uint64_t allocated_bytes_persistent, persistent_capacity;
    uint32_t mem_capacity_reached, node_capacity_reached, node_pool_overflow, alloc_overflow;
    uint32_t spill_overflow, backlog_overflow, dir_overflow;
    uint32_t free_node_blocks;      /* 8-node blocks freed by remo_collapse, taken first by splits */
    /* per-step counters (DeviceTimeline counter block semantics, summed over launches) */
    uint64_t batches, expand_iters, nodes_split, spilled_points, new_voxels;
    /* the same five, cumulative since the last reset */
    uint64_t total_batches, total_expand_iters, total_nodes_split, total_spilled_points, total_new_voxels;
    /* AccumGlobals */
    uint64_t morton_watermark, num_accumulated, sum_leaf_counts;
    uint32_t points_folded, leaves_folded, inner_with_sums;
    uint32_t pooled_grids;          /* grids freed by remo_collapse, taken first by splits */
    /* render diagnostics (DeviceDiagnostics) */
    uint32_t draw_items, draw_list_overflow;
    uint64_t draw_samples;
} remo_stats;

/* Canonical node record for dumps. Pool index order is not part of parity;
 * compare by (level, x, y, z). */
typedef struct remo_node_record {
    uint32_t level, x, y, z;
    int32_t first_child;            /* -1: leaf; else children are first_child + 0..7,
                                       child i = (cx<<2)|(cy<<1)|cz */
    uint32_t counter, num_points, num_voxels, num_voxels_stored;
    int32_t grid;                   /* occupancy-grid number, -1: none */
    uint32_t free_slot;             /* 1: a slot of a block remo_collapse pooled: not in the
                                       tree (its other fields are zero or stale)        */
    uint32_t pad1;
} remo_node_record;

/* Per-node Accumulate record. g/c/count/last_touched_batch/state/score have the
 * meaning of remo::NodeAccum (include/remo/RemoAccum.h:18-30); q1/q2 are the
 * exact sums of q = floor(l * 2^20) - 2^19, l the node-local coordinate in
 * [0,1] (node-centred fixed point, 2^-20 of the node extent). */
typedef struct remo_node_accum {
    int64_t q1[3];                  /* sum qx, qy, qz                    */
    int64_t q2[6];                  /* sum qxqx qxqy qxqz qyqy qyqz qzqz */
    double g[9];                    /* FP64 sums x y z xx xy xz yy yz zz */
    float c[6];                     /* FP32 colour sums r g b rr gg bb   */
    uint32_t count, last_touched_batch, state;
    float score;
    uint64_t morton_max;
} remo_node_accum;

/* Per-node Analysis result. Eigenvalues ascending, of the covariance in node
 * extent units; planarity = (l1 - l0) / l2, linearity = (l2 - l1) / l2,
 * surface variation = l0 / (l0 + l1 + l2). */
typedef struct remo_node_analysis {
    float lambda[3];
    float surface_variation, planarity, linearity;
    float coverage;                 /* projected extent, pixels^2 (clamped to the screen) */
    float score;
    uint64_t n;                     /* points under the node (rolled up) */
    uint32_t state;                 /* 0 open, 1 closed, 2 finalized     */
    uint32_t flags;                 /* bit1: learned score used (bit0 unused, 0)           */
} remo_node_analysis;

#define REMO_ANALYSIS_ANALYTIC 0
#define REMO_ANALYSIS_LEARNED  1

/* The learned score (no reference code: plans/04_NeuralPass.md is empty):
 * 13 -> 64 -> 64 -> 2 MLP with exact GELU, BF16 weights, FP32 accumulation,
 * predicting a node's final (surface variation, planarity). Parameters are FP32
 * W1[64][13], b1[64], W2[64][64], b2[64], W3[2][64], b3[2], in that order. */
#define REMO_MLP_INPUTS   13u
#define REMO_MLP_HIDDEN   64u
#define REMO_MLP_OUTPUTS  2u
#define REMO_MLP_PARAMS   5186u

typedef struct remo_ctx remo_ctx;
typedef struct remo_ring remo_ring;
typedef struct remo_mlp remo_mlp;

/* ---- lifecycle ---------------------------------------------------------- */

/* ILodPipeline::initPrograms + allocate (include/remo/ILodPipeline.h:124-127,
 * src/pipelines/RemolodPipeline.cpp:194-325): code objects are linked ahead of
 * time, so this only allocates every device buffer. */
int remo_create(const remo_config* cfg, const remo_cloud_meta* meta, remo_ctx** out);

/* ILodPipeline::release (src/pipelines/RemolodPipeline.cpp:327-342). */
void remo_destroy(remo_ctx* ctx);

/* ILodPipeline::reset + the reset kernel (kernels/remolod/remolod_reset.cu:23-88,
 * src/pipelines/RemolodPipeline.cpp:340-389): empty tree, root with a zeroed grid,
 * zeroed accumulators. The caller rewinds its source, as the reference does. */
int remo_reset(remo_ctx* ctx);

/* ILodPipeline::build -> kernel_construct (kernels/remolod/remolod_octree.cu:1062-1320):
 * folds batches [batchletIndex, min(+max_batches_per_step, uploaded)) from the ring,
 * stopping early on the step budget or the persistent-memory margin. Returns REMO_OK
 * while batches remain, REMO_DONE once num_batches_total are consumed. */
int remo_build_step(remo_ctx* ctx, const remo_batch_view* view, uint32_t* batches_consumed);

/* runAccumulator -> kernel_accumulate + kernel_accum_verify
 * (src/pipelines/RemolodPipeline.cpp:242-294, kernels/remolod/remolod_accum.cu:47-233). */
int remo_accumulate(remo_ctx* ctx, uint32_t batch_index);

/* Precision comparator (no reference counterpart): recomputes every leaf's sums
 * from its stored points with naive FP32 atomics -- out[node * 10 + k], k = 0 the
 * count, 1..3 sum x y z, 4..9 sum xx xy xz yy yz zz of the node-local coordinates
 * in [0,1] (device rows; inner nodes stay zero) -- so the int64 fixed-point sums
 * can be compared against FP32 atomics and an FP64 reference. */
int remo_accum_fp32_atomics(remo_ctx* ctx, float* out, uint32_t cap, uint32_t* num);

/* Analysis (README.md:80, wiki/02_KernelPipeline.md:205-224; no reference code):
 * roll-up of the exact sums, 3x3 eigen-solve, score = geometric x coverage,
 * work queue. mode: REMO_ANALYSIS_ANALYTIC or REMO_ANALYSIS_LEARNED. */
int remo_analysis(remo_ctx* ctx, const remo_shared_uniforms* u, int mode);

/* Limit of the incremental roll-up (no reference counterpart; a tuning and test
 * knob): at most lds_slots marked nodes keep their sums in LDS, where each adds
 * its partial sums to its marked ancestors at once; more take device memory and
 * a pass per level. 0 restores the default (384); larger values are clamped to
 * it. The results do not depend on the limit. */
int remo_set_rollup_limits(remo_ctx* ctx, uint32_t lds_slots);

/* ILodPipeline::render -> kernel_render selection + shared rasteriser
 * (kernels/remolod/remolod_render.cu:72-181, kernels/shared/remo_draw.cuh:118-167,
 * kernels/shared/remo_framebuffer.cuh:214-248). fb: device u64[width*height],
 * cleared and filled with (float_bits(depth) << 32) | rgba. */
int remo_render(remo_ctx* ctx, const remo_shared_uniforms* u, uint64_t* fb);

/* The `flat` pipeline's Rasterize (kernels/flat/flat_render.cu:9-76): every one
 * of n device points (e.g. a ring that holds the whole cloud) splatted into fb
 * with the same projection and packing as remo_render, no LOD -- the image
 * ground truth a LOD frame is compared against. Ordered on ctx's stream. */
int remo_render_flat(remo_ctx* ctx, const remo_point* points, uint64_t n, const remo_shared_uniforms* u,
                     uint64_t* fb);
/* Eye-dome lighting of a rendered fb in place when u->enable_edl
 * (remoApplyEDL, kernels/shared/remo_framebuffer.cuh:78-122); depth is kept. */
int remo_apply_edl(remo_ctx* ctx, const remo_shared_uniforms* u, uint64_t* fb);

/* Analysis + Rasterize of one frame as one captured hipGraph (no reference
 * counterpart: the reference launches kernel_render alone, eagerly). The graph is
 * captured on the first call and re-captured when the tree depth, the mode, the
 * framebuffer, its size or the launch shape change; the uniforms and the arrival
 * fraction reach its kernels through device memory, so a replay serves a new
 * camera. Results equal remo_analysis followed by remo_render. */
int remo_frame(remo_ctx* ctx, const remo_shared_uniforms* u, int mode, uint64_t* fb);

/* One launch shape of the per-frame kernels and its measured time. */
typedef struct remo_tune_entry {
    uint32_t node_blocks;    /* workgroups of the node-parallel Analysis / selection kernels */
    uint32_t raster_blocks;  /* workgroups of the raster kernel                              */
    uint32_t raster_tile;    /* samples per raster work unit                                 */
    float ms;                /* best of 3 timed runs of Analysis + Rasterize                 */
} remo_tune_entry;

/* Autotuner: times Analysis + Rasterize on the current tree and camera over a
 * grid of launch shapes and keeps the fastest for remo_analysis, remo_render and
 * remo_frame. Writes up to cap entries (NULL: count only) and *num. Results do
 * not depend on the shape. */
int remo_autotune_frame(remo_ctx* ctx, const remo_shared_uniforms* u, int mode, uint64_t* fb,
                        remo_tune_entry* out, uint32_t cap, uint32_t* num);
/* The launch shape in use (ms = 0). */
int remo_get_frame_shape(remo_ctx* ctx, remo_tune_entry* out);

/* ILodPipeline::stats / diagnostics (src/pipelines/RemolodPipeline.cpp:440-470, :603-726).
 * Runs the per-node stats sweep (kernel_construct's closing sweep,
 * remolod_octree.cu:1248-1319) and reads the counters back (synchronous). */
int remo_get_stats(remo_ctx* ctx, remo_stats* out);

/* remo_get_stats in two halves, so that a caller issuing more work after a build
 * step does not drain the stream to read the stats. remo_request_stats queues the
 * stats sweep and the read of the counters behind the work issued so far (no
 * synchronisation); remo_collect_stats waits for that read alone and fills `out`
 * (the draw counters are the last render's, as in remo_get_stats). The reference's
 * build kernel ends in the same sweep and its readStats copies the result
 * (remolod_octree.cu:1248-1319, src/pipelines/RemolodPipeline.cpp:440-470); its
 * plugin reads them once a frame, after the build, as RemolodHipPipeline::build
 * does with this pair (requested after Accumulate, collected after Analysis).
 * Collecting without a request first is REMO_ERR_STATE. */
int remo_request_stats(remo_ctx* ctx);
int remo_collect_stats(remo_ctx* ctx, remo_stats* out);
/* The last render's DeviceDiagnostics only -- draw items, samples drawn, draw
 * list overflow -- as RemolodPipeline::render reads them after kernel_render
 * (src/pipelines/RemolodPipeline.cpp:592-600): one small readback, no sweep. */
int remo_get_render_diagnostics(remo_ctx* ctx, uint32_t* draw_items, uint64_t* draw_samples, uint32_t* overflow);
const char* remo_last_error(const remo_ctx* ctx);
int remo_sync(remo_ctx* ctx);

/* ---- learned score ------------------------------------------------------ */
/* A training object: FP32 master parameters (Glorot-uniform from `seed`, zero
 * biases) with Adam state; batches of at most max_batch samples. */
int remo_mlp_create(uint64_t seed, uint32_t max_batch, remo_mlp** out);
void remo_mlp_destroy(remo_mlp* mlp);
int remo_mlp_get_params(remo_mlp* mlp, float* host, uint32_t n);        /* n = REMO_MLP_PARAMS */
int remo_mlp_set_params(remo_mlp* mlp, const float* host, uint32_t n);  /* resets Adam */
/* Inference on device rows x[n][13] -> y[n][2] with the BF16 weights. */
int remo_mlp_forward(remo_mlp* mlp, const float* x, uint32_t n, float* y);
/* Loss (mean squared error over n x 2 outputs) and its gradient w.r.t. the
 * parameters (host, REMO_MLP_PARAMS floats, may be NULL) without an update. */
int remo_mlp_gradient(remo_mlp* mlp, const float* x, const float* t, uint32_t n, float* loss, float* grad_host);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8, bias-corrected) on device rows
 * x[n][13], targets t[n][2]; *loss is the batch loss before the step. */
int remo_mlp_train_step(remo_mlp* mlp, const float* x, const float* t, uint32_t n, float lr, float* loss);
/* Loads the model into a context: remo_analysis(..., REMO_ANALYSIS_LEARNED) then
 * scores open nodes with it (closed nodes, and predictions above 0.05 surface
 * variation, take the analytic eigen-solve). */
int remo_set_score_model(remo_ctx* ctx, const remo_mlp* mlp);
/* The 13 inputs of every node from the last remo_analysis roll-up, as device
 * rows out[num][13] (training pairs are logged from these). */
int remo_mlp_features(remo_ctx* ctx, float* out, uint32_t cap, uint32_t* num);

/* ---- Refinement: colour filtering ----------------------------------------
 * The reference specifies Refinement without code (README.md:81,
 * wiki/02_KernelPipeline.md:228-242, constraints CLAUDE.md:181-190); this is its
 * colour-filtering step. Every gridded node's voxel list is rewritten in cell
 * order (x fastest, then y, then z), one voxel per occupied cell, at the cell
 * centre voxel sampling uses, coloured with the rounded mean of each RGBA channel
 * ((sum + n/2) / n) over every point of the node's subtree in that cell; the
 * node's voxel counts become the occupied-cell count. Results depend only on the
 * points in the tree, never on arrival order. The cell sums live in a sparse
 * accumulator per grid, one entry per occupied cell, indexed by the cell's rank
 * among the grid's set bits; a node's sums are its leaf children's points plus
 * its inner children's sums, so each pass reads every point once and every
 * child sum once.
 *
 * Incremental: a gridded node is dirty when a leaf below it gained points or its
 * voxel count changed since it was last refined; a call refines the dirty nodes
 * level by level from the deepest, up to `max_levels` levels (0: all) and while
 * the cost model's estimate stays within budget_ms (< 0: no limit, 0: nothing);
 * the last level may be cut to the nodes that fit (at least one node runs when
 * budget_ms > 0). What is left stays dirty for the next call. A node is left alone (nodes_skipped) when its voxel list is
 * incomplete, an inner child could not be refined, or the accumulator store is
 * full; it stays dirty. Cell sums are exact up to 16,843,009 points in one cell
 * (cell_overflow reports more). The store (24 B per entry) starts at about two
 * entries per voxel; when a pass would not fit beside the live sums it is
 * replaced first and every node is recomputed (environment variable
 * REMO_REFINE_STORE_ENTRIES caps its size). Synchronous: returns after the work ran. */
typedef struct remo_refine_stats {
    uint32_t levels_pending;    /* levels holding dirty gridded nodes at the start    */
    uint32_t levels_processed;  /* of those, refined by this call (deepest first)     */
    uint32_t nodes_dirty;       /* dirty gridded nodes at the start                   */
    uint32_t nodes_refined;     /* voxel lists rewritten                              */
    uint32_t nodes_skipped;     /* processed but left alone (see above)               */
    uint32_t complete;          /* 1: no dirty node was left for a later call         */
    uint64_t points_folded;     /* leaf points added into cell sums                   */
    uint64_t child_sums;        /* inner children whose cell sums were folded         */
    uint64_t voxels_written;
    uint32_t missing_cells;     /* points whose cell bit was clear (inconsistent tree) */
    uint32_t empty_cells;       /* set cells without points (written with colour 0)   */
    uint32_t arena_full;        /* 1: the accumulator store ran out (reset next call) */
    uint32_t cell_overflow;     /* 1: a cell held more than 16,843,009 points         */
    float estimate_ms;          /* the cost model's estimate for the work processed,
                                   corrected by recent calls' measured/estimated ratio */
    float device_ms;            /* measured: first to last kernel of the levels       */
} remo_refine_stats;

int remo_refine(remo_ctx* ctx, float budget_ms, uint32_t max_levels, remo_refine_stats* out);

/* Refinement's local deepening (README.md:81, wiki/02_KernelPipeline.md:232-236:
 * "local split and redistribute"; no reference code). Up to max_splits leaves
 * holding at least min_points stored points (and no points counted but not yet
 * stored) are split before they reach 50,000, the densest first (ties by
 * level, x, y, z), and their points are redistributed to the children exactly
 * as a split during Update does: the leaf gets a zeroed grid and voxels from its
 * own points, its chunks are recycled, the children take the points octant by
 * octant. The chosen leaves are a prefix of that order whose points fit both the
 * spill list they move through (spill_capacity) and the voxel backlog
 * (backlog_capacity); `chosen` says how many. The node pool's limit is respected
 * and reported like Update's (node_pool_overflow). Refused when the persistent
 * store is within Update's safety margin. Synchronous. */
typedef struct remo_deepen_stats {
    uint32_t candidates;        /* leaves that qualified                    */
    uint32_t chosen;            /* of those, offered to the split (<= max_splits) */
    uint32_t splits;            /* accepted by the node pool                */
    uint32_t refused;           /* 1: store near capacity, nothing done     */
    uint64_t points_moved;      /* stored points redistributed              */
    uint64_t new_voxels;        /* voxels the split leaves gained           */
} remo_deepen_stats;

int remo_deepen(remo_ctx* ctx, uint32_t max_splits, uint32_t min_points, remo_deepen_stats* out);

/* Refinement's collapsing (README.md:81 and wiki/02_KernelPipeline.md:238: "merge
 * a node back up", which "needs a grid pool first"; no reference code). Up to
 * max_collapses inner nodes below the root whose eight children are leaves with
 * every counted point stored, holding at most min(max_points, 50,000) points
 * together, are merged into one leaf each, the sparsest first (ties by level,
 * x, y, z); the chosen are a prefix of that order whose points fit the spill list
 * they move through (spill_capacity). A merged node takes its children's points
 * (its chunks from the chunk queue, as in Update's insertion), returns its
 * children's point chunks to that queue, its occupancy grid to the grid pool and
 * the grid's voxel chunks to the chunk queue, and drops its voxels (a leaf draws
 * its points); its children's 8-node block goes to the node-block pool. Splits
 * (Update, the fixed-depth arm, remo_deepen) take pooled blocks and grids before
 * they grow the node pool or the persistent store, so a collapsed region grows
 * back without new memory. Accumulate folds a merged leaf from scratch; the next
 * Analysis rolls up every node. Synchronous. */
typedef struct remo_collapse_stats {
    uint32_t candidates;        /* inner nodes that qualified                   */
    uint32_t collapsed;         /* merged (a prefix of the sparsest, <= max_collapses) */
    uint64_t points_moved;      /* stored points moved into the merged leaves   */
    uint32_t grids_pooled;      /* grids returned to the pool                   */
    uint32_t voxels_dropped;    /* voxels the merged nodes held                 */
} remo_collapse_stats;

int remo_collapse(remo_ctx* ctx, uint32_t max_collapses, uint32_t max_points, remo_collapse_stats* out);

/* remo_deepen for callers that cannot allocate device memory (code under graph
 * capture, serving loops, harnesses that forbid hipMalloc); no counterpart in
 * the reference. Same choice, splits and results as remo_deepen, which allocates
 * its candidate buffer on its first call; this variant allocates nothing and
 * keeps the candidate records in a fixed device array of the library (4.8 MB per
 * GPU, held until the process ends): trees of more than 200,000 nodes are
 * refused with REMO_ERR_CAPACITY. Every call on one GPU shares that array, so
 * calls that overlap on one GPU (contexts used from several threads) race. */
int remo_deepen_noalloc(remo_ctx* ctx, uint32_t max_splits, uint32_t min_points, remo_deepen_stats* out);

/* ---- timing scopes (GpuProfiler / GpuScope, include/remo/GpuProfiler.h:62-158) ----
 * When enabled, every phase launch is bracketed by a pair of HIP events on the
 * context's stream; remo_profile_read harvests finished pairs into per-scope
 * totals. Scope names follow the reference's (src/pipelines/RemolodPipeline.cpp:474-479):
 * remolod.construct.{count,split,voxelSampling,allocChunks,insertPoints,insertVoxels},
 * remolod.accumulate, remolod.analysis, remolod.render, remolod.frame (remo_frame)
 * remolod.refine (remo_refine) and flat.render (remo_render_flat,
 * src/pipelines/FlatPipeline.cpp:89-92). */
typedef struct remo_profile_entry {
    char name[48];
    double total_ms;
    uint64_t launches;
} remo_profile_entry;

/* scope_mask: bit i times scope i in the order listed above (count = bit 0 ...
 * render = bit 8, frame = bit 9, refine = bit 10, flat.render = bit 11); 0 disables,
 * 0xFFFFFFFF times all. */
int remo_profile_enable(remo_ctx* ctx, uint32_t scope_mask);
int remo_profile_read(remo_ctx* ctx, remo_profile_entry* out, uint32_t cap, uint32_t* num);
int remo_profile_clear(remo_ctx* ctx);

/* ---- canonical dumps (parity; host buffers) ----------------------------- */
int remo_dump_nodes(remo_ctx* ctx, remo_node_record* out, uint32_t cap, uint32_t* num);
/* Points of node `node` in storage order (leaf points, or voxels when voxels != 0). */
int remo_dump_node_samples(remo_ctx* ctx, uint32_t node, int voxels,
                           remo_point* out, uint32_t cap, uint32_t* num);
/* The 65,536-word occupancy grid of node `node` (REMO_ERR_INVALID if it has none). */
int remo_dump_grid(remo_ctx* ctx, uint32_t node, uint32_t* out_words);
int remo_dump_accums(remo_ctx* ctx, remo_node_accum* out, uint32_t cap, uint32_t* num);
int remo_dump_analysis(remo_ctx* ctx, remo_node_analysis* out, uint32_t cap, uint32_t* num);

/* The exact sums the last Analysis rolled up for every node (no reference
 * counterpart: the reference specifies the roll-up, README.md:80,
 * wiki/02_KernelPipeline.md:205-224, without code). Coordinates are integers in
 * units U = 2^-(20+D) of the root extent, D the deepest level, centred on the
 * node's own centre; s1 = sum of q, s2 = sum of q q^T (xx xy xz yy yz zz), each a
 * two's-complement 128-bit integer stored as [0] low and [1] high 64 bits;
 * e = 20 + D - level (the node extent is 2^e U); closed: the closure flag. */
typedef struct remo_node_rollup {
    int64_t s1[3][2];
    int64_t s2[6][2];
    uint64_t n;
    uint32_t e;
    uint32_t closed;
} remo_node_rollup;
int remo_dump_rollup(remo_ctx* ctx, remo_node_rollup* out, uint32_t cap, uint32_t* num);

/* ---- the input ring (PointSource, include/remo/PointSource.h:69-115) ---- */
int remo_ring_create(uint32_t slot_capacity, uint32_t num_slots, remo_ring** out);
void remo_ring_destroy(remo_ring* ring);
/* Synchronous upload of batch `batch_index` into slot batch_index % num_slots;
 * advances the uploaded cursor to batch_index + 1 if it was below. */
int remo_ring_upload(remo_ring* ring, uint32_t batch_index, const remo_point* host, uint32_t n);
int remo_ring_view(const remo_ring* ring, uint32_t num_batches_total, remo_batch_view* out);
/* PointSource::rewind: uploaded cursor back to 0 (waits for pushes in flight). */
int remo_ring_rewind(remo_ring* ring);
/* Asynchronous upload on the ring's own copy stream, overlapping the consumer's
 * kernels (the reference's ring has no second stream, README.md:185-191). `host`
 * should be pinned (remo_host_alloc) for the copy to be truly asynchronous and must
 * stay valid until remo_ring_sync. When the slot still holds an earlier batch, the
 * copy first waits for `consumer`'s last build_step (may be NULL if the caller
 * orders it otherwise). The producer keeps batch_index < consumed + num_slots
 * (PointSource's backpressure contract, include/remo/PointSource.h:95-103). */
int remo_ring_push(remo_ring* ring, uint32_t batch_index, const remo_point* host, uint32_t n,
                   remo_ctx* consumer);
/* Makes `consumer`'s stream wait for the pushes of batches
 * [first_batch, first_batch + num_batches): call before remo_build_step. */
int remo_ring_order_before(remo_ring* ring, remo_ctx* consumer, uint32_t first_batch, uint32_t num_batches);
/* Waits for every push issued so far. */
int remo_ring_sync(remo_ring* ring);

/* ---- device memory plumbing -------------------------------------------- */
int remo_dev_alloc(size_t bytes, void** out);
int remo_dev_free(void* p);
/* Pinned (page-locked) host memory, so ring pushes from it are asynchronous DMA. */
int remo_host_alloc(size_t bytes, void** out);
int remo_host_free(void* p);
int remo_memcpy_h2d(void* dst, const void* src, size_t bytes);
int remo_memcpy_d2h(void* dst, const void* src, size_t bytes);
/* Device-to-device copy by a plain streaming kernel (remo_copy_kernel): the
 * known-size reference that HBM byte counters are calibrated against. bytes must
 * be a multiple of 16; synchronous. */
int remo_dev_copy(void* dst, const void* src, size_t bytes);
/* Device-scope atomicOr throughput on `bytes` of device memory (a power of two,
 * >= 4 KiB) by a plain microbenchmark kernel (remo_atomic_kernel): the peak the
 * Update kernels' atomic rates are reported against. mode 0: scattered
 * addresses, returning; 1: scattered, non-returning; 2: all 64 lanes of a wave
 * on one word. 2^20 lanes issue per_lane atomics each; synchronous. */
int remo_atomic_rate(void* words, size_t bytes, uint32_t mode, uint32_t per_lane, double* atomics_per_s);
int remo_device_count(int* out);
/* Free and total device memory: CudaContext::freeMemory / totalMemory, which
 * App::computeBudget turns into DeviceBudget (src/shell/App.cpp:97-114). */
int remo_device_mem_info(size_t* free_bytes, size_t* total_bytes);

/* ---- synthetic scenes (makeSyntheticSource's role, src/io/PointSource.cpp:535-584) ----
 * Seed-generated lidar-like scene, scan-line ordered, translated to box min 0. */
typedef struct remo_scene_spec {
    double extent;                  /* ground square side, metres (40)          */
    uint32_t num_boxes;             /* axis-aligned boxes, 2-10 m sides (20)    */
    uint32_t num_cylinders;         /* r 0.15 m, h 6 m (50)                     */
    double cluster_fraction;        /* share of points in Gaussian clusters (0) */
    double cluster_volume_fraction; /* volume the clusters cover (0.05)         */
    double cluster_sigma;           /* 0.5 m                                    */
    double ground_sigma;            /* 0.02 m                                   */
    double scanline;                /* scan-line pitch, metres (0.05)           */
    uint32_t shuffle_window;        /* 4096                                     */
    uint32_t threads;               /* 0: all hardware threads                  */
} remo_scene_spec;

void remo_scene_default(remo_scene_spec* spec);
int remo_scene_generate(const remo_scene_spec* spec, uint64_t seed, uint64_t num_points,
                        remo_point* out, remo_cloud_meta* meta);

#ifdef __cplusplus
}
#endif

#endif

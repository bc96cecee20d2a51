// Host side of the C ABI: context lifecycle, the per-step batch loop, dumps,
// the input ring and device-memory plumbing.
//
// Mirrors RemolodPipeline (src/pipelines/RemolodPipeline.cpp): allocate() sizes
// every buffer up front, reset() rebuilds the root, build() folds up to 20
// batches or a device-time budget per call, runAccumulator() folds new points.
// Everything is ordered on one HIP stream owned by the context.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "remolod_internal.h"
#include "remolod_analysis.h"
#include "remolod_mlp.h"
#include "remolod_refine.h"

using namespace remo_hip;

namespace {

enum ScopeId : int {
    kScopeCount, kScopeSplit, kScopeVoxel, kScopeAlloc, kScopeInsertPoints, kScopeInsertVoxels,
    kScopeAccumulate, kScopeAnalysis, kScopeRender, kScopeFrame, kScopeRefine, kScopeFlat, kNumScopes
};
const char* const kScopeNames[kNumScopes] = {
    "remolod.construct.count", "remolod.construct.split", "remolod.construct.voxelSampling",
    "remolod.construct.allocChunks", "remolod.construct.insertPoints",
    "remolod.construct.insertVoxels", "remolod.accumulate", "remolod.analysis", "remolod.render",
    "remolod.frame", "remolod.refine", "flat.render",
};

// HIP-event pairs per launch, harvested after the stream has passed them.
struct Profiler {
    uint32_t mask = 0;  // scopes being timed
    std::vector<hipEvent_t> pool;
    struct Pair {
        int scope;
        hipEvent_t a, b;
    };
    std::vector<Pair> pending;
    double totalMs[kNumScopes] = {};
    uint64_t launches[kNumScopes] = {};

    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    void harvest() {
        for (Pair& p : pending) {
            float ms = 0.0f;
            if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
                totalMs[p.scope] += ms;
                launches[p.scope] += 1;
            }
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
    }
    void destroy() {
        harvest();
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
        pool.clear();
    }
};

// RAII scope: records a start event now and an end event when it closes.
struct Scope {
    Profiler& p;
    int id;
    hipStream_t s;
    hipEvent_t a = nullptr;
    Scope(Profiler& prof, int scope, hipStream_t stream) : p(prof), id(scope), s(stream) {
        if (p.mask & (1u << scope)) {
            a = p.get();
            if (a) (void)hipEventRecord(a, s);
        }
    }
    ~Scope() {
        if (!a) return;
        hipEvent_t b = p.get();
        if (!b) return;
        (void)hipEventRecord(b, s);
        p.pending.push_back({id, a, b});
    }
};

// A deepening candidate as kDeepenList writes it.
struct DeepenHostRec {
    uint32_t numPoints, level, x, y, z, node;
};

// The allocation-free deepening (remo_deepen_noalloc) keeps its candidate
// records and counters in this file-scope device array instead of memory it
// allocates: 256 B of counters, then up to kStaticDeepenRecords records
// (4.8 MB; every GPU the library runs on holds its own copy until the process
// ends). Host code never passes `scratch` itself to a kernel -- its host-side
// name is not a device address -- but the address hipGetSymbolAddress returns
// for the current device, looked up once per device at remo_create.
constexpr uint32_t kStaticDeepenRecords = 200000;
constexpr size_t kStaticCounterBytes = 256;
constexpr size_t kStaticDeepenBytes = kStaticCounterBytes + size_t(kStaticDeepenRecords) * kDeepenRecordBytes;
static __device__ __align__(256) uint8_t scratch[kStaticDeepenBytes];

std::mutex g_scratchMutex;
std::vector<void*> g_scratchByDevice;  // device ordinal -> address of `scratch` there

// The current device's address of `scratch`, found on the first call per device.
hipError_t staticScratchOf(int dev, void** out) {
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (dev < 0) return hipErrorInvalidDevice;
    if (size_t(dev) >= g_scratchByDevice.size()) g_scratchByDevice.resize(size_t(dev) + 1, nullptr);
    if (!g_scratchByDevice[size_t(dev)]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratchByDevice[size_t(dev)] = p;
    }
    *out = g_scratchByDevice[size_t(dev)];
    return hipSuccess;
}

}  // namespace

struct remo_ctx {
    remo_config cfg;
    remo_cloud_meta meta;
    DevTree t;
    hipStream_t stream = nullptr;
    Globals* hostG = nullptr;  // pinned mirror
    uint32_t* hostScratch = nullptr;  // pinned, 16 words after the mirror
    Globals* hostGQueued = nullptr;   // pinned, after the scratch: the counters remo_request_stats reads
    hipEvent_t statsReady = nullptr;  // recorded after that read
    bool statsQueued = false;
    uint64_t budgetTicks = 0;         // step budget in device wall-clock ticks
    const uint32_t* uploadedPtr = nullptr;  // ring cursor last read, and its value
    uint32_t uploadedSeen = 0;
    LaunchShape shape;
    std::string err;
    bool dead = false;
    bool needsReset = true;
    uint32_t batchletIndex = 0;
    uint32_t batchesTotal = 0;        // num_batches_total of the last build_step's ring
    hipEvent_t built = nullptr;       // recorded after each build_step's last kernel
    std::vector<void*> allocations;
    AnalysisState analysis;
    RefineState refine;
    void* deepenRecords = nullptr;    // [maxNodes] candidate records (first remo_deepen)
    uint32_t* deepenCount = nullptr;  // device counter; deepenCount[1] stays 0: an empty batch's size
    std::unique_ptr<DeepenHostRec[]> deepenHost;  // [maxNodes] host copy of the records (sorted there)
    std::unique_ptr<uint32_t[]> deepenIds;        // [maxNodes] the chosen leaves
    uint8_t* staticScratch = nullptr;             // this device's address of `scratch`
    AccumWork accWork;
    uint32_t accumParity = 0;  // which of the two tile counters the next Accumulate plans into
    Profiler prof;
};

// A slot's size and the producer cursor, written on the copy stream after the
// slot's data (so a consumer that sees the cursor sees the data).
__global__ void kRingMark(uint32_t* sizes, uint32_t slot, uint32_t n, uint32_t* uploaded, uint32_t cursor) {
    sizes[slot] = n;
    if (*uploaded < cursor) *uploaded = cursor;
}
constexpr uint32_t kNoBatch = 0xFFFFFFFFu;

struct remo_ring {
    remo_point* slots = nullptr;
    uint32_t* sizes = nullptr;
    uint32_t* uploaded = nullptr;
    uint32_t slotCapacity = 0, numSlots = 0, hostUploaded = 0;
    std::vector<uint32_t> hostSizes;
    // asynchronous pushes: a copy stream, and per slot the event after its last push
    hipStream_t copy = nullptr;
    std::vector<hipEvent_t> slotEvent;
    std::vector<uint32_t> slotBatch;  // batch index last pushed into each slot (kNoBatch: none)
};

namespace {

bool isSticky(hipError_t e) {
    switch (e) {
        case hipErrorIllegalAddress:
        case hipErrorLaunchFailure:
        case hipErrorAssert:
        case hipErrorLaunchTimeOut:
        case hipErrorECCNotCorrectable:
            return true;
        default:
            return false;
    }
}

int fail(remo_ctx* c, hipError_t e, const char* what) {
    if (e == hipSuccess) return REMO_OK;
    c->err = std::string(what) + ": " + hipGetErrorString(e);
    if (isSticky(e)) {
        c->dead = true;
        return REMO_ERR_DEVICE;
    }
    return e == hipErrorOutOfMemory ? REMO_ERR_OOM : REMO_ERR_INVALID;
}

#define CK(call, what)                              \
    do {                                            \
        const hipError_t e_ = (call);               \
        if (e_ != hipSuccess) return fail(c, e_, what); \
    } while (0)

template <typename T>
int devAlloc(remo_ctx* c, T** p, uint64_t count, const char* what) {
    void* ptr = nullptr;
    const uint64_t bytes = std::max<uint64_t>(count * sizeof(T), 16);
    const hipError_t e = hipMalloc(&ptr, bytes);
    if (e != hipSuccess) {
        c->err = std::string("hipMalloc failed for ") + what + " (" +
                 std::to_string(bytes >> 20) + " MiB): " + hipGetErrorString(e);
        return REMO_ERR_OOM;
    }
    c->allocations.push_back(ptr);
    *p = static_cast<T*>(ptr);
    return REMO_OK;
}

void releaseAll(remo_ctx* c) {
    for (void* p : c->allocations) (void)hipFree(p);
    c->allocations.clear();
    if (c->hostG) (void)hipHostFree(c->hostG);
    c->hostG = nullptr;
    if (c->built) (void)hipEventDestroy(c->built);
    c->built = nullptr;
    if (c->statsReady) (void)hipEventDestroy(c->statsReady);
    c->statsReady = nullptr;
    if (c->stream) (void)hipStreamDestroy(c->stream);
    c->stream = nullptr;
}

int readGlobals(remo_ctx* c) {
    CK(hipMemcpyAsync(c->hostG, c->t.g, sizeof(Globals), hipMemcpyDeviceToHost, c->stream),
       "reading device counters");
    CK(hipStreamSynchronize(c->stream), "device work");
    return REMO_OK;
}

// The fraction of the cloud's batches ingested (a learned-score input).
float arrivalOf(const remo_ctx* c) {
    return c->batchesTotal ? std::min(1.0f, float(c->batchletIndex) / float(c->batchesTotal)) : 1.0f;
}

void fillStats(const remo_ctx* c, const Globals& g, remo_stats* s) {
    std::memset(s, 0, sizeof(*s));
    s->num_nodes = g.numNodes - 8u * g.numFreeBlocks;
    s->num_inner = g.sInner;
    s->num_leaves = g.sLeaves;
    s->num_nonempty_leaves = g.sNonempty;
    s->num_points = g.sPoints;
    s->num_voxels = g.sVoxels;
    s->num_points_processed = g.numPointsProcessed;
    s->batchlet_index = g.batchletIndex;
    s->num_chunks_points = g.sChunksPoints;
    s->num_chunks_voxels = g.sChunksVoxels;
    s->max_points_per_node = g.sMaxPoints;
    s->counter_underflows = g.sUnderflows;
    s->max_depth = g.sMaxDepth;
    s->num_allocated_chunks = uint64_t(g.numAllocatedChunks);
    s->chunk_pool_size = uint64_t(g.chunkPoolSize);
    s->allocated_bytes_persistent = g.arenaOffset;
    s->persistent_capacity = c->t.arenaCapacity;
    s->mem_capacity_reached = g.memCapacityReached;
    s->node_capacity_reached = g.numNodes >= c->t.maxNodes && g.numFreeBlocks == 0 ? 1u : 0u;
    s->free_node_blocks = g.numFreeBlocks;
    s->pooled_grids = g.numFreeGrids;
    s->node_pool_overflow = g.nodePoolOverflow;
    s->alloc_overflow = g.allocOverflow;
    s->spill_overflow = g.spillOverflow;
    s->backlog_overflow = g.backlogOverflow;
    s->dir_overflow = g.dirOverflow | g.candidateOverflow;
    s->batches = g.batches;
    s->expand_iters = g.expandIters;
    s->nodes_split = g.nodesSplit;
    s->spilled_points = g.spilledPoints;
    s->new_voxels = g.newVoxels;
    s->total_batches = g.totBatches;
    s->total_expand_iters = g.totExpandIters;
    s->total_nodes_split = g.totNodesSplit;
    s->total_spilled_points = g.totSpilledPoints;
    s->total_new_voxels = g.totNewVoxels;
    s->morton_watermark = g.mortonWatermark;
    s->num_accumulated = g.numAccumulated;
    s->sum_leaf_counts = g.sumLeafCounts;
    s->points_folded = g.pointsFolded;
    s->leaves_folded = g.leavesFolded;
    s->inner_with_sums = g.innerWithSums;
    s->draw_items = c->analysis.drawItems;
    s->draw_list_overflow = c->analysis.drawListOverflow;
    s->draw_samples = c->analysis.drawSamples;
}

}  // namespace

extern "C" {

int remo_create(const remo_config* cfgIn, const remo_cloud_meta* meta, remo_ctx** out) {
    if (!cfgIn || !meta || !out) return REMO_ERR_INVALID;
    *out = nullptr;
    remo_ctx* c = new (std::nothrow) remo_ctx();
    if (!c) return REMO_ERR_OOM;
    c->cfg = *cfgIn;
    c->meta = *meta;
    remo_config& cfg = c->cfg;
    // reference defaults (src/pipelines/RemolodPipeline.cpp:154-167,
    // remolod_octree.cu:26-40, :1138, :1169)
    if (!cfg.max_nodes) cfg.max_nodes = REMO_DEFAULT_MAX_NODES;
    if (!cfg.max_batches_per_step) cfg.max_batches_per_step = 20;
    if (cfg.step_budget_ms == 0.0f) cfg.step_budget_ms = 10.0f;
    if (!cfg.slot_capacity) cfg.slot_capacity = REMO_DEFAULT_SLOT_CAPACITY;
    if (!cfg.ring_slots) cfg.ring_slots = REMO_DEFAULT_RING_SLOTS;
    if (!cfg.spill_capacity) cfg.spill_capacity = 10000000u;
    if (!cfg.backlog_capacity) cfg.backlog_capacity = 10000000u;
    // whole segments (the backlog is cut into kBacklogShards equal segments)
    cfg.backlog_capacity = uint32_t(std::min<uint64_t>(
        (uint64_t(cfg.backlog_capacity) + kBacklogShards - 1) / kBacklogShards * kBacklogShards, 0xFFFFFFC0ull));
    if (!cfg.closure_batches) cfg.closure_batches = 8u;
    if (!cfg.persistent_bytes) {
        // the same room above the store margin for both expand arms
        cfg.persistent_bytes = std::max<uint64_t>(uint64_t(48.0 * double(meta->num_points)), 512ull << 20) +
                               (safetyMargin(cfg.fixed_depth) - kSafetyMargin);
    }
    if (meta->num_points == 0 || cfg.fixed_depth > REMO_MAX_FIXED_DEPTH || cfg.max_nodes > kTopNodeMask) {
        delete c;
        return REMO_ERR_INVALID;  // "cloud is empty", no such fixed depth, or a node pool past 2^27
    }
    const float size = std::max(std::max(meta->box_size[0], meta->box_size[1]), meta->box_size[2]);
    if (!(size > 0.0f)) {
        delete c;
        return REMO_ERR_INVALID;
    }
    c->deepenHost.reset(new (std::nothrow) DeepenHostRec[cfg.max_nodes]);
    c->deepenIds.reset(new (std::nothrow) uint32_t[cfg.max_nodes]);
    if (!c->deepenHost || !c->deepenIds) {
        delete c;
        return REMO_ERR_OOM;
    }

    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        delete c;
        return REMO_ERR_DEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        delete c;
        return REMO_ERR_DEVICE;
    }
    {
        void* p = nullptr;
        if (staticScratchOf(dev, &p) != hipSuccess) {
            delete c;
            return REMO_ERR_DEVICE;
        }
        c->staticScratch = static_cast<uint8_t*>(p);
    }
    const uint32_t cus = std::max(1, prop.multiProcessorCount);
    c->shape.pointBlocks = std::max(8u, (cus * 8u) & ~7u);
    c->shape.nodeBlocks = std::max(8u, (cus * 2u) & ~7u);

    static_assert(sizeof(Globals) % 8 == 0, "the pinned block holds two mirrors");
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->built, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->statsReady, hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&c->hostG), 2 * sizeof(Globals) + 64, hipHostMallocDefault) !=
            hipSuccess) {
        releaseAll(c);
        delete c;
        return REMO_ERR_DEVICE;
    }
    std::memset(c->hostG, 0, 2 * sizeof(Globals) + 64);
    c->hostScratch = reinterpret_cast<uint32_t*>(c->hostG + 1);
    c->hostGQueued = reinterpret_cast<Globals*>(reinterpret_cast<uint8_t*>(c->hostG + 1) + 64);
    int clockKHz = 0;
    if (hipDeviceGetAttribute(&clockKHz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || clockKHz <= 0)
        clockKHz = 100000;
    c->budgetTicks = cfg.step_budget_ms > 0.0f ? uint64_t(double(cfg.step_budget_ms) * double(clockKHz)) : 0;

    DevTree& t = c->t;
    std::memset(&t, 0, sizeof(t));
    const uint64_t N = cfg.max_nodes;
    t.maxNodes = cfg.max_nodes;
    t.maxGrids = cfg.max_nodes / 8 + 2;
    t.octreeSize = size;
    t.arenaCapacity = cfg.persistent_bytes;
    t.chunkQueueCap = cfg.persistent_bytes / kChunkAllocBytes + 1;
    t.spillCap = cfg.spill_capacity;
    t.backlogCap = cfg.backlog_capacity;
    t.backlogSegCap = cfg.backlog_capacity / kBacklogShards;
    t.candidateCap = cfg.max_nodes;
    t.fixedDepth = cfg.fixed_depth;
    int rc = REMO_OK;
    auto A = [&](auto** p, uint64_t n, const char* what) {
        if (rc == REMO_OK) rc = devAlloc(c, p, n, what);
    };
    A(&t.firstChild, N, "node children");
    A(&t.walk, N, "node walk records");
    A(&t.parent, N, "node parents");
    A(&t.level, N, "node levels");
    A(&t.nx, N, "node X");
    A(&t.ny, N, "node Y");
    A(&t.nz, N, "node Z");
    A(&t.counter, N, "node counters");
    A(&t.numPoints, N, "node point counts");
    A(&t.numVoxels, N, "node voxel counts");
    A(&t.numVoxelsStored, N, "node stored voxel counts");
    A(&t.gridNo, N, "node grids");
    A(&t.storedOct, N * 8, "stored octant histograms");
    A(&t.batchOct, N * 8, "batch octant histograms");
    A(&t.pending, N, "pending flags");
    A(&t.pointDir, N * kPointDir, "point chunk directories");
    A(&t.topCell, kTopCells, "the top-level descent index");
    A(&t.gridOffset, t.maxGrids, "grid offsets");
    A(&t.voxelDir, uint64_t(t.maxGrids) * kVoxelDir, "voxel chunk directories");
    A(&t.voxelChunks, t.maxGrids, "voxel chunk counts");
    A(&t.arena, t.arenaCapacity, "the persistent octree store");
    A(&t.chunkQueue, t.chunkQueueCap, "the chunk queue");
    A(&t.spill, t.spillCap, "spilled points");
    A(&t.pointLeaf, uint64_t(cfg.slot_capacity) + t.spillCap, "the leaf of each point of a batch");
    A(&t.backlogPts, t.backlogCap, "the voxel backlog");
    A(&t.backlogGrid, t.backlogCap, "the voxel backlog targets (grid numbers)");
    A(&t.backlogFill, uint64_t(kBacklogShards) * kShardStride, "the voxel backlog segment counters");
    A(&t.backlogRank, t.backlogCap, "the voxel backlog ranks");
    A(&t.voxelShardCount, uint64_t(kBacklogShards) * t.maxGrids, "voxel counts per grid and backlog segment");
    A(&t.voxelShardStart, uint64_t(kBacklogShards) * t.maxGrids, "voxel slot starts per grid and backlog segment");
    A(&t.candidates, t.candidateCap, "split candidates");
    A(&t.acceptedSpill, uint64_t(t.candidateCap) * 4, "accepted splits");
    A(&t.statsPartials, uint64_t(kStatsBlocks) * kStatsWords, "stats sweep partials");
    A(&t.freeBlocks, N / 8 + 1, "the node-block pool");
    A(&t.freeGrids, t.maxGrids, "the grid pool");
    A(&t.dirtyList, N, "the analysis dirty list");
    A(&t.dirtyCount, 1, "the analysis dirty count");
    A(&t.dirtyListed, N, "the analysis dirty flags");
    A(&t.g, 1, "device counters");
    A(&t.accums, N, "the accumulator side array");
    if (t.fixedDepth) A(&t.cellCount, fixedPyramidEntries(t.fixedDepth), "the fixed-depth cell counts");
    {
        AccumWork& w = c->accWork;
        w.tileCap = N + (uint64_t(cfg.slot_capacity) * cfg.max_batches_per_step + cfg.spill_capacity) /
                            AccumWork::kTile + 1;
        w.tileCap = std::min<uint64_t>(w.tileCap, 1ull << 24);
        A(&w.tiles, w.tileCap, "accumulate tiles");
        A(&w.partials, w.tileCap, "accumulate partial sums");
        A(&w.planned, N, "accumulate planned leaves");
        A(&w.numPlanned, 2, "accumulate planned leaf counts");
        A(&w.numTiles, 2, "accumulate tile counts");
    }
    c->analysis.closureMargin = cfg.closure_batches;
    c->analysis.shape = FrameShape{c->shape.nodeBlocks, c->shape.pointBlocks, 1000u};
    if (rc == REMO_OK) rc = analysisAllocate(c->analysis, N, [&](void** p, uint64_t bytes, const char* what) {
        uint8_t* q = nullptr;
        const int r = devAlloc(c, &q, bytes, what);
        *p = q;
        return r;
    });
    if (rc != REMO_OK) {
        std::string e = c->err;
        releaseAll(c);
        delete c;
        (void)e;
        return rc;
    }
    t.rollup = c->analysis.rollup;
    t.closeAt = c->analysis.closeAt;
    t.shapeStale = c->analysis.shapeStale;
    t.closureMargin = c->analysis.closureMargin;
    *out = c;
    return remo_reset(c);
}

void remo_destroy(remo_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->prof.destroy();
    analysisRelease(c->analysis);
    refineRelease(c->refine);
    releaseAll(c);
    delete c;
}

const char* remo_last_error(const remo_ctx* c) { return c ? c->err.c_str() : "null context"; }

int remo_sync(remo_ctx* c) {
    if (!c) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    CK(hipStreamSynchronize(c->stream), "device work");
    return REMO_OK;
}

int remo_reset(remo_ctx* c) {
    if (!c) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    DevTree& t = c->t;
    CK(hipMemsetAsync(t.g, 0, sizeof(Globals), c->stream), "reset");
    CK(hipMemsetAsync(t.accums, 0, sizeof(remo_node_accum) * uint64_t(t.maxNodes), c->stream), "reset");
    // node indices restart at 0: every chunk directory starts empty (splits rely on it)
    CK(hipMemsetAsync(t.pointDir, 0, 4ull * kPointDir * uint64_t(t.maxNodes), c->stream), "reset");
    CK(hipMemsetAsync(c->accWork.numTiles, 0, 2 * sizeof(uint32_t), c->stream), "reset");
    CK(hipMemsetAsync(c->accWork.numPlanned, 0, 2 * sizeof(uint32_t), c->stream), "reset");
    c->accumParity = 0;
    CK(hipMemsetAsync(t.backlogFill, 0, 4ull * kBacklogShards * kShardStride, c->stream), "reset");
    CK(hipMemsetAsync(t.voxelShardCount, 0, 4ull * kBacklogShards * t.maxGrids, c->stream), "reset");
    CK(hipMemsetAsync(t.dirtyListed, 0, 4ull * t.maxNodes, c->stream), "reset");
    CK(hipMemsetAsync(t.dirtyCount, 0, 4, c->stream), "reset");
    launchResetTree(t, c->stream);
    refineReset(c->refine, t, c->stream);
    CK(hipGetLastError(), "reset kernel");
    const int rc = readGlobals(c);
    if (rc != REMO_OK) return rc;
    c->batchletIndex = 0;
    c->uploadedPtr = nullptr;
    c->uploadedSeen = 0;
    c->needsReset = false;
    analysisReset(c->analysis);
    return REMO_OK;
}

int remo_build_step(remo_ctx* c, const remo_batch_view* v, uint32_t* consumed) {
    if (!c || !v) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    if (!v->slots || !v->batch_sizes || !v->num_batches_uploaded) return REMO_ERR_INVALID;
    // The ring check of RemolodPipeline::build (src/pipelines/RemolodPipeline.cpp:315-328):
    // batch B is read from slot B % ring_slots, so the producer's ring must be that
    // deep, or hold the whole cloud without wrapping.
    const uint32_t R = c->cfg.ring_slots;
    const bool ringOk = v->num_slots == R || (v->num_batches_total <= R && v->num_slots >= v->num_batches_total);
    if (!ringOk || v->slot_capacity != c->cfg.slot_capacity) {
        c->err = "ring of " + std::to_string(v->num_slots) + " slots x " +
                 std::to_string(v->slot_capacity) + " points does not match the context (" +
                 std::to_string(R) + " x " + std::to_string(c->cfg.slot_capacity) + ")";
        return REMO_ERR_RING;
    }
    c->batchesTotal = v->num_batches_total;
    // The producer cursor only grows, so a value read earlier from the same ring is
    // a valid lower bound; it is re-read only when it would hold this step back.
    if (v->num_batches_uploaded != c->uploadedPtr ||
        c->uploadedSeen < c->batchletIndex + c->cfg.max_batches_per_step) {
        CK(hipMemcpyAsync(c->hostScratch, v->num_batches_uploaded, 4, hipMemcpyDeviceToHost, c->stream),
           "reading the uploaded-batch cursor");
        CK(hipStreamSynchronize(c->stream), "device work");
        c->uploadedSeen = c->hostScratch[0];
        c->uploadedPtr = v->num_batches_uploaded;
    }
    const uint32_t uploaded = c->uploadedSeen;
    DevTree& t = c->t;
    hipStream_t s = c->stream;
    const uint32_t first = c->batchletIndex;
    const uint32_t numBatches = uploaded > first ? std::min(uploaded - first, c->cfg.max_batches_per_step) : 0u;
    // step start time and per-step counters (DeviceTimeline counter block: per
    // launch); the first batch's begin kernel does it when there is one
    if (numBatches == 0) launchBeginStep(t, s);
    for (uint32_t bi = first; bi < first + numBatches; ++bi) {
        const uint32_t slot = bi % v->num_slots;
        BatchRef b{v->slots + uint64_t(slot) * v->slot_capacity, v->batch_sizes + slot};
        launchBeginBatch(t, b, bi != first, c->budgetTicks, bi == first, s);
        if (t.fixedDepth) {
            // the analytic-bucketing arm: one pass over the batch, no re-counting
            {
                Scope sc(c->prof, kScopeSplit, s);
                launchFixedExpand(t, b, c->shape, s);
            }
            CK(hipGetLastError(), "fixed-depth expand kernels");
            int rc = readGlobals(c);
            if (rc != REMO_OK) return rc;
        }
        uint32_t round = 0;
        for (; !t.fixedDepth;) {
            {
                Scope sc(c->prof, kScopeCount, s);
                launchCount(t, b, round, c->shape, s);
            }
            {
                Scope sc(c->prof, kScopeSplit, s);
                launchSplitDecide(t, round, c->shape, s);
                launchSplitApply(t, c->shape, s);
                launchSpillAndZero(t, c->shape, s);
            }
            CK(hipGetLastError(), "expand kernels");
            int rc = readGlobals(c);
            if (rc != REMO_OK) return rc;
            if (c->hostG->stop || c->hostG->numPending == 0) break;
            if (++round >= kMaxRounds) break;
        }
        if (c->hostG->stop) break;
        {
            Scope sc(c->prof, kScopeVoxel, s);
            launchVoxelSample(t, b, c->shape, s);
        }
        {
            Scope sc(c->prof, kScopeAlloc, s);
            launchAllocChunks(t, c->shape, s);
        }
        {
            Scope sc(c->prof, kScopeInsertPoints, s);
            launchInsertPoints(t, b, c->shape, s, true);
        }
        {
            Scope sc(c->prof, kScopeInsertVoxels, s);
            launchInsertVoxels(t, b, c->shape, s);
        }
        CK(hipGetLastError(), "insert kernels");
        c->batchletIndex = bi + 1;
    }
    // The per-node stats sweep (remolod_octree.cu:1248-1319) runs when stats are
    // read (remo_get_stats): nothing on the path consumes it. The consumed-batch
    // cursor is already known here: every batch either ran or hit `stop`, which
    // the round readback reported.
    // a producer pushing into a slot this step read waits for this event
    CK(hipEventRecord(c->built, c->stream), "build event");
    if (consumed) *consumed = c->batchletIndex;
    return (v->num_batches_total > 0 && c->batchletIndex >= v->num_batches_total) ? REMO_DONE : REMO_OK;
}

int remo_accumulate(remo_ctx* c, uint32_t batchIndex) {
    if (!c) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    if (c->cfg.flags & REMO_FLAG_NO_ACCUM) return REMO_OK;
    {
        Scope sc(c->prof, kScopeAccumulate, c->stream);
        launchAccumulate(c->t, c->accWork, batchIndex, c->accumParity, c->shape, c->stream);
        c->accumParity ^= 1u;
    }
    CK(hipGetLastError(), "accumulate kernels");
    return REMO_OK;
}

int remo_analysis(remo_ctx* c, const remo_shared_uniforms* u, int mode) {
    if (!c || !u) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    if (mode != REMO_ANALYSIS_ANALYTIC && mode != REMO_ANALYSIS_LEARNED) return REMO_ERR_INVALID;
    int rc;
    {
        Scope sc(c->prof, kScopeAnalysis, c->stream);
        rc = runAnalysis(c->analysis, c->t, *u, mode, c->hostG->maxLevel, c->hostG->numNodes, arrivalOf(c), c->stream,
                         &c->err);
    }
    if (rc != REMO_OK) return rc;
    CK(hipGetLastError(), "analysis kernels");
    return REMO_OK;
}

int remo_refine(remo_ctx* c, float budgetMs, uint32_t maxLevels, remo_refine_stats* out) {
    if (!c || !out) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    std::memset(out, 0, sizeof(*out));
    RefineState& r = c->refine;
    // The sums store: two entries per voxel created so far (a node takes 1.125
    // entries per occupied cell, and new ones when it outgrows them). When sums
    // were handed out past three quarters of it, or it ran out, it is replaced and
    // every node is recomputed.
    const bool fresh = !r.allocated;
    const bool compact = !fresh && (r.host->arenaFull || r.host->arenaUsed > r.accEntries / 4 * 3);
    if (fresh || compact) {
        int rc = readGlobals(c);
        if (rc != REMO_OK) return rc;
        const uint64_t v = c->hostG->totNewVoxels;
        uint64_t want = std::max<uint64_t>(std::max<uint64_t>(1ull << 22, 2 * v + (1ull << 20)),
                                           fresh ? 0ull : r.accEntries);
        // REMO_REFINE_STORE_ENTRIES caps the store (memory-constrained callers;
        // tests use it to exercise the full-store path)
        if (const char* cap = std::getenv("REMO_REFINE_STORE_ENTRIES")) {
            const unsigned long long e = std::strtoull(cap, nullptr, 10);
            if (e > 0) want = e;
        }
        if (fresh) {
            rc = refineAllocate(r, c->t, want, [&](void** p, uint64_t bytes, const char* what) {
                uint8_t* q = nullptr;
                const int e = devAlloc(c, &q, bytes, what);
                *p = q;
                return e;
            }, c->stream);
            if (rc != 0) {
                if (c->err.empty()) c->err = "refinement: allocation failed";
                return rc == -3 ? REMO_ERR_DEVICE : REMO_ERR_OOM;
            }
        } else {
            rc = refineGrow(r, c->t, want, c->stream, &c->err);
            if (rc != 0) return rc == -3 ? REMO_ERR_DEVICE : REMO_ERR_OOM;
        }
    }
    RefineStatsOut so;
    int rc;
    {
        Scope sc(c->prof, kScopeRefine, c->stream);
        rc = runRefine(r, c->t, c->hostG->numNodes, budgetMs, maxLevels, c->shape, c->stream, &so, &c->err);
        if (rc == kRefineReplaceStore) {
            // a fresh store (at least as large), every node dirty, then the pass
            int g = readGlobals(c);
            if (g != REMO_OK) return g;
            uint64_t want = std::max<uint64_t>(r.accEntries, 2 * c->hostG->totNewVoxels + (1ull << 20));
            if (const char* cap = std::getenv("REMO_REFINE_STORE_ENTRIES")) {
                const unsigned long long e = std::strtoull(cap, nullptr, 10);
                if (e > 0) want = e;
            }
            g = refineGrow(r, c->t, want, c->stream, &c->err);
            if (g != 0) return g == -3 ? REMO_ERR_DEVICE : REMO_ERR_OOM;
            rc = runRefine(r, c->t, c->hostG->numNodes, budgetMs, maxLevels, c->shape, c->stream, &so, &c->err);
        }
    }
    if (rc != 0) {
        CK(hipGetLastError(), "refinement kernels");
        CK(hipStreamSynchronize(c->stream), "refinement kernels");
        return REMO_ERR_DEVICE;
    }
    CK(hipGetLastError(), "refinement kernels");
    out->levels_pending = so.levelsPending;
    out->levels_processed = so.levelsProcessed;
    out->nodes_dirty = so.nodesDirty;
    out->nodes_refined = so.nodesRefined;
    out->nodes_skipped = so.nodesSkipped;
    out->complete = so.complete;
    out->points_folded = so.pointsFolded;
    out->child_sums = so.childSums;
    out->voxels_written = so.voxelsWritten;
    out->missing_cells = so.missingCells;
    out->empty_cells = so.emptyCells;
    out->arena_full = so.arenaFull;
    out->cell_overflow = so.cellOverflow;
    out->estimate_ms = so.estimateMs;
    out->device_ms = so.deviceMs;
    return REMO_OK;
}

// remo_deepen's work with the candidate records in `records` (room for `cap`)
// and the two counter words in `count`, device memory the caller provides.
// Host memory: the context's deepenHost / deepenIds, allocated at create.
static int deepenWith(remo_ctx* c, uint32_t maxSplits, uint32_t minPoints, void* records, uint32_t* count,
                      uint32_t cap, remo_deepen_stats* out) {
    DevTree& t = c->t;
    hipStream_t s = c->stream;
    int rc = readGlobals(c);
    if (rc != REMO_OK) return rc;
    if (c->hostG->memCapacityReached || c->hostG->arenaOffset + safetyMargin(t.fixedDepth) >= t.arenaCapacity) {
        out->refused = 1;
        return REMO_OK;
    }
    if (maxSplits == 0) return REMO_OK;
    launchDeepenList(t, records, count, cap, std::max(1u, minPoints), c->shape, s);
    CK(hipGetLastError(), "deepening list kernel");
    CK(hipMemcpyAsync(c->hostScratch, count, 4, hipMemcpyDeviceToHost, s), "deepening count");
    CK(hipStreamSynchronize(s), "deepening list");
    const uint32_t n = c->hostScratch[0];
    out->candidates = n;
    if (n > cap) {
        c->err = "deepening: " + std::to_string(n) + " candidate leaves exceed the " + std::to_string(cap) +
                 " records the candidate buffer holds";
        return REMO_ERR_CAPACITY;
    }
    if (n == 0) return REMO_OK;
    DeepenHostRec* recs = c->deepenHost.get();
    static_assert(sizeof(DeepenHostRec) == kDeepenRecordBytes, "record layout");
    CK(hipMemcpy(recs, records, sizeof(DeepenHostRec) * n, hipMemcpyDeviceToHost), "deepening records");
    // the densest first; ties by canonical key, so the choice is deterministic
    std::sort(recs, recs + n, [](const DeepenHostRec& a, const DeepenHostRec& b) {
        if (a.numPoints != b.numPoints) return a.numPoints > b.numPoints;
        if (a.level != b.level) return a.level < b.level;
        if (a.x != b.x) return a.x < b.x;
        if (a.y != b.y) return a.y < b.y;
        return a.z < b.z;
    });
    // the chosen leaves' points must fit the spill list they move through, and
    // their new voxels (at most one per point) the voxel backlog
    const uint64_t pointBudget = std::min<uint64_t>(t.spillCap, t.backlogCap);
    uint32_t chosen = 0;
    uint64_t points = 0;
    const uint32_t limit = std::min(std::min(n, maxSplits), t.candidateCap);
    while (chosen < limit && points + recs[chosen].numPoints <= pointBudget) points += recs[chosen++].numPoints;
    out->chosen = chosen;
    if (chosen == 0) return REMO_OK;
    uint32_t* ids = c->deepenIds.get();
    for (uint32_t k = 0; k < chosen; ++k) ids[k] = recs[k].node;
    CK(hipMemcpy(t.candidates, ids, 4ull * chosen, hipMemcpyHostToDevice), "deepening candidates");
    launchDeepenSplit(t, chosen, BatchRef{t.spill, count + 1}, c->shape, s);
    CK(hipGetLastError(), "deepening kernels");
    rc = readGlobals(c);
    if (rc != REMO_OK) return rc;
    const Globals& g = *c->hostG;
    out->splits = g.numAccepted;
    out->points_moved = std::min<uint64_t>(std::min<uint64_t>(g.numSpilled, t.spillCap), g.spillLimit);
    out->new_voxels = std::min<uint64_t>(g.numBacklog, t.backlogCap);
    return REMO_OK;
}

int remo_deepen(remo_ctx* c, uint32_t maxSplits, uint32_t minPoints, remo_deepen_stats* out) {
    if (!c || !out) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    std::memset(out, 0, sizeof(*out));
    if (maxSplits != 0 && !c->deepenRecords) {
        int rc = devAlloc(c, reinterpret_cast<uint8_t**>(&c->deepenRecords),
                          uint64_t(c->t.maxNodes) * kDeepenRecordBytes, "deepening candidates");
        if (rc == REMO_OK) rc = devAlloc(c, &c->deepenCount, 2, "deepening counter");
        if (rc != REMO_OK) return rc;
    }
    return deepenWith(c, maxSplits, minPoints, c->deepenRecords, c->deepenCount, c->t.maxNodes, out);
}

int remo_deepen_noalloc(remo_ctx* c, uint32_t maxSplits, uint32_t minPoints, remo_deepen_stats* out) {
    if (!c || !out) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    std::memset(out, 0, sizeof(*out));
    if (!c->staticScratch) {
        c->err = "deepening without allocation: no scratch address for this device";
        return REMO_ERR_STATE;
    }
    const int rc = readGlobals(c);
    if (rc != REMO_OK) return rc;
    // every candidate is a leaf, so a tree within the records' capacity always fits
    if (c->hostG->numNodes > kStaticDeepenRecords) {
        c->err = "deepening without allocation: the tree has " + std::to_string(c->hostG->numNodes) +
                 " nodes; its fixed candidate buffer holds " + std::to_string(kStaticDeepenRecords);
        return REMO_ERR_CAPACITY;
    }
    uint32_t* count = reinterpret_cast<uint32_t*>(c->staticScratch);
    void* records = c->staticScratch + kStaticCounterBytes;
    return deepenWith(c, maxSplits, minPoints, records, count, kStaticDeepenRecords, out);
}

int remo_collapse(remo_ctx* c, uint32_t maxCollapses, uint32_t maxPoints, remo_collapse_stats* out) {
    if (!c || !out) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    std::memset(out, 0, sizeof(*out));
    if (maxCollapses == 0) return REMO_OK;
    DevTree& t = c->t;
    hipStream_t s = c->stream;
    // the candidate records share remo_deepen's buffer (an inner node's record
    // has the leaf record's layout)
    if (!c->deepenRecords) {
        int rc = devAlloc(c, reinterpret_cast<uint8_t**>(&c->deepenRecords), uint64_t(t.maxNodes) * kDeepenRecordBytes,
                          "collapse candidates");
        if (rc == REMO_OK) rc = devAlloc(c, &c->deepenCount, 2, "collapse counter");
        if (rc != REMO_OK) return rc;
    }
    const uint32_t limit = std::min(maxPoints, kMaxPointsPerNode);
    launchCollapseList(t, c->deepenRecords, c->deepenCount, t.maxNodes, limit, c->shape, s);
    CK(hipGetLastError(), "collapse list kernel");
    CK(hipMemcpyAsync(c->hostScratch, c->deepenCount, 4, hipMemcpyDeviceToHost, s), "collapse count");
    CK(hipStreamSynchronize(s), "collapse list");
    const uint32_t n = std::min(c->hostScratch[0], t.maxNodes);
    out->candidates = n;
    if (n == 0) return REMO_OK;
    DeepenHostRec* recs = c->deepenHost.get();
    CK(hipMemcpy(recs, c->deepenRecords, sizeof(DeepenHostRec) * n, hipMemcpyDeviceToHost), "collapse records");
    // the sparsest first; ties by canonical key, so the choice is deterministic
    std::sort(recs, recs + n, [](const DeepenHostRec& a, const DeepenHostRec& b) {
        if (a.numPoints != b.numPoints) return a.numPoints < b.numPoints;
        if (a.level != b.level) return a.level < b.level;
        if (a.x != b.x) return a.x < b.x;
        if (a.y != b.y) return a.y < b.y;
        return a.z < b.z;
    });
    // the merged nodes' points move through the spill list
    uint32_t chosen = 0;
    uint64_t points = 0;
    const uint32_t most = std::min(std::min(n, maxCollapses), t.candidateCap);
    while (chosen < most && points + recs[chosen].numPoints <= t.spillCap) points += recs[chosen++].numPoints;
    if (chosen == 0) return REMO_OK;
    uint32_t* ids = c->deepenIds.get();
    for (uint32_t k = 0; k < chosen; ++k) ids[k] = recs[k].node;
    int rc = readGlobals(c);
    if (rc != REMO_OK) return rc;
    CK(hipMemcpy(t.candidates, ids, 4ull * chosen, hipMemcpyHostToDevice), "collapse candidates");
    const uint32_t gridsBefore = c->hostG->numFreeGrids;
    launchCollapse(t, chosen, BatchRef{t.spill, c->deepenCount + 1}, c->refine.refPoints, c->refine.refVoxels,
                   c->shape, s);
    CK(hipGetLastError(), "collapse kernels");
    // every node's roll-up record is rebuilt: the merged leaves' sums are folded anew
    c->analysis.needFull = true;
    rc = readGlobals(c);
    if (rc != REMO_OK) return rc;
    const Globals& g = *c->hostG;
    out->collapsed = chosen;
    out->points_moved = std::min<uint64_t>(std::min<uint64_t>(g.numSpilled, t.spillCap), g.spillLimit);
    out->grids_pooled = g.numFreeGrids - gridsBefore;
    out->voxels_dropped = g.droppedVoxels;
    return REMO_OK;
}

int remo_render_flat(remo_ctx* c, const remo_point* points, uint64_t n, const remo_shared_uniforms* u, uint64_t* fb) {
    if (!c || !u || !fb || (n && !points)) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    if (!(u->width >= 1.0f && u->height >= 1.0f)) return REMO_ERR_INVALID;
    {
        Scope sc(c->prof, kScopeFlat, c->stream);
        launchRenderFlat(points, n, *u, fb, c->shape.pointBlocks, c->stream);
    }
    CK(hipGetLastError(), "flat render kernels");
    return REMO_OK;
}

int remo_apply_edl(remo_ctx* c, const remo_shared_uniforms* u, uint64_t* fb) {
    if (!c || !u || !fb) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    if (u->enable_edl == 0) return REMO_OK;
    launchApplyEdl(fb, *u, c->stream);
    CK(hipGetLastError(), "EDL kernel");
    return REMO_OK;
}

int remo_accum_fp32_atomics(remo_ctx* c, float* out, uint32_t cap, uint32_t* num) {
    if (!c || !num) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    int rc = readGlobals(c);
    if (rc != REMO_OK) return rc;
    *num = c->hostG->numNodes;
    if (!out) return REMO_OK;
    launchAccumFp32Atomics(c->t, out, std::min(cap, c->hostG->numNodes), c->shape, c->stream);
    CK(hipGetLastError(), "fp32 atomics kernel");
    CK(hipStreamSynchronize(c->stream), "fp32 atomics kernel");
    return REMO_OK;
}

int remo_set_rollup_limits(remo_ctx* c, uint32_t ldsSlots) {
    if (!c) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    const AnalysisState defaults{};
    c->analysis.rollLdsSlots = ldsSlots == 0 ? defaults.rollLdsSlots : std::min(ldsSlots, defaults.rollLdsSlots);
    return REMO_OK;
}

int remo_set_score_model(remo_ctx* c, const remo_mlp* mlp) {
    if (!c || !mlp) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    CK(hipStreamSynchronize(c->stream), "device work");
    const int rc = mlpCopyWeights(mlp, c->analysis.mlpWeights);
    if (rc != REMO_OK) return rc;
    c->analysis.mlpLoaded = true;
    return REMO_OK;
}

int remo_mlp_features(remo_ctx* c, float* out, uint32_t cap, uint32_t* num) {
    if (!c || !num) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    int rc = readGlobals(c);
    if (rc != REMO_OK) return rc;
    *num = c->hostG->numNodes;
    if (!out) return REMO_OK;
    rc = runFeatures(c->analysis, c->t, out, cap, c->stream);
    if (rc != REMO_OK) return rc;
    CK(hipGetLastError(), "feature kernel");
    CK(hipStreamSynchronize(c->stream), "feature kernel");
    return REMO_OK;
}

int remo_render(remo_ctx* c, const remo_shared_uniforms* u, uint64_t* fb) {
    if (!c || !u || !fb) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    if (!(u->width >= 1.0f) || !(u->height >= 1.0f)) return REMO_ERR_INVALID;
    int rc;
    {
        Scope sc(c->prof, kScopeRender, c->stream);
        rc = runRender(c->analysis, c->t, *u, fb, c->stream, &c->err);
    }
    if (rc != REMO_OK) return rc;
    CK(hipGetLastError(), "render kernels");
    return REMO_OK;
}

int remo_frame(remo_ctx* c, const remo_shared_uniforms* u, int mode, uint64_t* fb) {
    if (!c || !u || !fb) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    if (mode != REMO_ANALYSIS_ANALYTIC && mode != REMO_ANALYSIS_LEARNED) return REMO_ERR_INVALID;
    if (!(u->width >= 1.0f) || !(u->height >= 1.0f)) return REMO_ERR_INVALID;
    // the tree depth (read back by every build_step) fixes the number of roll-up
    // launches in the graph
    int rc;
    {
        Scope sc(c->prof, kScopeFrame, c->stream);
        rc = runFrame(c->analysis, c->t, *u, mode, c->hostG->maxLevel, c->hostG->numNodes, arrivalOf(c), fb,
                      c->stream, &c->err);
    }
    if (rc != REMO_OK) return rc;
    CK(hipGetLastError(), "frame graph");
    return REMO_OK;
}

int remo_autotune_frame(remo_ctx* c, const remo_shared_uniforms* u, int mode, uint64_t* fb, remo_tune_entry* out,
                        uint32_t cap, uint32_t* num) {
    if (!c || !u || !fb || !num) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    if (mode != REMO_ANALYSIS_ANALYTIC && mode != REMO_ANALYSIS_LEARNED) return REMO_ERR_INVALID;
    int rc = readGlobals(c);
    if (rc != REMO_OK) return rc;
    int dev = 0;
    hipDeviceProp_t prop;
    CK(hipGetDevice(&dev), "device");
    CK(hipGetDeviceProperties(&prop, dev), "device properties");
    const uint32_t cus = std::max(1, prop.multiProcessorCount);
    const uint32_t nbs[3] = {cus, 2 * cus, 4 * cus};
    const uint32_t rbs[3] = {4 * cus, 8 * cus, 16 * cus};
    const uint32_t tiles[3] = {1000u, 2000u, 4000u};
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0), "event");
    CK(hipEventCreate(&e1), "event");
    const float arrival = arrivalOf(c);
    std::vector<remo_tune_entry> res;
    FrameShape best = c->analysis.shape;
    float bestMs = 1e30f;
    for (uint32_t nb : nbs)
        for (uint32_t rb : rbs)
            for (uint32_t tile : tiles) {
                c->analysis.shape = FrameShape{nb, rb, tile};
                float ms = 1e30f;
                for (int rep = 0; rep < 4; ++rep) {  // the first run warms up
                    (void)hipEventRecord(e0, c->stream);
                    rc = runAnalysis(c->analysis, c->t, *u, mode, c->hostG->maxLevel, c->hostG->numNodes, arrival,
                                     c->stream, &c->err);
                    if (rc == REMO_OK) rc = runRender(c->analysis, c->t, *u, fb, c->stream, &c->err);
                    (void)hipEventRecord(e1, c->stream);
                    if (rc != REMO_OK) break;
                    if (hipEventSynchronize(e1) != hipSuccess) {
                        rc = REMO_ERR_DEVICE;
                        break;
                    }
                    float t = 0.0f;
                    (void)hipEventElapsedTime(&t, e0, e1);
                    if (rep > 0) ms = std::min(ms, t);
                }
                if (rc != REMO_OK) break;
                res.push_back(remo_tune_entry{nb, rb, tile, ms});
                if (ms < bestMs) {
                    bestMs = ms;
                    best = c->analysis.shape;
                }
            }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->analysis.shape = best;
    if (rc != REMO_OK) return rc;
    CK(hipGetLastError(), "autotune");
    *num = uint32_t(res.size());
    if (out)
        for (uint32_t i = 0; i < std::min<uint32_t>(cap, uint32_t(res.size())); ++i) out[i] = res[i];
    return REMO_OK;
}

int remo_get_frame_shape(remo_ctx* c, remo_tune_entry* out) {
    if (!c || !out) return REMO_ERR_INVALID;
    const FrameShape& f = c->analysis.shape;
    *out = remo_tune_entry{f.nodeBlocks, f.rasterBlocks, f.rasterTile, 0.0f};
    return REMO_OK;
}

int remo_get_stats(remo_ctx* c, remo_stats* out) {
    if (!c || !out) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    launchStatsSweep(c->t, c->shape, c->stream);
    CK(hipGetLastError(), "stats sweep");
    // the counters and the draw diagnostics in one synchronisation
    CK(hipMemcpyAsync(c->hostG, c->t.g, sizeof(Globals), hipMemcpyDeviceToHost, c->stream),
       "reading device counters");
    analysisQueueDiagnostics(c->analysis, c->hostScratch + 8, c->stream);
    CK(hipStreamSynchronize(c->stream), "device work");
    analysisTakeDiagnostics(c->analysis, c->hostScratch + 8);
    fillStats(c, *c->hostG, out);
    return REMO_OK;
}

int remo_request_stats(remo_ctx* c) {
    if (!c) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    launchStatsSweep(c->t, c->shape, c->stream);
    CK(hipGetLastError(), "stats sweep");
    CK(hipMemcpyAsync(c->hostGQueued, c->t.g, sizeof(Globals), hipMemcpyDeviceToHost, c->stream),
       "reading device counters");
    CK(hipEventRecord(c->statsReady, c->stream), "stats event");
    c->statsQueued = true;
    return REMO_OK;
}

int remo_collect_stats(remo_ctx* c, remo_stats* out) {
    if (!c || !out) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    if (!c->statsQueued) {
        c->err = "remo_collect_stats without a remo_request_stats before it";
        return REMO_ERR_STATE;
    }
    CK(hipEventSynchronize(c->statsReady), "device work");
    c->statsQueued = false;
    fillStats(c, *c->hostGQueued, out);
    return REMO_OK;
}

int remo_get_render_diagnostics(remo_ctx* c, uint32_t* items, uint64_t* samples, uint32_t* overflow) {
    if (!c) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    analysisReadDiagnostics(c->analysis, c->stream);
    CK(hipGetLastError(), "render diagnostics");
    if (items) *items = c->analysis.drawItems;
    if (samples) *samples = c->analysis.drawSamples;
    if (overflow) *overflow = c->analysis.drawListOverflow;
    return REMO_OK;
}

int remo_dump_nodes(remo_ctx* c, remo_node_record* out, uint32_t cap, uint32_t* num) {
    if (!c || !num) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    int rc = readGlobals(c);
    if (rc != REMO_OK) return rc;
    const uint32_t n = c->hostG->numNodes;
    *num = n;
    if (!out || cap < n) return out ? REMO_ERR_INVALID : REMO_OK;
    std::vector<int32_t> fc(n);
    std::vector<uint32_t> lv(n), X(n), Y(n), Z(n), cnt(n), np(n), nv(n), nvs(n), gr(n), par(n);
    auto D = [&](void* dst, const void* src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream);
    };
    const DevTree& t = c->t;
    CK(D(fc.data(), t.firstChild, 4ull * n), "dump");
    CK(D(lv.data(), t.level, 4ull * n), "dump");
    CK(D(X.data(), t.nx, 4ull * n), "dump");
    CK(D(Y.data(), t.ny, 4ull * n), "dump");
    CK(D(Z.data(), t.nz, 4ull * n), "dump");
    CK(D(cnt.data(), t.counter, 4ull * n), "dump");
    CK(D(np.data(), t.numPoints, 4ull * n), "dump");
    CK(D(nv.data(), t.numVoxels, 4ull * n), "dump");
    CK(D(nvs.data(), t.numVoxelsStored, 4ull * n), "dump");
    CK(D(gr.data(), t.gridNo, 4ull * n), "dump");
    CK(D(par.data(), t.parent, 4ull * n), "dump");
    CK(hipStreamSynchronize(c->stream), "dump");
    for (uint32_t i = 0; i < n; ++i) {
        remo_node_record& r = out[i];
        std::memset(&r, 0, sizeof(r));
        r.level = lv[i];
        r.x = X[i];
        r.y = Y[i];
        r.z = Z[i];
        r.first_child = fc[i];
        r.counter = cnt[i];
        r.num_points = np[i];
        r.num_voxels = nv[i];
        r.num_voxels_stored = nvs[i];
        r.grid = gr[i] == kNone ? -1 : int32_t(gr[i]);
        r.free_slot = (i != 0 && par[i] == kNone) ? 1u : 0u;
    }
    return REMO_OK;
}

int remo_dump_node_samples(remo_ctx* c, uint32_t node, int voxels, remo_point* out, uint32_t cap,
                           uint32_t* num) {
    if (!c || !num) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    int rc = readGlobals(c);
    if (rc != REMO_OK) return rc;
    if (node >= c->hostG->numNodes) return REMO_ERR_INVALID;
    const DevTree& t = c->t;
    uint32_t n = 0, gi = kNone;
    CK(hipMemcpy(&n, (voxels ? t.numVoxelsStored : t.numPoints) + node, 4, hipMemcpyDeviceToHost), "dump");
    CK(hipMemcpy(&gi, t.gridNo + node, 4, hipMemcpyDeviceToHost), "dump");
    *num = n;
    if (!out) return REMO_OK;
    if (cap < n) return REMO_ERR_INVALID;
    if (n == 0) return REMO_OK;
    std::vector<uint32_t> dir;
    if (voxels) {
        if (gi == kNone) return REMO_ERR_INVALID;
        dir.resize(kVoxelDir);
        CK(hipMemcpy(dir.data(), t.voxelDir + uint64_t(gi) * kVoxelDir, 4ull * kVoxelDir,
                     hipMemcpyDeviceToHost), "dump");
    } else {
        dir.resize(kPointDir);
        CK(hipMemcpy(dir.data(), t.pointDir + uint64_t(node) * kPointDir, 4ull * kPointDir,
                     hipMemcpyDeviceToHost), "dump");
    }
    for (uint32_t k = 0; k * kPointsPerChunk < n; ++k) {
        const uint32_t m = std::min(kPointsPerChunk, n - k * kPointsPerChunk);
        if (k >= dir.size() || dir[k] == 0) return REMO_ERR_CAPACITY;
        CK(hipMemcpy(out + uint64_t(k) * kPointsPerChunk, t.arena + (uint64_t(dir[k]) << 4),
                     sizeof(remo_point) * m, hipMemcpyDeviceToHost), "dump");
    }
    return REMO_OK;
}

int remo_dump_grid(remo_ctx* c, uint32_t node, uint32_t* words) {
    if (!c || !words) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    int rc = readGlobals(c);
    if (rc != REMO_OK) return rc;
    if (node >= c->hostG->numNodes) return REMO_ERR_INVALID;
    uint32_t gi = kNone;
    CK(hipMemcpy(&gi, c->t.gridNo + node, 4, hipMemcpyDeviceToHost), "dump");
    if (gi == kNone) return REMO_ERR_INVALID;
    uint64_t off = 0;
    CK(hipMemcpy(&off, c->t.gridOffset + gi, 8, hipMemcpyDeviceToHost), "dump");
    CK(hipMemcpy(words, c->t.arena + off, 4ull * kGridWords, hipMemcpyDeviceToHost), "dump");
    return REMO_OK;
}

int remo_dump_accums(remo_ctx* c, remo_node_accum* out, uint32_t cap, uint32_t* num) {
    if (!c || !num) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    int rc = readGlobals(c);
    if (rc != REMO_OK) return rc;
    const uint32_t n = c->hostG->numNodes;
    *num = n;
    if (!out) return REMO_OK;
    if (cap < n) return REMO_ERR_INVALID;
    CK(hipMemcpy(out, c->t.accums, sizeof(remo_node_accum) * n, hipMemcpyDeviceToHost), "dump");
    return REMO_OK;
}

int remo_dump_analysis(remo_ctx* c, remo_node_analysis* out, uint32_t cap, uint32_t* num) {
    if (!c || !num) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    int rc = readGlobals(c);
    if (rc != REMO_OK) return rc;
    const uint32_t n = c->hostG->numNodes;
    *num = n;
    if (!out) return REMO_OK;
    if (cap < n) return REMO_ERR_INVALID;
    CK(hipMemcpy(out, c->analysis.results, sizeof(remo_node_analysis) * n, hipMemcpyDeviceToHost), "dump");
    return REMO_OK;
}

int remo_dump_rollup(remo_ctx* c, remo_node_rollup* out, uint32_t cap, uint32_t* num) {
    if (!c || !num) return REMO_ERR_INVALID;
    if (c->dead) return REMO_ERR_DEVICE;
    int rc = readGlobals(c);
    if (rc != REMO_OK) return rc;
    const uint32_t n = c->hostG->numNodes;
    *num = n;
    if (!out) return REMO_OK;
    if (cap < n) return REMO_ERR_INVALID;
    static_assert(sizeof(RollupSums) == sizeof(remo_node_rollup), "roll-up record layout");
    std::vector<uint8_t> closed(n);
    if (n) {
        // the records are kept in the root frame; the dump gives each in its node's frame
        void* tmp = nullptr;
        CK(hipMalloc(&tmp, sizeof(RollupSums) * n), "dump");
        launchCentredRollups(c->analysis, c->t, static_cast<RollupSums*>(tmp), n, c->stream);
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipMemcpy(out, tmp, sizeof(RollupSums) * n, hipMemcpyDeviceToHost);
        hipFree(tmp);
        CK(e, "dump");
    }
    CK(hipMemcpy(closed.data(), c->analysis.closed, n, hipMemcpyDeviceToHost), "dump");
    // RollupSums ends in {n, e, pad}: the pad word carries the closure flag here
    for (uint32_t i = 0; i < n; ++i) out[i].closed = closed[i];
    return REMO_OK;
}

int remo_profile_enable(remo_ctx* c, uint32_t mask) {
    if (!c) return REMO_ERR_INVALID;
    c->prof.mask = mask;
    return REMO_OK;
}

int remo_profile_read(remo_ctx* c, remo_profile_entry* out, uint32_t cap, uint32_t* num) {
    if (!c || !num) return REMO_ERR_INVALID;
    c->prof.harvest();
    *num = kNumScopes;
    if (!out) return REMO_OK;
    for (uint32_t i = 0; i < kNumScopes && i < cap; ++i) {
        std::memset(&out[i], 0, sizeof(out[i]));
        std::strncpy(out[i].name, kScopeNames[i], sizeof(out[i].name) - 1);
        out[i].total_ms = c->prof.totalMs[i];
        out[i].launches = c->prof.launches[i];
    }
    return REMO_OK;
}

int remo_profile_clear(remo_ctx* c) {
    if (!c) return REMO_ERR_INVALID;
    c->prof.harvest();
    for (int i = 0; i < kNumScopes; ++i) {
        c->prof.totalMs[i] = 0.0;
        c->prof.launches[i] = 0;
    }
    return REMO_OK;
}

// ---- ring ---------------------------------------------------------------------

int remo_ring_create(uint32_t slotCapacity, uint32_t numSlots, remo_ring** out) {
    if (!out || !slotCapacity || !numSlots) return REMO_ERR_INVALID;
    *out = nullptr;
    remo_ring* r = new (std::nothrow) remo_ring();
    if (!r) return REMO_ERR_OOM;
    r->slotCapacity = slotCapacity;
    r->numSlots = numSlots;
    r->hostSizes.assign(numSlots, 0);
    if (hipMalloc(&r->slots, sizeof(remo_point) * uint64_t(slotCapacity) * numSlots) != hipSuccess ||
        hipMalloc(&r->sizes, 4ull * numSlots) != hipSuccess || hipMalloc(&r->uploaded, 4) != hipSuccess) {
        remo_ring_destroy(r);
        return REMO_ERR_OOM;
    }
    if (hipMemset(r->sizes, 0, 4ull * numSlots) != hipSuccess || hipMemset(r->uploaded, 0, 4) != hipSuccess) {
        remo_ring_destroy(r);
        return REMO_ERR_DEVICE;
    }
    *out = r;
    return REMO_OK;
}

void remo_ring_destroy(remo_ring* r) {
    if (!r) return;
    if (r->copy) {
        (void)hipStreamSynchronize(r->copy);
        for (hipEvent_t e : r->slotEvent)
            if (e) (void)hipEventDestroy(e);
        (void)hipStreamDestroy(r->copy);
    }
    if (r->slots) (void)hipFree(r->slots);
    if (r->sizes) (void)hipFree(r->sizes);
    if (r->uploaded) (void)hipFree(r->uploaded);
    delete r;
}

int remo_ring_upload(remo_ring* r, uint32_t batchIndex, const remo_point* host, uint32_t n) {
    if (!r || (!host && n) || n > r->slotCapacity) return REMO_ERR_INVALID;
    const uint32_t slot = batchIndex % r->numSlots;
    if (n && hipMemcpy(r->slots + uint64_t(slot) * r->slotCapacity, host, sizeof(remo_point) * uint64_t(n),
                       hipMemcpyHostToDevice) != hipSuccess)
        return REMO_ERR_DEVICE;
    r->hostSizes[slot] = n;
    if (hipMemcpy(r->sizes + slot, &n, 4, hipMemcpyHostToDevice) != hipSuccess) return REMO_ERR_DEVICE;
    if (batchIndex + 1 > r->hostUploaded) {
        r->hostUploaded = batchIndex + 1;
        if (hipMemcpy(r->uploaded, &r->hostUploaded, 4, hipMemcpyHostToDevice) != hipSuccess)
            return REMO_ERR_DEVICE;
    }
    return REMO_OK;
}

int remo_ring_view(const remo_ring* r, uint32_t numBatchesTotal, remo_batch_view* out) {
    if (!r || !out) return REMO_ERR_INVALID;
    std::memset(out, 0, sizeof(*out));
    out->slots = r->slots;
    out->batch_sizes = r->sizes;
    out->num_batches_uploaded = r->uploaded;
    out->slot_capacity = r->slotCapacity;
    out->num_slots = r->numSlots;
    out->num_batches_total = numBatchesTotal;
    return REMO_OK;
}

int remo_ring_push(remo_ring* r, uint32_t batchIndex, const remo_point* host, uint32_t n, remo_ctx* consumer) {
    if (!r || (!host && n) || n > r->slotCapacity) return REMO_ERR_INVALID;
    if (!r->copy) {
        if (hipStreamCreateWithFlags(&r->copy, hipStreamNonBlocking) != hipSuccess) return REMO_ERR_DEVICE;
        r->slotEvent.assign(r->numSlots, nullptr);
        r->slotBatch.assign(r->numSlots, kNoBatch);
        for (auto& e : r->slotEvent)
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return REMO_ERR_DEVICE;
    }
    const uint32_t slot = batchIndex % r->numSlots;
    // the slot's previous batch may still be read by the consumer's last build_step
    if (r->slotBatch[slot] != kNoBatch && consumer && consumer->built &&
        hipStreamWaitEvent(r->copy, consumer->built, 0) != hipSuccess)
        return REMO_ERR_DEVICE;
    if (n && hipMemcpyAsync(r->slots + uint64_t(slot) * r->slotCapacity, host, sizeof(remo_point) * uint64_t(n),
                            hipMemcpyHostToDevice, r->copy) != hipSuccess)
        return REMO_ERR_DEVICE;
    r->hostSizes[slot] = n;
    r->hostUploaded = std::max(r->hostUploaded, batchIndex + 1);
    hipLaunchKernelGGL(kRingMark, dim3(1), dim3(1), 0, r->copy, r->sizes, slot, n, r->uploaded, batchIndex + 1);
    if (hipGetLastError() != hipSuccess || hipEventRecord(r->slotEvent[slot], r->copy) != hipSuccess)
        return REMO_ERR_DEVICE;
    r->slotBatch[slot] = batchIndex;
    return REMO_OK;
}

int remo_ring_order_before(remo_ring* r, remo_ctx* consumer, uint32_t firstBatch, uint32_t numBatches) {
    if (!r || !consumer) return REMO_ERR_INVALID;
    if (!r->copy) return REMO_OK;  // nothing was pushed asynchronously
    for (uint32_t k = 0; k < std::min(numBatches, r->numSlots); ++k) {
        const uint32_t b = firstBatch + k;
        const uint32_t slot = b % r->numSlots;
        if (r->slotBatch[slot] != b) continue;  // not pushed (yet)
        if (hipStreamWaitEvent(consumer->stream, r->slotEvent[slot], 0) != hipSuccess) return REMO_ERR_DEVICE;
    }
    return REMO_OK;
}

int remo_ring_sync(remo_ring* r) {
    if (!r) return REMO_ERR_INVALID;
    if (r->copy && hipStreamSynchronize(r->copy) != hipSuccess) return REMO_ERR_DEVICE;
    return REMO_OK;
}

int remo_ring_rewind(remo_ring* r) {
    if (!r) return REMO_ERR_INVALID;
    if (r->copy) {
        if (hipStreamSynchronize(r->copy) != hipSuccess) return REMO_ERR_DEVICE;
        r->slotBatch.assign(r->numSlots, kNoBatch);
    }
    r->hostUploaded = 0;
    return hipMemset(r->uploaded, 0, 4) == hipSuccess ? REMO_OK : REMO_ERR_DEVICE;
}

// ---- plumbing -----------------------------------------------------------------

int remo_dev_alloc(size_t bytes, void** out) {
    if (!out) return REMO_ERR_INVALID;
    return hipMalloc(out, std::max<size_t>(bytes, 16)) == hipSuccess ? REMO_OK : REMO_ERR_OOM;
}

int remo_dev_free(void* p) { return hipFree(p) == hipSuccess ? REMO_OK : REMO_ERR_DEVICE; }

int remo_host_alloc(size_t bytes, void** out) {
    if (!out) return REMO_ERR_INVALID;
    return hipHostMalloc(out, std::max<size_t>(bytes, 16), hipHostMallocDefault) == hipSuccess ? REMO_OK
                                                                                               : REMO_ERR_OOM;
}

int remo_host_free(void* p) { return hipHostFree(p) == hipSuccess ? REMO_OK : REMO_ERR_DEVICE; }

int remo_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? REMO_OK : REMO_ERR_DEVICE;
}

int remo_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? REMO_OK : REMO_ERR_DEVICE;
}

}  // extern "C"

__global__ void remo_copy_kernel(uint4* __restrict__ dst, const uint4* __restrict__ src, uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

// Device-scope atomicOr throughput: every lane issues `perLane` atomics on
// 32-bit words of a buffer of mask + 1 words. Mode 0: scattered addresses, the
// result used (returning atomics, as voxel sampling's climb issues them); mode 1:
// scattered, result unused (non-returning); mode 2: every lane of a wave on one
// word (a hot address). Addresses come from a per-lane 64-bit LCG.
__global__ void remo_atomic_kernel(uint32_t* __restrict__ words, uint64_t mask, uint32_t perLane, uint32_t mode,
                                   uint32_t* __restrict__ sink) {
    const uint64_t tid = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    uint64_t x = tid * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    uint32_t acc = 0;
    const uint64_t waveWord = ((tid >> 6) * 0x9E3779B97F4A7C15ull >> 20) & mask;
    for (uint32_t k = 0; k < perLane; ++k) {
        x = x * 6364136223846793005ull + 1442695040888963407ull;
        const uint64_t idx = mode == 2 ? waveWord : ((x >> 24) & mask);
        const uint32_t bit = 1u << ((x >> 59) & 31u);
        if (mode == 1)
            __hip_atomic_fetch_or(words + idx, bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else
            acc |= atomicOr(words + idx, bit);
    }
    if (acc == 0xFFFFFFFFu && (tid & 0xFFFFF) == 0) sink[tid >> 20] = acc;  // keeps the results live
}

extern "C" {

int remo_dev_copy(void* dst, const void* src, size_t bytes) {
    if (!dst || !src || (bytes & 15)) return REMO_ERR_INVALID;
    hipLaunchKernelGGL(remo_copy_kernel, dim3(4096), dim3(256), 0, 0, static_cast<uint4*>(dst),
                       static_cast<const uint4*>(src), uint64_t(bytes / 16));
    if (hipGetLastError() != hipSuccess) return REMO_ERR_DEVICE;
    return hipDeviceSynchronize() == hipSuccess ? REMO_OK : REMO_ERR_DEVICE;
}

int remo_atomic_rate(void* words, size_t bytes, uint32_t mode, uint32_t per_lane, double* atomics_per_s) {
    if (!words || !atomics_per_s || mode > 2 || per_lane == 0) return REMO_ERR_INVALID;
    const uint64_t n = bytes / 4;
    if (n < 1024 || (n & (n - 1)) != 0) return REMO_ERR_INVALID;  // a power of two of words
    constexpr uint32_t kBlocks = 4096, kThreads = 256;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return REMO_ERR_DEVICE;
    uint32_t* w = static_cast<uint32_t*>(words);
    // the first launch warms up; the sink is the buffer's last word
    hipLaunchKernelGGL(remo_atomic_kernel, dim3(kBlocks), dim3(kThreads), 0, 0, w, n - 1, 4u, mode, w + n - 1);
    (void)hipEventRecord(e0, 0);
    hipLaunchKernelGGL(remo_atomic_kernel, dim3(kBlocks), dim3(kThreads), 0, 0, w, n - 1, per_lane, mode, w + n - 1);
    (void)hipEventRecord(e1, 0);
    int rc = REMO_OK;
    float ms = 0.0f;
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
        rc = REMO_ERR_DEVICE;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == REMO_OK) *atomics_per_s = double(kBlocks) * kThreads * per_lane / (double(ms) * 1e-3);
    return rc;
}

int remo_device_count(int* out) {
    if (!out) return REMO_ERR_INVALID;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *out = n;
    return REMO_OK;
}

int remo_device_mem_info(size_t* free_bytes, size_t* total_bytes) {
    if (!free_bytes || !total_bytes) return REMO_ERR_INVALID;
    *free_bytes = *total_bytes = 0;
    if (hipMemGetInfo(free_bytes, total_bytes) != hipSuccess) return REMO_ERR_DEVICE;
    return REMO_OK;
}

}  // extern "C"

// Internal device-side layout of one RemoLOD octree on a gfx950 GPU.
//
// Nodes are structure-of-arrays so that the hot descent reads one 4-byte word
// per level (first_child) out of an array that stays L2-resident, instead of
// chasing 152-byte Node records (kernels/remolod/remolod_structures.cuh:70-135).
// All 8 children of a split are allocated contiguously, so child i of node n is
// first_child[n] + i with i = (cx << 2) | (cy << 1) | cz, the reference's order.
//
// Point and voxel storage keeps the reference's 1000-point chunks and its
// persistent-store byte strides (AllocatorGlobal, kernels/simlod/utils.h.cu:185-199:
// 16,032 B per chunk, 262,160 B per 128^3 occupancy grid, first offset 16), but
// replaces the singly linked chunk list with a per-node chunk directory, so slot
// s of a node is directory[s / 1000] in one load rather than a list walk
// (remolod_octree.cu:845-857).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/remolod.h"

namespace remo_hip {

constexpr uint32_t kMaxPointsPerNode = 50000;
constexpr uint32_t kPointsPerChunk = 1000;
constexpr uint32_t kGridSize = 128;
constexpr uint32_t kGridWords = 65536;
constexpr int kMaxDepth = 20;
constexpr uint64_t kChunkAllocBytes = 16032;   // 16 * ((16,016 + 16) / 16)
constexpr uint64_t kGridAllocBytes = 262160;   // 16 * ((262,144 + 16) / 16)
constexpr uint64_t kArenaFirstOffset = 16;
// The store margin the batch loop stops within (remolod_octree.cu:1191-1195):
// 200 MB for the iterative expand, 1 GB for the analytic-bucketing arm, whose
// batches split hundreds of nodes at once.
constexpr uint64_t kSafetyMargin = 200000000ull;
constexpr uint64_t kSafetyMarginFixed = 1000000000ull;
__host__ __device__ inline uint64_t safetyMargin(uint32_t fixedDepth) {
    return fixedDepth ? kSafetyMarginFixed : kSafetyMargin;
}
constexpr uint32_t kPointDir = 256;    // chunk slots per leaf: 256,000 points (fixed-depth leaves exceed 50,000)
constexpr uint32_t kVoxelDir = 2112;   // >= ceil(2^21 / 1000) chunk slots per grid
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kMaxRounds = 20;    // the expand loop bound, remolod_octree.cu:411
// The per-batch voxel backlog is cut into segments, each with its own reservation
// counter on its own 64-byte line: one shared counter took every wave's
// returning atomic on one address (~20 us per 1M-point batch).
constexpr uint32_t kBacklogShards = 64;
constexpr uint32_t kShardStride = 16;  // u32 words between two segment counters
// Descents start from a dense index of the tree's top levels: for every cell of
// level kTopLevel, the deepest node of level <= kTopLevel that contains it (its
// index | its level << kTopLevelShift). The split kernel keeps it current.
constexpr uint32_t kTopLevel = 4;
constexpr uint32_t kTopCells = 1u << (3 * kTopLevel);
constexpr uint32_t kTopLevelShift = 27;  // node indices below 2^27
constexpr uint32_t kTopNodeMask = (1u << kTopLevelShift) - 1u;

// Device-resident counters; one instance per context.
struct Globals {
    uint32_t numNodes;
    uint32_t numGrids;
    uint64_t arenaOffset;          // persistent-store bump offset (reference strides)
    int64_t numAllocatedChunks;    // point chunks in use == chunk-queue stack pointer
    int64_t chunkPoolSize;         // high-water mark of numAllocatedChunks
    uint32_t numSpilled;
    uint32_t numBacklog;           // backlog entries held this batch (published by kInsertVoxels)
    uint32_t numCandidates;        // split candidates of the current round
    uint32_t numAccepted;          // accepted splits of the current round
    uint32_t numPending;           // new children that must split again this batch
    uint32_t stop;                 // set when the margin or the step budget is reached
    uint32_t batchletIndex;
    uint32_t memCapacityReached;
    uint64_t numPointsProcessed;
    // health flags
    uint32_t allocOverflow, nodePoolOverflow, spillOverflow, backlogOverflow;
    uint32_t dirOverflow, candidateOverflow, budgetStop, pad1;
    uint64_t stepStart;            // wall_clock64() at the start of the step
    uint32_t maxLevel;             // deepest node level (kept by the splits)
    uint32_t spillLimit;           // lowest spill offset of a split that did not fit (kNone: none);
                                   // spill slots from here on were never written
    // per-step counters
    uint64_t batches, expandIters, nodesSplit, spilledPoints, newVoxels;
    // the same, cumulative since reset
    uint64_t totBatches, totExpandIters, totNodesSplit, totSpilledPoints, totNewVoxels;
    // AccumGlobals (include/remo/RemoAccum.h:33-40)
    uint64_t mortonWatermark, numAccumulated, sumLeafCounts;
    uint32_t pointsFolded, leavesFolded, innerWithSums;
    uint32_t statsDone;            // stats-sweep workgroups finished (the last reduces, then clears it)
    // stats sweep
    uint32_t sInner, sLeaves, sNonempty, sMaxPoints;
    uint64_t sPoints, sVoxels;
    uint32_t sChunksPoints, sChunksVoxels, sUnderflows, sMaxDepth;
    // pools fed by Refinement's collapse, drawn from first by splits
    uint32_t numFreeBlocks;        // 8-node child blocks on DevTree::freeBlocks
    uint32_t numFreeGrids;         // grids on DevTree::freeGrids
    uint32_t gridsTaken;           // pooled grids taken by the current split round
    uint32_t droppedVoxels;        // voxels of the nodes the last collapse merged
    uint64_t totCollapsed;
    // kernel_accum_verify's sums as Accumulate's plan gathers them (its finish
    // publishes them into sumLeafCounts / innerWithSums and clears these)
    uint64_t sumLeafCountsAcc;
    uint32_t innerWithSumsAcc, pad3;
};

struct RollupSums;  // remolod_analysis.h

// Everything a kernel needs, passed by value (one kernel argument).
struct DevTree {
    // node structure-of-arrays, capacity maxNodes
    int32_t* firstChild;
    uint2* walk;            // per node {first child or kNone, grid offset >> 4 or 0}: one load per level
    uint32_t* parent;       // kNone for the root
    uint32_t* level;
    uint32_t* nx;
    uint32_t* ny;
    uint32_t* nz;
    uint32_t* counter;
    uint32_t* numPoints;
    uint32_t* numVoxels;
    uint32_t* numVoxelsStored;
    uint32_t* gridNo;       // kNone: no grid
    uint32_t* storedOct;    // [n * 8] octant histogram of the stored points (leaves)
    uint32_t* batchOct;     // [n * 8] octant histogram counted this round
    uint32_t* pending;      // 1: created this batch with > 50,000 points
    uint32_t* pointDir;     // [n * kPointDir] chunk refs (arena offset >> 4), 0 = none
    uint32_t* topCell;      // [kTopCells] the descent's starting node per level-kTopLevel cell
    // grids, capacity maxGrids
    uint64_t* gridOffset;   // arena offset of grid g
    uint32_t* voxelDir;     // [g * kVoxelDir] chunk refs
    uint32_t* voxelChunks;  // [g] chunks in the directory
    // persistent store
    uint8_t* arena;
    uint64_t arenaCapacity;
    uint32_t* chunkQueue;   // recycled point chunks (remolod_octree.cu:349-360, :763-770)
    uint64_t chunkQueueCap;
    // per-batch scratch
    uint32_t* pointLeaf;    // [slot capacity + spill capacity] each point's leaf | its level << kTopLevelShift,
                            // from the voxel sampler's descent, for the insertion that follows it
    remo_point* spill;
    uint32_t spillCap;
    uint32_t backlogCap;
    remo_point* backlogPts;  // [backlogCap]: segment s holds [s * backlogSegCap, (s + 1) * backlogSegCap)
    uint32_t* backlogGrid;   // [backlogCap] the grid number of each entry's node
    uint32_t* backlogFill;   // [kBacklogShards * kShardStride] entries reserved per segment
    uint32_t backlogSegCap;  // backlogCap / kBacklogShards
    uint32_t* backlogRank;   // [backlogCap] an entry's rank among its segment's entries of its grid
    uint32_t* voxelShardCount;  // [maxGrids * kBacklogShards] this batch's entries per (grid, segment)
    uint32_t* voxelShardStart;  // [maxGrids * kBacklogShards] first voxel slot of each (grid, segment) range
    uint32_t* candidates;   // split candidates of a round
    uint32_t* acceptedSpill;// [candidateCap * 4] (node, spill offset, stored points, chunk-queue top) of accepted splits
    uint32_t candidateCap;
    uint32_t maxNodes;
    uint32_t maxGrids;
    float octreeSize;
    uint32_t fixedDepth;    // 0: iterative expand; else the analytic-bucketing arm's depth
    uint32_t* cellCount;    // fixed-depth cell-count pyramid, levels 0..fixedDepth
    uint32_t* statsPartials;  // [kStatsBlocks * kStatsWords] per-workgroup partials of the stats sweep
    // Analysis' dirty list: every node whose roll-up sums changed since the last
    // Analysis -- leaves Accumulate folded, nodes that split -- listed once.
    uint32_t* dirtyList;    // [maxNodes]
    uint32_t* dirtyCount;   // [1]
    uint32_t* dirtyListed;  // [maxNodes] 1 while the node is in the list
    // Collapse's pools: first node index of each freed 8-node block, number of
    // each freed grid (stacks; splits take from the top)
    uint32_t* freeBlocks;   // [maxNodes / 8 + 1]
    uint32_t* freeGrids;    // [maxGrids]
    Globals* g;
    remo_node_accum* accums;
    // Analysis' per-node roll-up records, written where a leaf's sums change
    // (Accumulate's finish, a split's new children) and by Analysis for inner
    // nodes; a record whose shape is not yet solved is flagged stale.
    RollupSums* rollup;     // [maxNodes]
    uint32_t* closeAt;      // [maxNodes] batch index from which the node is closed
    uint32_t* shapeStale;   // [maxNodes]
    uint32_t closureMargin; // batches untouched before a leaf is closed
};

// The stats sweep reduces per workgroup into statsPartials, then one workgroup
// adds the partials (no same-address atomics).
constexpr uint32_t kStatsBlocks = 512;
constexpr uint32_t kStatsWords = 16;

// Partial sums of one Accumulate tile.
struct AccumPartial {
    long long q[9];
    double g[9];
    float c[6];
    unsigned long long morton;
    uint32_t pad[6];
};

// Work lists of one Accumulate launch.
struct AccumWork {
    static constexpr uint32_t kTile = 2048;  // points per tile
    uint4* tiles;            // (node, begin, end, index within node)
    AccumPartial* partials;  // one per tile
    uint4* planned;          // [maxNodes] the leaves planned this launch (leaf, first tile, tiles, new count)
    uint32_t* numPlanned;    // [2] leaves planned, by launch parity (a launch clears the other's)
    uint32_t* numTiles;      // [2] tiles planned, by launch parity
    uint64_t tileCap;
};

// A batch in the ring: batch B at slots + (B % numSlots) * slotCapacity,
// batch_sizes[B % numSlots] points.
struct BatchRef {
    const remo_point* points;
    const uint32_t* sizePtr;
};

// ---- device helpers -------------------------------------------------------

// Saturating float -> u32 (NaN and negatives to 0), the device cvt's rule.
__device__ __forceinline__ uint32_t f2u_sat(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return uint32_t(f);
}

// remolod_octree.cu:163-169 with octreeMin = 0: pX = u32(2^28 * x / size).
// The 20-bit coordinate the descent uses is pX >> 8: 2^28 * x / size is
// 2^8 * (2^20 * x / size) exactly, since scaling by a power of two commutes
// with IEEE rounding, and floor(256 f) >> 8 == floor(f).
__device__ __forceinline__ uint32_t quant28(float c, float size) {
    const float num = 268435456.0f * c;
    return f2u_sat(__fdiv_rn(num, size));
}

// The top-index cell of a 20-bit coordinate. A coordinate at the box maximum
// quantises to 2^20, whose bit 20 the descent never reads (it wraps to cell 0
// there): the mask keeps that wrap, and the index in bounds.
__device__ __forceinline__ uint32_t topCellOf(uint32_t X, uint32_t Y, uint32_t Z) {
    constexpr uint32_t sh = kMaxDepth - kTopLevel;
    constexpr uint32_t m = (1u << kTopLevel) - 1u;
    return (((X >> sh) & m) << (2 * kTopLevel)) | (((Y >> sh) & m) << kTopLevel) | ((Z >> sh) & m);
}

// Root -> leaf descent on the 20-bit coordinate (remolod_octree.cu:182-199) for
// N points of one lane in lock-step, entered at the top index: the N chains of
// dependent firstChild loads are in flight together instead of one after
// another, and the levels above the index cost one load.
template <uint32_t N>
__device__ __forceinline__ void descendN(const int32_t* __restrict__ firstChild, const uint32_t* __restrict__ top,
                                         const uint32_t (&X)[N], const uint32_t (&Y)[N], const uint32_t (&Z)[N],
                                         const bool (&valid)[N], uint32_t (&leaf)[N], uint32_t (&levelOut)[N]) {
    bool live[N];
    for (uint32_t j = 0; j < N; ++j) {
        const uint32_t e = valid[j] ? top[topCellOf(X[j], Y[j], Z[j])] : 0u;
        leaf[j] = e & kTopNodeMask;
        levelOut[j] = e >> kTopLevelShift;
        live[j] = valid[j];
    }
    for (uint32_t step = 0; step < uint32_t(kMaxDepth); ++step) {
        int32_t fc[N];
        for (uint32_t j = 0; j < N; ++j) fc[j] = live[j] ? firstChild[leaf[j]] : -1;
        bool any = false;
        for (uint32_t j = 0; j < N; ++j) {
            if (fc[j] < 0 || levelOut[j] >= uint32_t(kMaxDepth)) {
                live[j] = false;
                continue;
            }
            const uint32_t sh = kMaxDepth - levelOut[j] - 1;
            const uint32_t ci = (((X[j] >> sh) & 1u) << 2) | (((Y[j] >> sh) & 1u) << 1) | ((Z[j] >> sh) & 1u);
            leaf[j] = uint32_t(fc[j]) + ci;
            levelOut[j] += 1;
            any = true;
        }
        if (!any) break;
    }
}

// A slot of an 8-node block that a collapse returned to the pool: no parent and
// not the root. It is in no pass until a split takes the block (every field a
// split does not rewrite is zero: no points, voxels, grid or sums).
__device__ __forceinline__ bool isFreeSlot(const DevTree& t, uint32_t i) {
    return i != 0 && t.parent[i] == kNone;
}

__device__ __forceinline__ uint32_t octantAt(uint32_t X, uint32_t Y, uint32_t Z, uint32_t level) {
    const uint32_t sh = kMaxDepth - level - 1;
    return (((X >> sh) & 1u) << 2) | (((Y >> sh) & 1u) << 1) | ((Z >> sh) & 1u);
}

// XCD-aware block index: consecutive hardware block ids go round-robin to the 8
// XCDs, so renumber blocks so that each XCD owns one contiguous range of the
// logical index space (its L2 then holds that range's nodes and grid words).
__device__ __forceinline__ uint32_t xcdBlock() {
    const uint32_t b = blockIdx.x, n = gridDim.x;
    if ((n & 7u) != 0u) return b;
    return (b & 7u) * (n >> 3) + (b >> 3);
}

__device__ __forceinline__ uint32_t laneId() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanesBelow() {
    const uint32_t l = laneId();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

__device__ __forceinline__ remo_point* chunkPtr(uint8_t* arena, uint32_t ref) {
    return reinterpret_cast<remo_point*>(arena + (uint64_t(ref) << 4));
}

// Puts `node` on Analysis' dirty list unless it is already there. Producers
// (Accumulate's finish, the split kernel) hand each node to one thread, so the
// flag needs no atomic; the list slot is one atomic per call.
__device__ __forceinline__ void listDirty(const DevTree& t, uint32_t node) {
    if (t.dirtyListed[node] != 0) return;
    t.dirtyListed[node] = 1;
    t.dirtyList[atomicAdd(t.dirtyCount, 1u)] = node;
}

// ---- host launchers --------------------------------------------------------
struct LaunchShape {
    uint32_t pointBlocks;   // blocks of the point-parallel kernels
    uint32_t nodeBlocks;    // blocks of the node-parallel kernels
};

void launchResetTree(const DevTree& t, hipStream_t s);
void launchBeginStep(const DevTree& t, hipStream_t s);
// beginStep: this batch is the step's first, so it also opens the step (kBeginStep's work).
void launchBeginBatch(const DevTree& t, BatchRef b, bool checkBudget, uint64_t budgetTicks, bool beginStep,
                      hipStream_t s);
void launchCount(const DevTree& t, BatchRef b, uint32_t round, const LaunchShape& ls, hipStream_t s);
void launchSplitDecide(const DevTree& t, uint32_t round, const LaunchShape& ls, hipStream_t s);
void launchSplitApply(const DevTree& t, const LaunchShape& ls, hipStream_t s);
void launchSpillAndZero(const DevTree& t, const LaunchShape& ls, hipStream_t s);
// The analytic-bucketing expand of one batch (fixed depth): bucket, pyramid,
// fixedDepth split rounds, counter seeding.
void launchFixedExpand(const DevTree& t, BatchRef b, const LaunchShape& ls, hipStream_t s);
// Entries of the fixed-depth cell-count pyramid for depth d (levels 0..d).
inline uint64_t fixedPyramidEntries(uint32_t d) { return ((1ull << (3 * (d + 1))) - 1ull) / 7ull; }
void launchVoxelSample(const DevTree& t, BatchRef b, const LaunchShape& ls, hipStream_t s);
void launchAllocChunks(const DevTree& t, const LaunchShape& ls, hipStream_t s);
// leavesKnown: kVoxelSample ran over the same points since the tree last changed,
// so their leaves are taken from pointLeaf instead of a descent.
void launchInsertPoints(const DevTree& t, BatchRef b, const LaunchShape& ls, hipStream_t s, bool leavesKnown);
// closeBatch: also the batch loop's closing update (batch index, processed points)
void launchInsertVoxels(const DevTree& t, BatchRef b, const LaunchShape& ls, hipStream_t s, bool closeBatch = true);
// Refinement's local deepening: list the leaves that may split (records of 24 B:
// numPoints, level, x, y, z, node; at most cap written, count[0] = all found,
// count[1] = 0), then split the chosen ones (already in t.candidates) and
// redistribute their points; `empty` is a batch of size 0.
constexpr uint32_t kDeepenRecordBytes = 24;
void launchDeepenList(const DevTree& t, void* out, uint32_t* count, uint32_t cap, uint32_t minPoints,
                      const LaunchShape& ls, hipStream_t s);
void launchDeepenSplit(const DevTree& t, uint32_t numChosen, BatchRef empty, const LaunchShape& ls, hipStream_t s);
// Refinement's collapse: list the inner nodes whose 8 children are settled
// leaves holding at most maxPoints together (records of 24 B: points, level, x,
// y, z, node; count[0] = all found, count[1] = 0), then merge the chosen ones
// (already in t.candidates): points through the spill list into the merged
// leaf, chunks and grids to their pools. refPoints / refVoxels: Refinement's
// per-node marks (null when it never ran), reset for the merged nodes and the
// freed slots.
void launchCollapseList(const DevTree& t, void* out, uint32_t* count, uint32_t cap, uint32_t maxPoints,
                        const LaunchShape& ls, hipStream_t s);
void launchCollapse(const DevTree& t, uint32_t numChosen, BatchRef empty, uint32_t* refPoints, uint32_t* refVoxels,
                    const LaunchShape& ls, hipStream_t s);
void launchStatsSweep(const DevTree& t, const LaunchShape& ls, hipStream_t s);
void launchAccumFp32Atomics(const DevTree& t, float* out, uint32_t cap, const LaunchShape& ls, hipStream_t s);
// parity: this launch's tile counter (launches alternate, see AccumWork::numTiles)
void launchAccumulate(const DevTree& t, const AccumWork& w, uint32_t batchIndex, uint32_t parity,
                      const LaunchShape& ls, hipStream_t s);

}  // namespace remo_hip

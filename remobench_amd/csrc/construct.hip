// Update: the per-batch progressive octree build on gfx950.
//
// Follows kernel_construct / addBatch (kernels/remolod/remolod_octree.cu:940-1320)
// phase for phase -- count, split, voxel-sample, allocate chunks, insert points,
// insert voxels -- with changes that remove work without changing the tree
// (node set, counters, occupied cells and per-leaf point multisets are the
// reference's):
//
//  * Every leaf keeps an octant histogram of its stored points (storedOct). A
//    split therefore knows its children's counters at once (stored + this
//    batch's octant counts), so the reference's re-count of the spilled points
//    (expand iterations 1..N, 21.5% of construct, wiki/03_ExpandStageFindings.md)
//    only happens for a child that itself holds > 50,000 points.
//  * Counting and inserting aggregate a tile by (leaf, octant) in an LDS table,
//    one global atomic per key (the reference's labeled_partition,
//    remolod_octree.cu:205-219, at workgroup scope).
//  * Descents enter at a dense index of the level-4 cells; voxel sampling climbs
//    only while cells are new; new voxels go to a backlog cut into 64 segments
//    and are counted, ranked and inserted without per-voxel atomics.
//  * Chunks are found through a per-node directory, not a list walk.
//
// Each phase is its own kernel on one stream; kernel boundaries are the
// reference's grid.sync()s.

#include <algorithm>

#include "remolod_analysis.h"

namespace remo_hip {

namespace {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kPerThread = 4;
constexpr uint32_t kTile = kBlock * kPerThread;

// Point i of the batch-then-spill stream.
__device__ __forceinline__ remo_point streamPoint(const DevTree& t, BatchRef b, uint32_t batchSize,
                                                  uint32_t i) {
    return i < batchSize ? b.points[i] : t.spill[i - batchSize];
}

// Spilled points actually held. On overflow the counter runs past the capacity,
// and a split whose range straddles the capacity writes nothing, so the list
// ends at the lowest offset of a split that did not fit.
__device__ __forceinline__ uint32_t spilledCount(const DevTree& t) {
    return min(min(t.g->numSpilled, t.spillCap), t.g->spillLimit);
}

__device__ __forceinline__ uint32_t batchSizeOf(const DevTree& t, BatchRef b) {
    return t.g->stop ? 0u : *b.sizePtr;
}

// Blocks cover contiguous tile ranges, XCD-aware, so a wave sees 64
// consecutive (scan-line coherent) points.
struct TileRange {
    uint32_t begin, end;
};
template <uint32_t Tile = kTile>
__device__ __forceinline__ TileRange myTiles(uint32_t total) {
    const uint32_t numTiles = (total + Tile - 1) / Tile;
    const uint32_t lb = xcdBlock();
    const uint64_t b = uint64_t(numTiles) * lb / gridDim.x;
    const uint64_t e = uint64_t(numTiles) * (lb + 1) / gridDim.x;
    return {uint32_t(b), uint32_t(e)};
}

// Block-level aggregation table in LDS: open addressing on a 32-bit key, so a
// tile of 1,024 scan-line-coherent points pays one global atomic per distinct
// key instead of one per point (or per wave group).
constexpr uint32_t kAggSlots = 2048;
constexpr uint32_t kEmpty32 = 0xFFFFFFFFu;

// Keys and counts; the insert kernels add each key's reserved base.
struct KeyCounts {
    uint32_t key[kAggSlots];
    uint32_t cnt[kAggSlots];
};
struct BlockAgg : KeyCounts {
    uint32_t base[kAggSlots];
};

__device__ __forceinline__ void aggClear(KeyCounts& a) {
    for (uint32_t i = threadIdx.x; i < kAggSlots; i += blockDim.x) {
        a.key[i] = kEmpty32;
        a.cnt[i] = 0;
    }
}

__device__ __forceinline__ uint32_t aggHash(uint32_t key) { return (key * 2654435761u) >> 21; }

// Finds or claims the slot of `key` and takes a rank in it; kEmpty32 when the
// probe sequence is exhausted (the caller then goes to global memory directly).
__device__ __forceinline__ uint32_t aggInsert(KeyCounts& a, uint32_t key, uint32_t* rank) {
    uint32_t h = aggHash(key);
    for (uint32_t probe = 0; probe < 64; ++probe) {
        const uint32_t prev = atomicCAS(&a.key[h], kEmpty32, key);
        if (prev == kEmpty32 || prev == key) {
            *rank = atomicAdd(&a.cnt[h], 1u);
            return h;
        }
        h = (h + 1) & (kAggSlots - 1);
    }
    return kEmpty32;
}

// ---------------------------------------------------------------------------
// reset (kernels/remolod/remolod_reset.cu:23-88): root at level 0 with a grid.
__global__ void kResetRoot(DevTree t) {
    Globals* g = t.g;
    t.firstChild[0] = -1;
    t.parent[0] = kNone;
    t.level[0] = 0;
    t.nx[0] = t.ny[0] = t.nz[0] = 0;
    t.counter[0] = 0;
    t.numPoints[0] = 0;
    t.numVoxels[0] = 0;
    t.numVoxelsStored[0] = 0;
    t.pending[0] = 0;
    for (int i = 0; i < 8; ++i) {
        t.storedOct[i] = 0;
        t.batchOct[i] = 0;
    }
    for (uint32_t i = 0; i < kPointDir; ++i) t.pointDir[i] = 0;
    g->numNodes = 1;
    g->spillLimit = kNone;
    g->arenaOffset = kArenaFirstOffset;
    g->numGrids = 1;
    t.gridOffset[0] = g->arenaOffset;
    g->arenaOffset += kGridAllocBytes;
    t.voxelChunks[0] = 0;
    t.gridNo[0] = 0;
    t.walk[0] = make_uint2(kNone, uint32_t(t.gridOffset[0] >> 4));
}

// Zeroes `count` 256 KiB grids listed by offsets (uint4 stores).
__global__ void kZeroRootGrid(DevTree t) {
    uint4* w = reinterpret_cast<uint4*>(t.arena + t.gridOffset[0]);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < kGridWords / 4;
         i += gridDim.x * blockDim.x)
        w[i] = make_uint4(0, 0, 0, 0);
    // every top-index cell starts at the root (node 0, level 0)
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < kTopCells; i += gridDim.x * blockDim.x)
        t.topCell[i] = 0u;
}

__global__ void kBeginStep(DevTree t) {
    Globals* g = t.g;
    g->stepStart = wall_clock64();
    g->batches = g->expandIters = g->nodesSplit = g->spilledPoints = g->newVoxels = 0;
    g->stop = 0;
    g->budgetStop = 0;
}

// The batch's opening update by one thread; 0 when the step's budget stops it.
__device__ uint32_t beginBatchThread0(Globals* g, uint64_t capacity, uint64_t margin, uint32_t checkBudget,
                                      uint64_t budgetTicks, uint32_t beginStep) {
    if (beginStep) {
        // the step's first batch also opens the step (kBeginStep, one launch fewer)
        g->stepStart = wall_clock64();
        g->batches = g->expandIters = g->nodesSplit = g->spilledPoints = g->newVoxels = 0;
        g->stop = 0;
        g->budgetStop = 0;
    }
    // remolod_octree.cu:1230-1240: after a batch, stop once the step has used its
    // device-time budget. Checked here, before the next batch touches the tree.
    if (checkBudget && budgetTicks && wall_clock64() - g->stepStart > budgetTicks) {
        g->stop = 1;
        g->budgetStop = 1;
        return 0;
    }
    g->numSpilled = 0;
    g->spillLimit = kNone;
    g->numBacklog = 0;
    g->numCandidates = 0;
    g->numAccepted = 0;
    g->numPending = 0;
    // remolod_octree.cu:1181-1210: stop before a batch once the persistent store
    // is within the safety margin of its capacity.
    const bool reached = g->arenaOffset + margin >= capacity;
    g->memCapacityReached = reached ? 1u : 0u;
    g->stop = reached ? 1u : 0u;
    return 1;
}

// Launched with kBacklogShards threads: thread 0 opens the batch, every thread
// clears one backlog segment counter once the batch is going ahead.
__global__ void kBeginBatch(DevTree t, BatchRef b, uint64_t capacity, uint32_t checkBudget,
                            uint64_t budgetTicks, uint32_t beginStep) {
    Globals* g = t.g;
    __shared__ uint32_t opened;
    if (threadIdx.x == 0)
        opened = beginBatchThread0(g, capacity, safetyMargin(t.fixedDepth), checkBudget, budgetTicks, beginStep);
    __syncthreads();
    if (opened) t.backlogFill[threadIdx.x * kShardStride] = 0;
    (void)b;
}

// ---------------------------------------------------------------------------
// Count (doCounting, remolod_octree.cu:142-301). Round 0 counts this batch's
// points into the leaves that existed before it; round r > 0 counts batch and
// spilled points into the children created in round r-1 that hold > 50,000.
// Counts go to the octant histogram batchOct[leaf][octant], aggregated per
// tile in LDS.
__global__ __launch_bounds__(kBlock) void kCount(DevTree t, BatchRef b, uint32_t round) {
    __shared__ KeyCounts agg;  // 16 KB: eight workgroups of a CU resident at once
    const uint32_t batchSize = batchSizeOf(t, b);
    if (t.g->stop) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // per-round counters; nothing in this kernel reads them
        t.g->numCandidates = 0;
        t.g->numAccepted = 0;
        t.g->numPending = 0;
    }
    const uint32_t total = batchSize + (round ? spilledCount(t) : 0u);
    const TileRange tr = myTiles(total);
    const float size = t.octreeSize;
    for (uint32_t tile = tr.begin; tile < tr.end; ++tile) {
        aggClear(agg);
        __syncthreads();
        uint32_t X[kPerThread], Y[kPerThread], Z[kPerThread], leaf[kPerThread], level[kPerThread];
        bool valid[kPerThread];
        for (uint32_t j = 0; j < kPerThread; ++j) {
            const uint32_t i = tile * kTile + j * kBlock + threadIdx.x;
            valid[j] = i < total;
            X[j] = Y[j] = Z[j] = 0;
            if (!valid[j]) continue;
            const remo_point p = streamPoint(t, b, batchSize, i);
            X[j] = quant28(p.x, size) >> 8;
            Y[j] = quant28(p.y, size) >> 8;
            Z[j] = quant28(p.z, size) >> 8;
        }
        descendN(t.firstChild, t.topCell, X, Y, Z, valid, leaf, level);
        bool counted[kPerThread];
        for (uint32_t j = 0; j < kPerThread; ++j) counted[j] = valid[j] && (round == 0 || t.pending[leaf[j]] != 0);
        for (uint32_t j = 0; j < kPerThread; ++j) {
            if (!counted[j]) continue;
            // a leaf at MAX_DEPTH never splits; its points count in octant 0
            const uint32_t key =
                leaf[j] * 8u + (level[j] < uint32_t(kMaxDepth) ? octantAt(X[j], Y[j], Z[j], level[j]) : 0u);
            uint32_t rank;
            if (aggInsert(agg, key, &rank) == kEmpty32) atomicAdd(&t.batchOct[key], 1u);
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < kAggSlots; k += blockDim.x) {
            const uint32_t key = agg.key[k];
            if (key != kEmpty32) atomicAdd(&t.batchOct[key], agg.cnt[k]);
        }
        __syncthreads();
    }
}

// Split decision per node. Round 0: a leaf whose counter crosses 50,000 with
// this batch is a candidate (remolod_octree.cu:213-218); every touched leaf's
// counter becomes counter + its batch count. Round r > 0: the pending children
// are the candidates (their counters are already their region counts).
__global__ __launch_bounds__(kBlock) void kSplitDecide(DevTree t, uint32_t round) {
    Globals* g = t.g;
    if (g->stop) return;
    const uint32_t n = g->numNodes;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (t.firstChild[i] >= 0) continue;
        bool cand = false;
        if (round != 0 && t.pending[i] == 0) continue;
        if (round == 0) {
            const uint4* bo = reinterpret_cast<const uint4*>(t.batchOct + 8ull * i);
            const uint4 a = bo[0], c = bo[1];
            const uint32_t add = a.x + a.y + a.z + a.w + c.x + c.y + c.z + c.w;
            if (add == 0) continue;
            const uint32_t old = t.counter[i];
            t.counter[i] = old + add;
            cand = old <= kMaxPointsPerNode && old + add > kMaxPointsPerNode;
        } else {
            cand = t.pending[i] != 0;
            t.pending[i] = 0;
        }
        if (cand && t.level[i] < uint32_t(kMaxDepth)) {
            const uint32_t slot = atomicAdd(&g->numCandidates, 1u);
            if (slot < t.candidateCap)
                t.candidates[slot] = i;
            else
                g->candidateOverflow = 1;
        } else {
            // not splitting: this round's octant counts are spent
            uint4* bo = reinterpret_cast<uint4*>(t.batchOct + 8ull * i);
            bo[0] = make_uint4(0, 0, 0, 0);
            bo[1] = make_uint4(0, 0, 0, 0);
        }
    }
}

__device__ __forceinline__ bool keyLess(const DevTree& t, uint32_t a, uint32_t b) {
    const uint32_t la = t.level[a], lb = t.level[b];
    if (la != lb) return la < lb;
    if (t.nx[a] != t.nx[b]) return t.nx[a] < t.nx[b];
    if (t.ny[a] != t.ny[b]) return t.ny[a] < t.ny[b];
    return t.nz[a] < t.nz[b];
}

// doSplitting (remolod_octree.cu:303-400). Candidates are ranked by canonical
// key (level, X, Y, Z); the node pool accepts the lowest-ranked ones that fit,
// so child indices and refusals are deterministic. Children take their
// counters from the octant histograms; a child still above 50,000 is pending.
// Candidate keys staged in LDS for ranking (one workgroup handles a round).
constexpr uint32_t kRankSlots = 2048;

__device__ __forceinline__ bool keyLess4(const uint4& a, const uint4& b) {
    if (a.x != b.x) return a.x < b.x;
    if (a.y != b.y) return a.y < b.y;
    if (a.z != b.z) return a.z < b.z;
    return a.w < b.w;
}

// Splits whose children are at most kTopLevel deep: one per node above that level.
constexpr uint32_t kTopSplitSlots = (kTopCells - 1u) / 7u;  // 1 + 8 + ... + 8^(kTopLevel-1)

__global__ __launch_bounds__(1024) void kSplitApply(DevTree t) {
    __shared__ uint4 keys[kRankSlots];  // (level, X, Y, Z) of each candidate
    __shared__ uint4 topSplits[kTopSplitSlots];  // (X, Y, Z, first child | child level << 27)
    __shared__ uint32_t numTopSplits;
    Globals* g = t.g;
    if (g->stop) return;
    if (threadIdx.x == 0) numTopSplits = 0;
    const uint32_t nc = min(g->numCandidates, t.candidateCap);
    const uint32_t base = g->numNodes;
    // blocks and grids a collapse pooled go first (kSpillAndZero settles the counts)
    const uint32_t pooledBlocks = g->numFreeBlocks;
    const uint32_t pooledGrids = g->numFreeGrids;
    // the deepest level before this launch: a split below it makes the next
    // Analysis roll up every node, so the children's records may take any frame
    const uint32_t D0 = g->maxLevel;
    const bool staged = nc <= kRankSlots;
    if (staged) {
        for (uint32_t k = threadIdx.x; k < nc; k += blockDim.x) {
            const uint32_t c = t.candidates[k];
            keys[k] = make_uint4(t.level[c], t.nx[c], t.ny[c], t.nz[c]);
        }
    }
    __syncthreads();
    // 8 lanes per candidate: each computes the same rank and child block, lane i
    // initialises child i, lane 0 does the node's own work
    const uint32_t sub = threadIdx.x & 7u;
    for (uint32_t ci = (blockIdx.x * blockDim.x + threadIdx.x) >> 3; ci < nc; ci += (gridDim.x * blockDim.x) >> 3) {
        const uint32_t node = t.candidates[ci];
        // the rank in key order: the group's lanes take every eighth candidate
        uint32_t rank = 0;
        if (staged) {
            const uint4 mine = keys[ci];
            for (uint32_t k = sub; k < nc; k += 8) rank += keyLess4(keys[k], mine) ? 1u : 0u;
        } else {
            for (uint32_t k = sub; k < nc; k += 8) rank += keyLess(t, t.candidates[k], node) ? 1u : 0u;
        }
        for (int m = 1; m < 8; m <<= 1) rank += uint32_t(__shfl_xor(int(rank), m));
        // the lowest ranks take the pooled blocks, the rest extend the pool's
        // used range: the node pool accepts a prefix of the key order
        const bool fromPool = rank < pooledBlocks;
        const uint64_t childOffset = fromPool ? uint64_t(t.freeBlocks[pooledBlocks - 1 - rank])
                                              : uint64_t(base) + 8ull * (rank - pooledBlocks);
        if (!fromPool && childOffset + 8 > t.maxNodes) {
            if (sub == 0) {
                g->nodePoolOverflow = 1;  // stays a leaf; its points stay where they are
                uint4* bo = reinterpret_cast<uint4*>(t.batchOct + 8ull * node);
                bo[0] = make_uint4(0, 0, 0, 0);
                bo[1] = make_uint4(0, 0, 0, 0);
            }
            continue;
        }
        const uint32_t c0 = uint32_t(childOffset);
        // everything the split reads, loaded before its first store (a load from
        // one of these arrays cannot move past a store to it, so reading the
        // octant counts child by child between the children's stores made eight
        // round trips in a row)
        const uint4* so4 = reinterpret_cast<const uint4*>(t.storedOct + 8ull * node);
        const uint4* bo4 = reinterpret_cast<const uint4*>(t.batchOct + 8ull * node);
        const uint4 sLo = so4[0], sHi = so4[1], bLo = bo4[0], bHi = bo4[1];
        const uint32_t np = t.numPoints[node];
        const uint32_t gno0 = t.gridNo[node];
        const uint4 key = staged ? keys[ci] : make_uint4(t.level[node], t.nx[node], t.ny[node], t.nz[node]);
        const uint32_t lvl = key.x + 1;
        const uint32_t X = key.y, Y = key.z, Z = key.w;
        const uint64_t gridOff0 = gno0 != kNone ? t.gridOffset[gno0] : 0ull;
        const uint32_t octCount[8] = {sLo.x + bLo.x, sLo.y + bLo.y, sLo.z + bLo.z, sLo.w + bLo.w,
                                      sHi.x + bHi.x, sHi.y + bHi.y, sHi.z + bHi.z, sHi.w + bHi.w};
        // child `sub`, by its own lane (the children's many stores were one
        // lane's long sequence of store instructions)
        {
            const uint32_t c = c0 + sub;
            const uint32_t cnt = octCount[0] * (sub == 0) + octCount[1] * (sub == 1) + octCount[2] * (sub == 2) +
                                 octCount[3] * (sub == 3) + octCount[4] * (sub == 4) + octCount[5] * (sub == 5) +
                                 octCount[6] * (sub == 6) + octCount[7] * (sub == 7);
            t.firstChild[c] = -1;
            t.walk[c] = make_uint2(kNone, 0u);
            t.parent[c] = node;
            t.level[c] = lvl;
            t.nx[c] = 2 * X + ((sub >> 2) & 1u);
            t.ny[c] = 2 * Y + ((sub >> 1) & 1u);
            t.nz[c] = 2 * Z + (sub & 1u);
            t.counter[c] = cnt;
            t.numPoints[c] = 0;
            t.numVoxels[c] = 0;
            t.numVoxelsStored[c] = 0;
            t.gridNo[c] = kNone;
            const bool pend = cnt > kMaxPointsPerNode && lvl < uint32_t(kMaxDepth);
            t.pending[c] = pend ? 1u : 0u;
            if (pend) atomicAdd(&g->numPending, 1u);
            uint4* so = reinterpret_cast<uint4*>(t.storedOct + 8ull * c);
            so[0] = so[1] = make_uint4(0, 0, 0, 0);
            uint4* bo = reinterpret_cast<uint4*>(t.batchOct + 8ull * c);
            bo[0] = bo[1] = make_uint4(0, 0, 0, 0);
            remo_node_accum* a = t.accums + c;
            *a = remo_node_accum{};
            // an empty leaf's roll-up record: no points, closed from the start
            RollupSums* r = t.rollup + c;
            for (int k = 0; k < 3; ++k) r->s1[k] = 0;
            for (int k = 0; k < 6; ++k) r->s2[k] = 0;
            r->n = 0;
            r->e = 20 + max(D0, lvl) - lvl;
            r->pad = 0;
            t.closeAt[c] = 0;
            markRecordChanged(t, c);
        }
        if (sub != 0) continue;
        // the split's counter reservations, issued together: each is a returning
        // atomic on one global counter, and waiting for one before issuing the
        // next made a round's splits a chain of five memory round trips
        const uint32_t nChunks = (np + kPointsPerChunk - 1) / kPointsPerChunk;
        const bool needGrid = gno0 == kNone;
        const uint32_t spillAt = np > 0 ? atomicAdd(&g->numSpilled, np) : 0u;
        const int64_t top = np > 0 ? int64_t(atomicAdd((unsigned long long*)&g->numAllocatedChunks,
                                                       (unsigned long long)(-int64_t(nChunks))))
                                   : 0;
        const uint32_t slot = atomicAdd(&g->numAccepted, 1u);
        // a pooled grid when there is one (pooled grids are alike: no voxel
        // chunks, already counted in the store), else a new one from the store
        uint32_t gi = 0, poolTake = kNone;
        unsigned long long gridAt = 0;
        if (needGrid && pooledGrids != 0) {
            poolTake = atomicAdd(&g->gridsTaken, 1u);
            if (poolTake >= pooledGrids) poolTake = kNone;
        }
        if (needGrid && poolTake != kNone) {
            gi = t.freeGrids[pooledGrids - 1 - poolTake];
            gridAt = t.gridOffset[gi];
        } else if (needGrid) {
            gi = atomicAdd(&g->numGrids, 1u);
            gridAt = atomicAdd((unsigned long long*)&g->arenaOffset, (unsigned long long)kGridAllocBytes);
        }
        atomicMax(&g->maxLevel, lvl);
        uint4* bo = reinterpret_cast<uint4*>(t.batchOct + 8ull * node);
        bo[0] = bo[1] = make_uint4(0, 0, 0, 0);
        uint4* so = reinterpret_cast<uint4*>(t.storedOct + 8ull * node);
        so[0] = so[1] = make_uint4(0, 0, 0, 0);

        // stored points -> spill list, chunks -> queue slots [top - nChunks, top):
        // both copied by kSpillAndZero, a workgroup per chunk (a loop over a
        // leaf's ~50 chunks here was a chain of dependent loads and stores)
        uint32_t spillOff = kNone;
        if (np > 0) {
            if (uint64_t(spillAt) + np <= t.spillCap) {
                spillOff = spillAt;
            } else {
                // its points are lost (flagged); later passes stop reading the
                // spill list where this range starts
                g->spillOverflow = 1;
                atomicMin(&g->spillLimit, spillAt);
            }
        }
        t.acceptedSpill[4 * slot] = node;
        t.acceptedSpill[4 * slot + 1] = spillOff;
        t.acceptedSpill[4 * slot + 2] = np;
        t.acceptedSpill[4 * slot + 3] = uint32_t(top);
        t.numPoints[node] = 0;  // remolod_octree.cu:362; kSpillAndZero copies from the recorded count

        uint32_t gno = gno0;
        uint64_t gridOff = gridOff0;
        if (needGrid && poolTake != kNone) {
            t.voxelChunks[gi] = 0;
            t.gridNo[node] = gi;
            gno = gi;
            gridOff = gridAt;
        } else if (needGrid) {
            if (gi < t.maxGrids && gridAt + kGridAllocBytes <= t.arenaCapacity) {
                t.gridOffset[gi] = gridAt;
                t.voxelChunks[gi] = 0;
                t.gridNo[node] = gi;
                gno = gi;
                gridOff = gridAt;
            } else {
                g->allocOverflow = 1;
            }
        }
        atomicAdd((unsigned long long*)&g->nodesSplit, 1ull);
        atomicAdd((unsigned long long*)&g->totNodesSplit, 1ull);
        // publish the children last: later kernels descend through them
        t.firstChild[node] = int32_t(c0);
        t.walk[node] = make_uint2(c0, gno == kNone ? 0u : uint32_t(gridOff >> 4));
        // its sums become its (new) children's: Analysis rolls it up again
        listDirty(t, node);
        if (lvl <= kTopLevel) topSplits[atomicAdd(&numTopSplits, 1u)] = make_uint4(X, Y, Z, c0 | (lvl << kTopLevelShift));
    }
    // the split nodes' top-index cells now start at the children holding them
    // (the whole block writes them: a split near the root covers thousands)
    __syncthreads();
    for (uint32_t q = 0; q < numTopSplits; ++q) {
        const uint4 sp = topSplits[q];
        const uint32_t c0 = sp.w & kTopNodeMask, lvl = sp.w >> kTopLevelShift;
        const uint32_t sb = kTopLevel - lvl;  // a child spans 2^sb top cells per axis
        const uint32_t m = (1u << sb) - 1u;
        for (uint32_t k = threadIdx.x; k < (8u << (3 * sb)); k += blockDim.x) {
            const uint32_t ci = k >> (3 * sb);
            const uint32_t cx = ((2 * sp.x + ((ci >> 2) & 1u)) << sb) | ((k >> (2 * sb)) & m);
            const uint32_t cy = ((2 * sp.y + ((ci >> 1) & 1u)) << sb) | ((k >> sb) & m);
            const uint32_t cz = ((2 * sp.z + (ci & 1u)) << sb) | (k & m);
            t.topCell[(cx << (2 * kTopLevel)) | (cy << kTopLevel) | cz] = (c0 + ci) | (lvl << kTopLevelShift);
        }
    }
}

// After the splits of a round (one launch): bump numNodes by the accepted
// splits (kSplitApply ranked against the value it had when it started), copy
// each accepted node's stored points to the spill list (the spill loop of
// doCounting, remolod_octree.cu:250-285) and zero its grid
// (remolod_octree.cu:372-387), including a grid the node already had (the
// root's, from reset). Blocks [0, numAccepted) copy; all blocks zero.
__global__ __launch_bounds__(kBlock) void kSpillAndZero(DevTree t, uint32_t countRound) {
    Globals* g = t.g;
    if (g->stop) return;
    const uint32_t na = g->numAccepted;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // the accepted splits took the pooled blocks first, then new ones
        const uint32_t pooled = g->numFreeBlocks, fromPool = min(na, pooled);
        g->numFreeBlocks = pooled - fromPool;
        g->numNodes += 8u * (na - fromPool);
        const uint32_t taken = g->gridsTaken;
        if (taken) {
            g->numFreeGrids -= min(taken, g->numFreeGrids);
            g->gridsTaken = 0;
        }
        if (countRound) {
            g->expandIters += 1;
            g->totExpandIters += 1;
        }
    }
    // one (accepted node, chunk) pair per block step, so a large node's copy
    // spreads over many workgroups
    for (uint32_t pair = blockIdx.x; pair < na * kPointDir; pair += gridDim.x) {
        const uint32_t a = pair / kPointDir, k = pair % kPointDir;
        const uint32_t node = t.acceptedSpill[4 * a];
        const uint32_t off = t.acceptedSpill[4 * a + 1];
        const uint32_t np = t.acceptedSpill[4 * a + 2];
        if (k * kPointsPerChunk >= np) continue;
        const uint32_t ref = t.pointDir[uint64_t(node) * kPointDir + k];
        if (threadIdx.x == 0) {
            // the chunk goes back to the queue (doSplitting, remolod_octree.cu:349-360)
            const int64_t qi = int64_t(t.acceptedSpill[4 * a + 3]) - 1 - int64_t(k);
            if (qi >= 0 && uint64_t(qi) < t.chunkQueueCap) t.chunkQueue[qi] = ref;
        }
        if (off == kNone) continue;
        const uint32_t m = min(kPointsPerChunk, np - k * kPointsPerChunk);
        const remo_point* src = chunkPtr(t.arena, ref);
        for (uint32_t s = threadIdx.x; s < m; s += blockDim.x) t.spill[off + k * kPointsPerChunk + s] = src[s];
    }
    const uint64_t perGrid = kGridWords / 4;
    const uint64_t total = uint64_t(na) * perGrid;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t node = t.acceptedSpill[4 * (i / perGrid)];
        const uint32_t gi = t.gridNo[node];
        if (gi == kNone) continue;
        reinterpret_cast<uint4*>(t.arena + t.gridOffset[gi])[i % perGrid] = make_uint4(0, 0, 0, 0);
    }
}

// ---------------------------------------------------------------------------
// The analytic-bucketing expand arm (expandFixed, remolod_octree.cu:464-675).
// With the depth D fixed, a point's destination cell is the top D bits of its
// 20-bit coordinate (masked, so a point on boxMax wraps to cell 0 as the descent
// does), so one bucketing pass counts the batch into a dense Morton-ordered
// array at level D; a pyramid sums it up to the root (the Morton order of
// childIndex makes a cell's parent cell >> 3); D rounds split every leaf at level
// L < D whose cell holds points, through the same split kernels as the iterative
// arm; and every leaf at level <= D adds its cell's count to its counter.
// Leaves at level D never split, whatever their count. A leaf above D that holds
// points (one whose split the node pool refused) is not split either: the
// reference's arm would drop its points.

__device__ __forceinline__ uint32_t fixedLevelOffset(uint32_t level) { return ((1u << (3 * level)) - 1u) / 7u; }

// fixedMorton (remolod_octree.cu:469-481): childIndex order, most significant level first.
__device__ __forceinline__ uint32_t fixedMorton(uint32_t x, uint32_t y, uint32_t z, uint32_t level) {
    uint32_t code = 0;
    for (uint32_t l = 0; l < level; ++l) {
        const uint32_t bit = level - 1 - l;
        code = (code << 3) | (((x >> bit) & 1u) << 2) | (((y >> bit) & 1u) << 1) | ((z >> bit) & 1u);
    }
    return code;
}

__global__ __launch_bounds__(kBlock) void kFixedClear(DevTree t, uint32_t entries) {
    if (t.g->stop) return;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < entries; i += gridDim.x * blockDim.x)
        t.cellCount[i] = 0;
}

// 1. bucket: the batch's points into level-D cells, aggregated per tile in LDS.
__global__ __launch_bounds__(kBlock) void kFixedBucket(DevTree t, BatchRef b) {
    __shared__ KeyCounts agg;
    const uint32_t batchSize = batchSizeOf(t, b);
    if (t.g->stop) return;
    const uint32_t D = t.fixedDepth;
    const uint32_t shift = uint32_t(kMaxDepth) - D, maxCoord = (1u << D) - 1u;
    const uint32_t base = fixedLevelOffset(D);
    const TileRange tr = myTiles(batchSize);
    const float size = t.octreeSize;
    for (uint32_t tile = tr.begin; tile < tr.end; ++tile) {
        aggClear(agg);
        __syncthreads();
        for (uint32_t j = 0; j < kPerThread; ++j) {
            const uint32_t i = tile * kTile + j * kBlock + threadIdx.x;
            if (i >= batchSize) continue;
            const remo_point p = streamPoint(t, b, batchSize, i);
            const uint32_t cx = ((quant28(p.x, size) >> 8) >> shift) & maxCoord;
            const uint32_t cy = ((quant28(p.y, size) >> 8) >> shift) & maxCoord;
            const uint32_t cz = ((quant28(p.z, size) >> 8) >> shift) & maxCoord;
            const uint32_t key = base + fixedMorton(cx, cy, cz, D);
            uint32_t rank;
            if (aggInsert(agg, key, &rank) == kEmpty32) atomicAdd(&t.cellCount[key], 1u);
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < kAggSlots; k += blockDim.x) {
            const uint32_t key = agg.key[k];
            if (key != kEmpty32) atomicAdd(&t.cellCount[key], agg.cnt[k]);
        }
        __syncthreads();
    }
}

// 2. pyramid: level `level` from level + 1.
__global__ __launch_bounds__(kBlock) void kFixedPyramid(DevTree t, uint32_t level) {
    if (t.g->stop) return;
    const uint32_t cells = 1u << (3 * level);
    uint32_t* dst = t.cellCount + fixedLevelOffset(level);
    const uint32_t* src = t.cellCount + fixedLevelOffset(level + 1);
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += gridDim.x * blockDim.x) {
        uint32_t sum = 0;
        for (uint32_t i = 0; i < 8; ++i) sum += src[8u * c + i];
        dst[c] = sum;
    }
}

__global__ void kFixedRound(DevTree t) {
    t.g->numCandidates = 0;
    t.g->numAccepted = 0;
    t.g->numPending = 0;
}

// 3. materialise, one round: every empty leaf at `level` whose cell holds points.
__global__ __launch_bounds__(kBlock) void kFixedDecide(DevTree t, uint32_t level) {
    Globals* g = t.g;
    if (g->stop) return;
    const uint32_t n = g->numNodes;
    const uint32_t base = fixedLevelOffset(level);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (t.firstChild[i] >= 0 || t.level[i] != level || t.numPoints[i] != 0 || isFreeSlot(t, i)) continue;
        if (t.cellCount[base + fixedMorton(t.nx[i], t.ny[i], t.nz[i], level)] == 0) continue;
        const uint32_t slot = atomicAdd(&g->numCandidates, 1u);
        if (slot < t.candidateCap)
            t.candidates[slot] = i;
        else
            g->candidateOverflow = 1;
    }
}

// 4. seed: every leaf at level <= D adds its cell's count (allocatePointChunks'
// sizing input); one expand iteration per batch.
__global__ __launch_bounds__(kBlock) void kFixedSeed(DevTree t) {
    Globals* g = t.g;
    if (g->stop) return;
    const uint32_t n = g->numNodes, D = t.fixedDepth;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        g->expandIters += 1;
        g->totExpandIters += 1;
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t L = t.level[i];
        if (t.firstChild[i] >= 0 || L > D || isFreeSlot(t, i)) continue;
        const uint32_t cnt = t.cellCount[fixedLevelOffset(L) + fixedMorton(t.nx[i], t.ny[i], t.nz[i], L)];
        if (cnt) t.counter[i] += cnt;
    }
}

// ---------------------------------------------------------------------------
// voxelSampling / sampleVoxel (remolod_octree.cu:677-740, :95-140): at every
// gridded node on the root->leaf path, set the point's 128^3 cell; whoever sets
// a bit appends the cell-centre voxel, coloured by its point.
//
// A point's path is a chain of dependent loads (the descent) and, per gridded
// level, an atomicOr whose return says whether this lane set the bit (and so
// owns that voxel). The chain is kept short:
//  * during the descent each level's cell word is peeked (a plain load, in
//    flight beside the next level's walk load), and the path's (node, grid,
//    peek) stay in registers at fixed level indices;
//  * after it, the atomicOr of every level whose peek was clear is issued at
//    once: one round trip for the whole path instead of one per level. A stale
//    clear peek only costs an atomic; a set bit is never stale (bits are only
//    set while this kernel runs). Every bit still gets exactly one owner, the
//    lane whose atomic set it, so each new cell becomes exactly one voxel;
//  * a node's coordinates at level L are the point's own 20-bit coordinate bits
//    19..20-L (the descent's child bits), so a voxel centre needs no node load.
// A leaf deeper than the register window climbs its levels below the window
// one at a time (stopping at the first bit already set: its setter covers the
// levels above), then takes the window like every other point.
//
// New voxels are staged per wave in LDS and reach the global backlog with one
// atomicAdd per flush, on the counter of the block's backlog segment, and are
// ranked with one more per (flush, grid) on that grid's counter for the segment
// (those issued together); kAllocChunks sums the counters into numVoxels.
// Bumping per-node counters in the climb sent every wave's atomics to the few
// lines that hold the counters of the nodes a batch touches (~33 us per
// 1M-point batch).
constexpr uint32_t kStage = 128;  // staged voxels per wave

struct VoxelStage {
    remo_point pts[kStage];
    uint32_t node[kStage];
};

// Lanes holding the same key form a group: its lowest lane leads, every lane
// gets its rank within the group and the leader the group's size (ballots and
// shuffles only).
__device__ __forceinline__ void groupByKey(uint32_t key, bool held, uint32_t& leader, uint32_t& rank,
                                           uint32_t& size) {
    const uint32_t lane = laneId();
    leader = lane;
    rank = 0;
    size = 0;
    uint64_t pend = __ballot(held);
    while (pend) {
        const int l = __ffsll((unsigned long long)pend) - 1;
        const uint32_t lk = __shfl(key, l);
        const uint64_t grp = __ballot(held && key == lk) & pend;
        if ((grp >> lane) & 1ull) {
            leader = uint32_t(l);
            rank = uint32_t(__popcll(grp & lanesBelow()));
        }
        if (lane == uint32_t(l)) size = uint32_t(__popcll(grp));
        pend &= ~grp;
    }
}

// Moves a wave's staged voxels to the global backlog: the grids of the staged
// voxels (loads in flight together), the segment reservation (one atomic), then
// the per-(grid, segment) ranks of the held entries, every grid's atomic at once.
__device__ __forceinline__ void flushStage(const DevTree& t, VoxelStage& st, uint32_t& count) {
    if (count == 0) return;
    // the wave's own LDS writes must land before other lanes read them
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    static_assert(kStage == 128, "two 64-entry halves");
    const uint32_t shard = blockIdx.x & (kBacklogShards - 1u);
    const uint32_t lane = laneId();
    const bool in0 = lane < count, in1 = lane + 64 < count;
    const uint32_t node0 = in0 ? st.node[lane] : 0u, node1 = in1 ? st.node[lane + 64] : 0u;
    const uint32_t grid0 = in0 ? t.gridNo[node0] : kNone, grid1 = in1 ? t.gridNo[node1] : kNone;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(&t.backlogFill[shard * kShardStride], count);
    base = __shfl(base, 0);
    const bool held0 = in0 && base + lane < t.backlogSegCap;
    const bool held1 = in1 && base + lane + 64 < t.backlogSegCap;
    uint32_t lead0, rank0, size0, lead1, rank1, size1;
    groupByKey(grid0, held0, lead0, rank0, size0);
    groupByKey(grid1, held1, lead1, rank1, size1);
    uint32_t gb0 = 0, gb1 = 0;
    if (held0 && lead0 == lane)
        gb0 = atomicAdd(&t.voxelShardCount[uint64_t(grid0) * kBacklogShards + shard], size0);
    if (held1 && lead1 == lane)
        gb1 = atomicAdd(&t.voxelShardCount[uint64_t(grid1) * kBacklogShards + shard], size1);
    gb0 = __shfl(gb0, lead0);
    gb1 = __shfl(gb1, lead1);
    for (uint32_t h = 0; h < 2; ++h) {
        const uint32_t k = lane + 64 * h;
        const bool in = h ? in1 : in0, held = h ? held1 : held0;
        const uint64_t slot = uint64_t(shard) * t.backlogSegCap + base + k;
        if (held) {
            t.backlogPts[slot] = st.pts[k];
            t.backlogGrid[slot] = h ? grid1 : grid0;
            t.backlogRank[slot] = h ? gb1 + rank1 : gb0 + rank0;
        } else if (in) {
            // a dropped voxel still counts (no backlog entry will)
            t.g->backlogOverflow = 1;
            atomicAdd(&t.numVoxels[h ? node1 : node0], 1u);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    count = 0;
}

// One point per lane: the voxel walk is a chain of dependent loads and
// returning atomics, so more waves in flight beat more points per lane
// (measured on the 100M-point scene: 232 us per 1M-point batch at one point per
// lane, 256 us at four, 258 us at eight).
constexpr uint32_t kVoxelTile = kBlock;
// The descent keeps its last kPathRegs ancestors (node, grid) in registers, so a
// climb step is a peek and an atomic, not also a parent and a grid load; a climb
// above them loads them. (It starts at the root, not at the top index: entered
// there, the climb lost the ancestors above it and cost 8 us more per batch.)
// Three or six ancestors time the same as four (172.9 and 171.1 against 172.7 us).
constexpr uint32_t kPathRegs = 4;

// Cell of a 28-bit point coordinate in the 128^3 grid of a node at `level`.
__device__ __forceinline__ uint32_t cellIndex(uint32_t pX, uint32_t pY, uint32_t pZ, uint32_t level) {
    const uint32_t sh = (kMaxDepth + 1) - level;
    return ((pX >> sh) & (kGridSize - 1)) + ((pY >> sh) & (kGridSize - 1)) * kGridSize +
           ((pZ >> sh) & (kGridSize - 1)) * kGridSize * kGridSize;
}

__device__ __forceinline__ uint32_t* cellWord(const DevTree& t, uint32_t gref, uint32_t vi) {
    return reinterpret_cast<uint32_t*>(t.arena + (uint64_t(gref) << 4)) + (vi >> 5);
}

// Stages the wave's new voxels of one step (isNew per lane) into LDS. Every lane
// of the wave calls it together, so the staged count stays wave-uniform.
__device__ __forceinline__ void stageLevel(const DevTree& t, VoxelStage& st, uint32_t& staged, bool isNew,
                                           uint32_t level, uint32_t nd, uint32_t vi, uint32_t X, uint32_t Y,
                                           uint32_t Z, uint32_t col, float size) {
    const uint64_t nm = __ballot(isNew);
    const uint32_t cnt = uint32_t(__popcll(nm));
    if (cnt == 0) return;
    if (staged + cnt > kStage) flushStage(t, st, staged);
    if (isNew) {
        // the node's coordinates at `level`: bits 19..20-level of the point's
        // 20-bit coordinate (a point on boxMax wraps to 0, as the descent does)
        const uint32_t sh = kMaxDepth - level, m = (1u << level) - 1u;
        const float nodeSize = size / float(1u << level);
        const float minx = float((X >> sh) & m) * nodeSize;
        const float miny = float((Y >> sh) & m) * nodeSize;
        const float minz = float((Z >> sh) & m) * nodeSize;
        const uint32_t cx = vi & (kGridSize - 1), cy = (vi >> 7) & (kGridSize - 1), cz = vi >> 14;
        remo_point v;
        v.x = minx + (nodeSize * (float(cx) + 0.5f)) / 128.0f;
        v.y = miny + (nodeSize * (float(cy) + 0.5f)) / 128.0f;
        v.z = minz + (nodeSize * (float(cz) + 0.5f)) / 128.0f;
        v.color = col;
        const uint32_t slot = staged + uint32_t(__popcll(nm & lanesBelow()));
        st.pts[slot] = v;
        st.node[slot] = nd;
    }
    staged += cnt;
}

__global__ __launch_bounds__(kBlock) void kVoxelSample(DevTree t, BatchRef b) {
    __shared__ VoxelStage stage[kBlock / 64];
    const uint32_t batchSize = batchSizeOf(t, b);
    Globals* g = t.g;
    if (g->stop) return;
    const uint32_t total = batchSize + spilledCount(t);
    const TileRange tr = myTiles<kVoxelTile>(total);
    const float size = t.octreeSize;
    VoxelStage& st = stage[threadIdx.x >> 6];
    uint32_t staged = 0;  // wave-uniform
    for (uint32_t tile = tr.begin; tile < tr.end; ++tile) {
        const uint32_t i = tile * kVoxelTile + threadIdx.x;
        bool active = i < total;
        uint32_t pX = 0, pY = 0, pZ = 0, col = 0;
        if (active) {
            const remo_point p = streamPoint(t, b, batchSize, i);
            pX = quant28(p.x, size);
            pY = quant28(p.y, size);
            pZ = quant28(p.z, size);
            col = p.color;
        }
        const uint32_t X = pX >> 8, Y = pY >> 8, Z = pZ >> 8;
        // descend to the leaf; pathNode[0] / pathGrid[0] is the leaf's parent
        uint32_t pathNode[kPathRegs], pathGrid[kPathRegs];
        for (uint32_t k = 0; k < kPathRegs; ++k) {
            pathNode[k] = kNone;
            pathGrid[k] = 0;
        }
        uint32_t node = 0, lvl = 0, gref = 0;
        if (active) {
            for (;;) {
                const uint2 rec = t.walk[node];
                if (rec.x == kNone || lvl == uint32_t(kMaxDepth)) {
                    gref = rec.y;
                    break;
                }
                for (uint32_t k = kPathRegs - 1; k > 0; --k) {
                    pathNode[k] = pathNode[k - 1];
                    pathGrid[k] = pathGrid[k - 1];
                }
                pathNode[0] = node;
                pathGrid[0] = rec.y;
                const uint32_t s2 = kMaxDepth - lvl - 1;
                node = rec.x + ((((X >> s2) & 1u) << 2) | (((Y >> s2) & 1u) << 1) | ((Z >> s2) & 1u));
                lvl += 1;
            }
            t.pointLeaf[i] = node | (lvl << kTopLevelShift);  // for the insertion (the tree stays as it is)
        }
        // The climb, from the leaf up: one loop with a single entry and a
        // uniform trip count, so every lane runs each step's staging together. A
        // node without a grid (the leaf, unless it is the root) is passed over; at
        // one with a grid the lane peeks its cell (a stale peek only costs an
        // atomic), sets it, and climbs on while the bit was new: a set bit at
        // level L implies the point's cells at every gridded ancestor are set (or
        // are being set by the lane that set it).
        for (uint32_t step = 0; step <= uint32_t(kMaxDepth); ++step) {
            if (__ballot(active) == 0) break;
            const uint32_t vi = cellIndex(pX, pY, pZ, lvl);
            const uint32_t mask = 1u << (vi & 31u);
            bool isNew = false, set = false;
            if (active && gref != 0) {
                uint32_t* word = cellWord(t, gref, vi);
                isNew = (*word & mask) == 0 && (atomicOr(word, mask) & mask) == 0;
                set = !isNew;
            }
            stageLevel(t, st, staged, isNew, lvl, node, vi, X, Y, Z, col, size);
            if (set || lvl == 0) active = false;
            if (active) {
                lvl -= 1;
                if (pathNode[0] != kNone) {
                    node = pathNode[0];
                    gref = pathGrid[0];
                    for (uint32_t k = 0; k + 1 < kPathRegs; ++k) {
                        pathNode[k] = pathNode[k + 1];
                        pathGrid[k] = pathGrid[k + 1];
                    }
                    pathNode[kPathRegs - 1] = kNone;
                } else {
                    node = t.parent[node];
                    gref = t.walk[node].y;
                }
            }
        }
    }
    flushStage(t, st, staged);
}

// Exclusive prefix over the backlog segments' fills, clipped to the segment size,
// in LDS; prefix[kBacklogShards] is the number of entries held. Every thread of
// the block calls it.
__device__ __forceinline__ void backlogPrefix(const DevTree& t, uint32_t* prefix) {
    if (threadIdx.x < 64) {
        static_assert(kBacklogShards == 64, "one lane per segment");
        const uint32_t lane = threadIdx.x;
        const uint32_t own = min(t.backlogFill[lane * kShardStride], t.backlogSegCap);
        uint32_t v = own;
        for (uint32_t off = 1; off < 64; off <<= 1) {
            const uint32_t o = __shfl_up(v, off);
            if (lane >= off) v += o;
        }
        prefix[lane + 1] = v;
        if (lane == 0) prefix[0] = 0;
    }
    __syncthreads();
}

// Backlog entry i (< prefix[kBacklogShards]) -> its slot in the segmented arrays.
__device__ __forceinline__ uint64_t backlogSlot(const DevTree& t, const uint32_t* prefix, uint32_t i) {
    uint32_t lo = 0, hi = kBacklogShards;  // prefix[lo] <= i < prefix[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (prefix[mid] <= i) lo = mid;
        else hi = mid;
    }
    return uint64_t(lo) * t.backlogSegCap + (i - prefix[lo]);
}

// Exclusive prefix sum over the wave's lanes; `total` gets the wave's sum.
template <typename T>
__device__ __forceinline__ T waveExclusiveSum(T v, T& total) {
    const uint32_t lane = laneId();
    T inc = v;
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const T o = __shfl_up(inc, off);
        if (lane >= off) inc += o;
    }
    total = __shfl(inc, 63);
    return inc - v;
}

// The lane that owns item e of a wave whose lanes own consecutive ranges
// starting at `before` (an exclusive prefix sum): the last lane whose range
// starts at or below e. Every lane of the wave must take part.
__device__ __forceinline__ uint32_t ownerLane(uint32_t before, uint32_t e) {
    uint32_t o = 0;
    for (uint32_t step = 32; step; step >>= 1)
        if (uint32_t(__shfl(int(before), int(o + step))) <= e) o += step;
    return o;
}

// allocatePointChunks + allocateVoxelChunks (remolod_octree.cu:742-792, :885-916).
// Point chunks come from the recycled queue while the in-use count is below the
// pool's high-water mark, from the persistent store beyond it. A wave reserves
// its lanes' chunks with one atomic on the chunk counter and one on the store's
// offset (one per node took a returning atomic on each of these two addresses).
__global__ __launch_bounds__(kBlock) void kAllocChunks(DevTree t) {
    Globals* g = t.g;
    if (g->stop) return;
    const uint32_t n = g->numNodes;
    const int64_t poolSize = g->chunkPoolSize;
    const uint32_t lane = laneId();
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < n; base += stride) {
        const uint32_t i = base + lane;
        const bool valid = i < n;
        // point chunks a leaf lacks for its counter
        uint32_t have = 0, add = 0;
        if (valid && t.firstChild[i] < 0) {
            const uint32_t np = t.numPoints[i], cnt = t.counter[i];
            if (np < cnt) {
                have = (np + kPointsPerChunk - 1) / kPointsPerChunk;
                const uint32_t need = (cnt + kPointsPerChunk - 1) / kPointsPerChunk;
                if (need > have) add = need - have;
            }
        }
        // voxel chunks a grid lacks; this batch's backlog entries of the grid take
        // the slots after the stored voxels, segment after segment (the flushes
        // ranked them per segment); the 64 counts are read as 16 vectors before
        // any store
        const uint32_t gi = valid ? t.gridNo[i] : kNone;
        uint32_t vHave = 0, vAdd = 0;
        if (gi != kNone) {
            uint4* c4 = reinterpret_cast<uint4*>(t.voxelShardCount + uint64_t(gi) * kBacklogShards);
            uint4 cnt[kBacklogShards / 4];
            for (uint32_t q = 0; q < kBacklogShards / 4; ++q) cnt[q] = c4[q];
            const uint32_t stored = t.numVoxelsStored[i];
            uint4* s4 = reinterpret_cast<uint4*>(t.voxelShardStart + uint64_t(gi) * kBacklogShards);
            uint32_t added = 0;
            for (uint32_t q = 0; q < kBacklogShards / 4; ++q) {
                const uint4 c = cnt[q];
                if ((c.x | c.y | c.z | c.w) == 0) continue;
                const uint32_t s0 = stored + added;
                const uint32_t s1 = s0 + c.x, s2 = s1 + c.y, s3 = s2 + c.z;
                added += c.x + c.y + c.z + c.w;
                s4[q] = make_uint4(s0, s1, s2, s3);
                c4[q] = make_uint4(0, 0, 0, 0);
            }
            uint32_t nv = t.numVoxels[i];
            if (added) {
                nv += added;
                t.numVoxels[i] = nv;
                t.numVoxelsStored[i] = stored + added;
            }
            const uint32_t need = (nv + kPointsPerChunk - 1) / kPointsPerChunk;
            vHave = t.voxelChunks[gi];
            if (need > vHave) vAdd = need - vHave;
        }
        // the wave's chunk indices, then its fresh memory (indices at or above the
        // pool size, and every voxel chunk)
        uint32_t addTotal;
        const uint32_t addBefore = waveExclusiveSum(add, addTotal);
        int64_t waveFirst = 0;
        if (addTotal) {
            if (lane == 0)
                waveFirst = int64_t(atomicAdd((unsigned long long*)&g->numAllocatedChunks,
                                              (unsigned long long)addTotal));
            waveFirst = __shfl(waveFirst, 0);
        }
        const int64_t first = waveFirst + addBefore;
        const int64_t freshFrom = first > poolSize ? first : poolSize;
        const int64_t last = first + add;
        const uint32_t fresh = (add && last > freshFrom) ? uint32_t(last - freshFrom) : 0u;
        unsigned long long bytesTotal;
        const unsigned long long bytesBefore =
            waveExclusiveSum((unsigned long long)(fresh + vAdd) * kChunkAllocBytes, bytesTotal);
        unsigned long long waveOff = 0;
        if (bytesTotal) {
            if (lane == 0) waveOff = atomicAdd((unsigned long long*)&g->arenaOffset, bytesTotal);
            waveOff = __shfl(waveOff, 0);
        }
        // the references, a lane per chunk: the wave's point chunks are numbered
        // 0..addTotal-1 in lane order, then its voxel chunks likewise (a loop over
        // its own chunks per lane put a queue load and a store per chunk in one
        // dependent chain)
        const uint64_t offFresh = waveOff + bytesBefore;
        for (uint32_t e0 = 0; e0 < addTotal; e0 += 64) {
            const uint32_t e = e0 + lane;
            const uint32_t o = ownerLane(addBefore, e);
            const uint32_t oBefore = __shfl(addBefore, int(o)), oHave = __shfl(have, int(o));
            const uint64_t oOff = __shfl(offFresh, int(o));
            if (e < addTotal) {
                const uint32_t k = e - oBefore;
                const int64_t qi = waveFirst + int64_t(e);
                uint32_t ref = 0;
                if (qi >= poolSize) {
                    // the owner's fresh chunks are its last ones, in order
                    const int64_t oFirst = waveFirst + int64_t(oBefore);
                    const int64_t oFreshFrom = oFirst > poolSize ? oFirst : poolSize;
                    const uint64_t at = oOff + uint64_t(qi - oFreshFrom) * kChunkAllocBytes;
                    if (at + kChunkAllocBytes <= t.arenaCapacity)
                        ref = uint32_t(at >> 4);
                    else
                        g->allocOverflow = 1;
                } else if (uint64_t(qi) < t.chunkQueueCap) {
                    ref = t.chunkQueue[qi];
                }
                const uint32_t slot = oHave + k;
                if (slot < kPointDir)
                    t.pointDir[uint64_t(base + o) * kPointDir + slot] = ref;
                else
                    g->dirOverflow = 1;
            }
        }
        uint32_t vTotal;
        const uint32_t vBefore = waveExclusiveSum(vAdd, vTotal);
        const uint64_t offVoxel = offFresh + uint64_t(fresh) * kChunkAllocBytes;
        for (uint32_t e0 = 0; e0 < vTotal; e0 += 64) {
            const uint32_t e = e0 + lane;
            const uint32_t o = ownerLane(vBefore, e);
            const uint32_t oBefore = __shfl(vBefore, int(o)), oHave = __shfl(vHave, int(o));
            const uint32_t oGrid = __shfl(gi, int(o));
            const uint64_t oOff = __shfl(offVoxel, int(o));
            if (e < vTotal) {
                const uint32_t k = e - oBefore;
                const uint64_t at = oOff + uint64_t(k) * kChunkAllocBytes;
                const uint32_t slot = oHave + k;
                if (at + kChunkAllocBytes > t.arenaCapacity) {
                    g->allocOverflow = 1;
                } else if (slot < kVoxelDir) {
                    t.voxelDir[uint64_t(oGrid) * kVoxelDir + slot] = uint32_t(at >> 4);
                } else {
                    g->dirOverflow = 1;
                }
            }
        }
        if (vAdd) {
            t.voxelChunks[gi] = vHave + vAdd;
        }
    }
}

// insertPoints (remolod_octree.cu:794-883): slot = atomicAdd(numPoints), taken
// once per (leaf, octant) key of a tile: the tile ranks its points in LDS, one
// lane per key reserves the range (and grows the stored-octant histogram), then
// every point lands at base + rank.
// Eight points per lane: a 2,048-point tile takes half the reservations of a
// 1,024-point one on the numPoints of the leaves a batch fills (48 -> 39 us per
// 1M-point batch; 12 per lane is no faster, 16 halves the occupancy).
constexpr uint32_t kInsertPerThread = 8;
constexpr uint32_t kInsertTile = kBlock * kInsertPerThread;
static_assert(kInsertTile <= kAggSlots, "a tile's keys fit the aggregation table");

__global__ __launch_bounds__(kBlock) void kInsertPoints(DevTree t, BatchRef b, uint32_t leavesKnown) {
    __shared__ BlockAgg agg;
    const uint32_t batchSize = batchSizeOf(t, b);
    Globals* g = t.g;
    if (g->stop) return;
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        // allocatePointChunks' closing update (remolod_octree.cu:789-791)
        if (g->numAllocatedChunks > g->chunkPoolSize) g->chunkPoolSize = g->numAllocatedChunks;
    }
    const uint32_t total = batchSize + spilledCount(t);
    const TileRange tr = myTiles<kInsertTile>(total);
    const float size = t.octreeSize;
    for (uint32_t tile = tr.begin; tile < tr.end; ++tile) {
        aggClear(agg);
        __syncthreads();
        remo_point p[kInsertPerThread];
        uint32_t leaf[kInsertPerThread], slotOf[kInsertPerThread], rank[kInsertPerThread];
        uint32_t X[kInsertPerThread], Y[kInsertPerThread], Z[kInsertPerThread], level[kInsertPerThread];
        bool valid[kInsertPerThread];
        for (uint32_t j = 0; j < kInsertPerThread; ++j) {
            const uint32_t i = tile * kInsertTile + j * kBlock + threadIdx.x;
            valid[j] = i < total;
            X[j] = Y[j] = Z[j] = 0;
            if (!valid[j]) continue;
            p[j] = streamPoint(t, b, batchSize, i);
            X[j] = quant28(p[j].x, size) >> 8;
            Y[j] = quant28(p[j].y, size) >> 8;
            Z[j] = quant28(p[j].z, size) >> 8;
        }
        if (leavesKnown) {  // (uniform) the voxel sampler's descent of the same points
            for (uint32_t j = 0; j < kInsertPerThread; ++j) {
                const uint32_t v = valid[j] ? t.pointLeaf[tile * kInsertTile + j * kBlock + threadIdx.x] : 0u;
                leaf[j] = v & kTopNodeMask;
                level[j] = v >> kTopLevelShift;
            }
        } else {
            descendN(t.firstChild, t.topCell, X, Y, Z, valid, leaf, level);
        }
        for (uint32_t j = 0; j < kInsertPerThread; ++j) {
            slotOf[j] = kEmpty32;
            if (!valid[j]) {
                leaf[j] = kNone;
                continue;
            }
            const uint32_t key =
                leaf[j] * 8u + (level[j] < uint32_t(kMaxDepth) ? octantAt(X[j], Y[j], Z[j], level[j]) : 0u);
            slotOf[j] = aggInsert(agg, key, &rank[j]);
            if (slotOf[j] == kEmpty32) {
                // table full: take a slot directly
                rank[j] = atomicAdd(&t.numPoints[leaf[j]], 1u);
                atomicAdd(&t.storedOct[key], 1u);
            }
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < kAggSlots; k += blockDim.x) {
            const uint32_t key = agg.key[k];
            if (key == kEmpty32) continue;
            const uint32_t c = agg.cnt[k];
            agg.base[k] = atomicAdd(&t.numPoints[key >> 3], c);
            atomicAdd(&t.storedOct[key], c);
        }
        __syncthreads();
        for (uint32_t j = 0; j < kInsertPerThread; ++j) {
            if (leaf[j] == kNone) continue;
            const uint32_t slot = slotOf[j] == kEmpty32 ? rank[j] : agg.base[slotOf[j]] + rank[j];
            const uint32_t k = slot / kPointsPerChunk;
            const uint32_t ref = k < kPointDir ? t.pointDir[uint64_t(leaf[j]) * kPointDir + k] : 0u;
            if (ref != 0)
                chunkPtr(t.arena, ref)[slot % kPointsPerChunk] = p[j];
            else
                g->dirOverflow = 1;
        }
        __syncthreads();
    }
}

// insertVoxels (remolod_octree.cu:918-938): every backlog entry goes to slot
// start(grid, segment) + rank, both taken before (the voxel flush, kAllocChunks).
// An entry carries its node's grid number, not the node (one dependent load
// fewer between the entry and its store).
__global__ __launch_bounds__(kBlock) void kInsertVoxels(DevTree t, BatchRef b, uint32_t closeBatch) {
    __shared__ uint32_t prefix[kBacklogShards + 1];
    Globals* g = t.g;
    if (g->stop) return;
    backlogPrefix(t, prefix);
    const uint32_t total = prefix[kBacklogShards];
    __syncthreads();  // every thread has read the counters before block 0 ends the batch
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (closeBatch) {
            // the batch loop's closing update (remolod_octree.cu:1223-1226)
            const uint64_t spilled = spilledCount(t);
            g->batchletIndex += 1;
            g->numPointsProcessed += *b.sizePtr;
            g->batches += 1;
            g->spilledPoints += spilled;
            g->totBatches += 1;
            g->totSpilledPoints += spilled;
        }
        g->numBacklog = total;
        g->newVoxels += total;
        g->totNewVoxels += total;
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const uint64_t at = backlogSlot(t, prefix, i);
        const uint32_t gi = t.backlogGrid[at];
        if (gi == kNone) continue;  // not reached: sampled nodes have grids
        const uint32_t seg = uint32_t(at / t.backlogSegCap);
        const uint32_t slot = t.voxelShardStart[uint64_t(gi) * kBacklogShards + seg] + t.backlogRank[at];
        const uint32_t k = slot / kPointsPerChunk;
        const uint32_t ref = k < kVoxelDir ? t.voxelDir[uint64_t(gi) * kVoxelDir + k] : 0u;
        if (ref != 0)
            chunkPtr(t.arena, ref)[slot % kPointsPerChunk] = t.backlogPts[at];
        else
            g->dirOverflow = 1;
    }
}

// The per-node stats sweep (remolod_octree.cu:1248-1319), in two passes: every
// workgroup reduces its nodes (wave shuffles, then LDS) into its own partial
// record, and one workgroup adds the records into Globals. One atomic per wave
// on the Globals line, as this once did, serialised ~2,000 waves on one address.
// Partial record words: 0 inner, 1 leaves, 2 nonempty, 3 chunksP, 4 chunksV,
// 5 underflows, 6-7 points, 8-9 voxels, 10 max points, 11 max depth.
struct StatsAcc {
    uint32_t inner, leaves, nonempty, chunksP, chunksV, under, maxPts, maxDepth;
    uint64_t pts, vox;
};

__device__ __forceinline__ void statsWaveReduce(StatsAcc& a) {
    for (int off = 32; off > 0; off >>= 1) {
        a.inner += __shfl_down(a.inner, off);
        a.leaves += __shfl_down(a.leaves, off);
        a.nonempty += __shfl_down(a.nonempty, off);
        a.chunksP += __shfl_down(a.chunksP, off);
        a.chunksV += __shfl_down(a.chunksV, off);
        a.under += __shfl_down(a.under, off);
        a.pts += __shfl_down(a.pts, off);
        a.vox += __shfl_down(a.vox, off);
        a.maxPts = max(a.maxPts, __shfl_down(a.maxPts, off));
        a.maxDepth = max(a.maxDepth, __shfl_down(a.maxDepth, off));
    }
}

// Lane 0 of every wave holds its wave's totals; the block's totals go to wave 0.
__device__ __forceinline__ void statsBlockReduce(StatsAcc& a) {
    __shared__ StatsAcc waves[kBlock / 64];
    statsWaveReduce(a);
    const uint32_t w = threadIdx.x / 64;
    if (laneId() == 0) waves[w] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t k = 1; k < kBlock / 64; ++k) {
            const StatsAcc& o = waves[k];
            a.inner += o.inner;
            a.leaves += o.leaves;
            a.nonempty += o.nonempty;
            a.chunksP += o.chunksP;
            a.chunksV += o.chunksV;
            a.under += o.under;
            a.pts += o.pts;
            a.vox += o.vox;
            a.maxPts = max(a.maxPts, o.maxPts);
            a.maxDepth = max(a.maxDepth, o.maxDepth);
        }
    }
}

// The per-node stats sweep in one launch: every workgroup reduces its nodes in
// LDS and writes one partial record; the workgroup that finishes last (a count
// on Globals, the partials released before it and acquired after) adds the
// records up into Globals and clears the count.
__global__ __launch_bounds__(kBlock) void kStatsSweep(DevTree t) {
    __shared__ bool last;
    Globals* g = t.g;
    const uint32_t n = g->numNodes;
    uint32_t inner = 0, leaves = 0, nonempty = 0, maxPts = 0, chunksP = 0, chunksV = 0, under = 0,
             maxDepth = 0;
    uint64_t pts = 0, vox = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (isFreeSlot(t, i)) continue;
        maxDepth = max(maxDepth, t.level[i]);
        if (t.firstChild[i] < 0) {
            const uint32_t np = t.numPoints[i];
            leaves++;
            pts += np;
            chunksP += (np + kPointsPerChunk - 1) / kPointsPerChunk;
            maxPts = max(maxPts, np);
            if (t.counter[i] < np) under++;
            if (np > 0) nonempty++;
        } else {
            const uint32_t nv = t.numVoxels[i];
            inner++;
            vox += nv;
            chunksV += (nv + kPointsPerChunk - 1) / kPointsPerChunk;
        }
    }
    StatsAcc a{inner, leaves, nonempty, chunksP, chunksV, under, maxPts, maxDepth, pts, vox};
    statsBlockReduce(a);
    if (threadIdx.x == 0) {
        uint32_t* p = t.statsPartials + uint64_t(blockIdx.x) * kStatsWords;
        p[0] = a.inner;
        p[1] = a.leaves;
        p[2] = a.nonempty;
        p[3] = a.chunksP;
        p[4] = a.chunksV;
        p[5] = a.under;
        p[6] = uint32_t(a.pts);
        p[7] = uint32_t(a.pts >> 32);
        p[8] = uint32_t(a.vox);
        p[9] = uint32_t(a.vox >> 32);
        p[10] = a.maxPts;
        p[11] = a.maxDepth;
        __threadfence();  // the partial before the count
        last = atomicAdd(&g->statsDone, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;  // uniform over the workgroup
    __threadfence();    // every other workgroup's partial after the count
    StatsAcc r{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t b = threadIdx.x; b < gridDim.x; b += blockDim.x) {
        const uint32_t* p = t.statsPartials + uint64_t(b) * kStatsWords;
        r.inner += p[0];
        r.leaves += p[1];
        r.nonempty += p[2];
        r.chunksP += p[3];
        r.chunksV += p[4];
        r.under += p[5];
        r.pts += uint64_t(p[6]) | (uint64_t(p[7]) << 32);
        r.vox += uint64_t(p[8]) | (uint64_t(p[9]) << 32);
        r.maxPts = max(r.maxPts, p[10]);
        r.maxDepth = max(r.maxDepth, p[11]);
    }
    statsBlockReduce(r);
    if (threadIdx.x == 0) {
        g->sInner = r.inner;
        g->sLeaves = r.leaves;
        g->sNonempty = r.nonempty;
        g->sChunksPoints = r.chunksP;
        g->sChunksVoxels = r.chunksV;
        g->sUnderflows = r.under;
        g->sPoints = r.pts;
        g->sVoxels = r.vox;
        g->sMaxPoints = r.maxPts;
        g->sMaxDepth = r.maxDepth;
        g->statsDone = 0;
    }
}

}  // namespace

// ---- launchers --------------------------------------------------------------

void launchResetTree(const DevTree& t, hipStream_t s) {
    hipLaunchKernelGGL(kResetRoot, dim3(1), dim3(1), 0, s, t);
    hipLaunchKernelGGL(kZeroRootGrid, dim3(64), dim3(kBlock), 0, s, t);
}

void launchBeginStep(const DevTree& t, hipStream_t s) {
    hipLaunchKernelGGL(kBeginStep, dim3(1), dim3(1), 0, s, t);
}

void launchBeginBatch(const DevTree& t, BatchRef b, bool checkBudget, uint64_t budgetTicks, bool beginStep,
                      hipStream_t s) {
    hipLaunchKernelGGL(kBeginBatch, dim3(1), dim3(kBacklogShards), 0, s, t, b, t.arenaCapacity, checkBudget ? 1u : 0u,
                       budgetTicks, beginStep ? 1u : 0u);
}

void launchCount(const DevTree& t, BatchRef b, uint32_t round, const LaunchShape& ls, hipStream_t s) {
    hipLaunchKernelGGL(kCount, dim3(ls.pointBlocks), dim3(kBlock), 0, s, t, b, round);
}

void launchSplitDecide(const DevTree& t, uint32_t round, const LaunchShape& ls, hipStream_t s) {
    hipLaunchKernelGGL(kSplitDecide, dim3(ls.nodeBlocks), dim3(kBlock), 0, s, t, round);
}

void launchSplitApply(const DevTree& t, const LaunchShape& ls, hipStream_t s) {
    (void)ls;
    hipLaunchKernelGGL(kSplitApply, dim3(1), dim3(1024), 0, s, t);
}

void launchSpillAndZero(const DevTree& t, const LaunchShape& ls, hipStream_t s) {
    (void)ls;
    hipLaunchKernelGGL(kSpillAndZero, dim3(1024), dim3(kBlock), 0, s, t, 1u);
}

void launchFixedExpand(const DevTree& t, BatchRef b, const LaunchShape& ls, hipStream_t s) {
    const uint32_t D = t.fixedDepth;
    const uint32_t entries = uint32_t(fixedPyramidEntries(D));
    hipLaunchKernelGGL(kFixedClear, dim3(std::min<uint32_t>(4096, (entries + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       s, t, entries);
    hipLaunchKernelGGL(kFixedBucket, dim3(ls.pointBlocks), dim3(kBlock), 0, s, t, b);
    for (int level = int(D) - 1; level >= 0; --level) {
        const uint32_t cells = 1u << (3 * level);
        hipLaunchKernelGGL(kFixedPyramid, dim3(std::min<uint32_t>(4096, (cells + kBlock - 1) / kBlock)), dim3(kBlock),
                           0, s, t, uint32_t(level));
    }
    for (uint32_t level = 0; level < D; ++level) {
        hipLaunchKernelGGL(kFixedRound, dim3(1), dim3(1), 0, s, t);
        hipLaunchKernelGGL(kFixedDecide, dim3(ls.nodeBlocks), dim3(kBlock), 0, s, t, level);
        hipLaunchKernelGGL(kSplitApply, dim3(1), dim3(1024), 0, s, t);
        hipLaunchKernelGGL(kSpillAndZero, dim3(1024), dim3(kBlock), 0, s, t, 0u);
    }
    hipLaunchKernelGGL(kFixedSeed, dim3(ls.nodeBlocks), dim3(kBlock), 0, s, t);
}

void launchVoxelSample(const DevTree& t, BatchRef b, const LaunchShape& ls, hipStream_t s) {
    hipLaunchKernelGGL(kVoxelSample, dim3(ls.pointBlocks), dim3(kBlock), 0, s, t, b);
}

void launchAllocChunks(const DevTree& t, const LaunchShape& ls, hipStream_t s) {
    hipLaunchKernelGGL(kAllocChunks, dim3(ls.nodeBlocks), dim3(kBlock), 0, s, t);
}

void launchInsertPoints(const DevTree& t, BatchRef b, const LaunchShape& ls, hipStream_t s, bool leavesKnown) {
    hipLaunchKernelGGL(kInsertPoints, dim3(ls.pointBlocks), dim3(kBlock), 0, s, t, b, leavesKnown ? 1u : 0u);
}

void launchInsertVoxels(const DevTree& t, BatchRef b, const LaunchShape& ls, hipStream_t s, bool closeBatch) {
    hipLaunchKernelGGL(kInsertVoxels, dim3(ls.pointBlocks), dim3(kBlock), 0, s, t, b, closeBatch ? 1u : 0u);
}

// ---------------------------------------------------------------------------
// Refinement's local deepening: leaves split outside Update, their stored points
// redistributed to the children through the spill list, exactly as a split
// during a batch would (kSplitApply, kSpillAndZero, then voxel sampling, chunk
// acquisition and insertion of the spilled points with an empty batch).

struct DeepenCandidate {
    uint32_t numPoints, level, x, y, z, node;
};

// count[0]: candidates found (may pass cap: only the first cap are written);
// count[1]: 0, the size of the empty batch the split path reads.
__global__ void kDeepenClear(uint32_t* count) {
    if (threadIdx.x < 2) count[threadIdx.x] = 0;
}

__global__ __launch_bounds__(kBlock) void kDeepenList(DevTree t, DeepenCandidate* out, uint32_t* count, uint32_t cap,
                                                      uint32_t minPoints) {
    const uint32_t n = t.g->numNodes;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (t.firstChild[i] >= 0 || t.level[i] >= uint32_t(kMaxDepth)) continue;
        const uint32_t np = t.numPoints[i];
        if (np < minPoints || t.counter[i] != np) continue;
        const uint32_t at = atomicAdd(count, 1u);
        if (at < cap) out[at] = DeepenCandidate{np, t.level[i], t.nx[i], t.ny[i], t.nz[i], i};
    }
}

// The chosen splits become this round's candidates; a fresh round of counters.
__global__ void kDeepenBegin(DevTree t, uint32_t numChosen) {
    Globals* g = t.g;
    t.backlogFill[threadIdx.x * kShardStride] = 0;  // kBacklogShards threads
    if (threadIdx.x != 0) return;
    g->numCandidates = numChosen;
    g->numAccepted = 0;
    g->numPending = 0;
    g->numSpilled = 0;
    g->spillLimit = kNone;
    g->numBacklog = 0;
    g->droppedVoxels = 0;
    g->stop = 0;
}

// ---------------------------------------------------------------------------
// Refinement's collapse (no reference code; README.md:81,
// wiki/02_KernelPipeline.md:238): an inner node whose 8 children are leaves is
// merged into one leaf. Its children's points pass through the spill list and
// are inserted as a batch of no new points would be (its counter is their count,
// kAllocChunks gives it chunks, kInsertPoints stores them); the children's point
// chunks go back to the chunk queue, the node's grid to the grid pool and the
// grid's voxel chunks to the chunk queue above its high-water mark (so the pool
// of recyclable chunks grows by them), and the children's 8-node block to the
// block pool. kSplitApply takes pooled blocks and grids before new ones.

// Inner nodes below the root whose children are all leaves with every counted
// point stored and no grid, holding at most maxPoints points together.
__global__ __launch_bounds__(kBlock) void kCollapseList(DevTree t, DeepenCandidate* out, uint32_t* count,
                                                        uint32_t cap, uint32_t maxPoints) {
    const uint32_t n = t.g->numNodes;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int32_t fc = t.firstChild[i];
        if (fc < 0 || i == 0) continue;  // free slots are leaves
        uint32_t total = 0;
        bool ok = true;
        for (uint32_t c = 0; c < 8; ++c) {
            const uint32_t ch = uint32_t(fc) + c;
            const uint32_t np = t.numPoints[ch];
            if (t.firstChild[ch] >= 0 || t.counter[ch] != np || t.gridNo[ch] != kNone) ok = false;
            total += np;
        }
        if (!ok || total > maxPoints) continue;
        const uint32_t at = atomicAdd(count, 1u);
        if (at < cap) out[at] = DeepenCandidate{total, t.level[i], t.nx[i], t.ny[i], t.nz[i], i};
    }
}

// Per chosen node: its ranges in the spill list (its children's points), in the
// chunk queue (their chunks, below the in-use count it lowers) and above the
// queue's high-water mark (its grid's voxel chunks, which raise it), recorded
// as (node, spill offset, queue top, pool position). Nothing is moved yet.
__global__ __launch_bounds__(kBlock) void kCollapsePlan(DevTree t, uint32_t numChosen) {
    Globals* g = t.g;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < numChosen; k += gridDim.x * blockDim.x) {
        const uint32_t node = t.candidates[k];
        const uint32_t fc = uint32_t(t.firstChild[node]);
        const uint32_t gi = t.gridNo[node];
        uint32_t np[8];
        for (uint32_t c = 0; c < 8; ++c) np[c] = t.numPoints[fc + c];
        // (a grid's directory holds every chunk its 128^3 cells can need)
        const uint32_t vChunks = gi != kNone ? min(t.voxelChunks[gi], kVoxelDir) : 0u;
        const uint32_t nv = t.numVoxels[node];
        if (nv) atomicAdd(&g->droppedVoxels, nv);
        uint32_t total = 0, chunks = 0;
        for (uint32_t c = 0; c < 8; ++c) {
            total += np[c];
            chunks += (np[c] + kPointsPerChunk - 1) / kPointsPerChunk;
        }
        const uint32_t spillAt = total ? atomicAdd(&g->numSpilled, total) : 0u;
        const int64_t top = chunks ? int64_t(atomicAdd((unsigned long long*)&g->numAllocatedChunks,
                                                       (unsigned long long)(-int64_t(chunks))))
                                   : 0;
        const int64_t poolAt =
            vChunks ? int64_t(atomicAdd((unsigned long long*)&g->chunkPoolSize, (unsigned long long)vChunks)) : 0;
        const bool fits = uint64_t(spillAt) + total <= t.spillCap;  // the caller sizes the choice to fit
        if (!fits) {
            g->spillOverflow = 1;
            atomicMin(&g->spillLimit, spillAt);
        }
        t.acceptedSpill[4 * k] = node;
        t.acceptedSpill[4 * k + 1] = fits ? spillAt : kNone;
        t.acceptedSpill[4 * k + 2] = uint32_t(top);
        t.acceptedSpill[4 * k + 3] = uint32_t(poolAt);
    }
}

// The children's points into the spill list and their chunks into the queue, one
// (chosen node, child, chunk) per block step; then the grids' voxel chunks.
constexpr uint32_t kCollapseChunkSlots = 64;  // >= 50 chunks of a child (at most 50,000 points)

__global__ __launch_bounds__(kBlock) void kCollapseMove(DevTree t, uint32_t numChosen) {
    const uint32_t pairs = numChosen * 8 * kCollapseChunkSlots;
    for (uint32_t pair = blockIdx.x; pair < pairs; pair += gridDim.x) {
        const uint32_t k = pair / (8 * kCollapseChunkSlots);
        const uint32_t c = (pair / kCollapseChunkSlots) % 8, j = pair % kCollapseChunkSlots;
        const uint32_t node = t.acceptedSpill[4 * k];
        const uint32_t off = t.acceptedSpill[4 * k + 1];
        const uint32_t fc = uint32_t(t.firstChild[node]);
        // this child's place among its siblings' points and chunks
        uint32_t pointsBefore = 0, chunksBefore = 0, np = 0;
        for (uint32_t q = 0; q < 8; ++q) {
            const uint32_t m = t.numPoints[fc + q];
            if (q < c) {
                pointsBefore += m;
                chunksBefore += (m + kPointsPerChunk - 1) / kPointsPerChunk;
            } else if (q == c) {
                np = m;
            }
        }
        if (j * kPointsPerChunk >= np) continue;
        const uint32_t ref = t.pointDir[uint64_t(fc + c) * kPointDir + j];
        if (threadIdx.x == 0) {
            const int64_t qi = int64_t(t.acceptedSpill[4 * k + 2]) - 1 - int64_t(chunksBefore + j);
            if (qi >= 0 && uint64_t(qi) < t.chunkQueueCap) t.chunkQueue[qi] = ref;
        }
        if (off == kNone) continue;
        const uint32_t m = min(kPointsPerChunk, np - j * kPointsPerChunk);
        const remo_point* src = chunkPtr(t.arena, ref);
        remo_point* dst = t.spill + off + pointsBefore + j * kPointsPerChunk;
        for (uint32_t q = threadIdx.x; q < m; q += blockDim.x) dst[q] = src[q];
    }
    const uint64_t slots = uint64_t(numChosen) * kVoxelDir;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < slots;
         i += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t k = uint32_t(i / kVoxelDir), v = uint32_t(i % kVoxelDir);
        const uint32_t gi = t.gridNo[t.acceptedSpill[4 * k]];
        if (gi == kNone || v >= t.voxelChunks[gi]) continue;
        const uint64_t qi = uint64_t(t.acceptedSpill[4 * k + 3]) + v;
        if (qi < t.chunkQueueCap) t.chunkQueue[qi] = t.voxelDir[uint64_t(gi) * kVoxelDir + v];
    }
}

// One workgroup per chosen node: the children become free slots, the node a
// leaf waiting for its points (counter = their count), its grid and block go to
// the pools, and the top-level descent index enters at the node again.
__global__ __launch_bounds__(kBlock) void kCollapseFinish(DevTree t, uint32_t numChosen, uint32_t* refPoints,
                                                          uint32_t* refVoxels) {
    __shared__ uint32_t npS[8];
    Globals* g = t.g;
    for (uint32_t k = blockIdx.x; k < numChosen; k += gridDim.x) {
        const uint32_t node = t.acceptedSpill[4 * k];
        const uint32_t fc = uint32_t(t.firstChild[node]);
        if (threadIdx.x < 8) npS[threadIdx.x] = t.numPoints[fc + threadIdx.x];
        __syncthreads();
        if (threadIdx.x < 8) {
            const uint32_t c = fc + threadIdx.x;
            const uint32_t used = (npS[threadIdx.x] + kPointsPerChunk - 1) / kPointsPerChunk;
            for (uint32_t j = 0; j < used && j < kPointDir; ++j) t.pointDir[uint64_t(c) * kPointDir + j] = 0;
            t.firstChild[c] = -1;
            t.walk[c] = make_uint2(kNone, 0u);
            t.parent[c] = kNone;  // a free slot from here on
            t.counter[c] = 0;
            t.numPoints[c] = 0;
            t.numVoxels[c] = 0;
            t.numVoxelsStored[c] = 0;
            t.gridNo[c] = kNone;
            t.pending[c] = 0;
            uint4* so = reinterpret_cast<uint4*>(t.storedOct + 8ull * c);
            so[0] = so[1] = make_uint4(0, 0, 0, 0);
            uint4* bo = reinterpret_cast<uint4*>(t.batchOct + 8ull * c);
            bo[0] = bo[1] = make_uint4(0, 0, 0, 0);
            t.accums[c] = remo_node_accum{};
            t.dirtyListed[c] = 0;
            if (refPoints) refPoints[c] = 0;
            if (refVoxels) refVoxels[c] = kNone;
        }
        const uint32_t lvl = t.level[node];
        if (threadIdx.x == 8) {
            uint32_t total = 0;
            for (uint32_t c = 0; c < 8; ++c) total += npS[c];
            const uint32_t gi = t.gridNo[node];
            t.firstChild[node] = -1;
            t.walk[node] = make_uint2(kNone, 0u);
            t.gridNo[node] = kNone;
            t.numVoxels[node] = 0;
            t.numVoxelsStored[node] = 0;
            t.counter[node] = total;
            t.numPoints[node] = 0;
            t.pending[node] = 0;
            uint4* so = reinterpret_cast<uint4*>(t.storedOct + 8ull * node);
            so[0] = so[1] = make_uint4(0, 0, 0, 0);
            uint4* bo = reinterpret_cast<uint4*>(t.batchOct + 8ull * node);
            bo[0] = bo[1] = make_uint4(0, 0, 0, 0);
            t.accums[node] = remo_node_accum{};
            if (refPoints) refPoints[node] = kNone;  // its parent folds it again
            if (refVoxels) refVoxels[node] = kNone;
            t.freeBlocks[atomicAdd(&g->numFreeBlocks, 1u)] = fc;
            if (gi != kNone) {
                t.voxelChunks[gi] = 0;
                t.freeGrids[atomicAdd(&g->numFreeGrids, 1u)] = gi;
            }
            atomicAdd((unsigned long long*)&g->totCollapsed, 1ull);
        }
        // descents into the node's region enter at the node again
        if (lvl < kTopLevel) {
            const uint32_t sb = kTopLevel - lvl;  // the node spans 2^sb top cells per axis
            const uint32_t X = t.nx[node], Y = t.ny[node], Z = t.nz[node];
            const uint32_t m = (1u << sb) - 1u;
            for (uint32_t q = threadIdx.x; q < (1u << (3 * sb)); q += blockDim.x) {
                const uint32_t cx = (X << sb) | ((q >> (2 * sb)) & m);
                const uint32_t cy = (Y << sb) | ((q >> sb) & m);
                const uint32_t cz = (Z << sb) | (q & m);
                t.topCell[(cx << (2 * kTopLevel)) | (cy << kTopLevel) | cz] = node | (lvl << kTopLevelShift);
            }
        }
        __syncthreads();  // npS is rewritten for the next node
    }
}

void launchCollapseList(const DevTree& t, void* out, uint32_t* count, uint32_t cap, uint32_t maxPoints,
                        const LaunchShape& ls, hipStream_t s) {
    hipLaunchKernelGGL(kDeepenClear, dim3(1), dim3(64), 0, s, count);
    hipLaunchKernelGGL(kCollapseList, dim3(ls.nodeBlocks), dim3(kBlock), 0, s, t, static_cast<DeepenCandidate*>(out),
                       count, cap, maxPoints);
}

void launchCollapse(const DevTree& t, uint32_t numChosen, BatchRef empty, uint32_t* refPoints, uint32_t* refVoxels,
                    const LaunchShape& ls, hipStream_t s) {
    hipLaunchKernelGGL(kDeepenBegin, dim3(1), dim3(kBacklogShards), 0, s, t, numChosen);
    const uint32_t planBlocks = std::max(1u, std::min(256u, (numChosen + kBlock - 1) / kBlock));
    hipLaunchKernelGGL(kCollapsePlan, dim3(planBlocks), dim3(kBlock), 0, s, t, numChosen);
    hipLaunchKernelGGL(kCollapseMove, dim3(1024), dim3(kBlock), 0, s, t, numChosen);
    hipLaunchKernelGGL(kCollapseFinish, dim3(std::max(1u, std::min(numChosen, 1024u))), dim3(kBlock), 0, s, t,
                       numChosen, refPoints, refVoxels);
    hipLaunchKernelGGL(kAllocChunks, dim3(ls.nodeBlocks), dim3(kBlock), 0, s, t);
    hipLaunchKernelGGL(kInsertPoints, dim3(ls.pointBlocks), dim3(kBlock), 0, s, t, empty, 0u);  // (no voxel sampling ran over these points)
    hipLaunchKernelGGL(kInsertVoxels, dim3(ls.pointBlocks), dim3(kBlock), 0, s, t, empty, 0u);
}

void launchDeepenList(const DevTree& t, void* out, uint32_t* count, uint32_t cap, uint32_t minPoints,
                      const LaunchShape& ls, hipStream_t s) {
    hipLaunchKernelGGL(kDeepenClear, dim3(1), dim3(64), 0, s, count);
    hipLaunchKernelGGL(kDeepenList, dim3(ls.nodeBlocks), dim3(kBlock), 0, s, t, static_cast<DeepenCandidate*>(out),
                       count, cap, minPoints);
}

void launchDeepenSplit(const DevTree& t, uint32_t numChosen, BatchRef empty, const LaunchShape& ls, hipStream_t s) {
    hipLaunchKernelGGL(kDeepenBegin, dim3(1), dim3(kBacklogShards), 0, s, t, numChosen);
    hipLaunchKernelGGL(kSplitApply, dim3(1), dim3(1024), 0, s, t);
    hipLaunchKernelGGL(kSpillAndZero, dim3(1024), dim3(kBlock), 0, s, t, 0u);
    hipLaunchKernelGGL(kVoxelSample, dim3(ls.pointBlocks), dim3(kBlock), 0, s, t, empty);
    hipLaunchKernelGGL(kAllocChunks, dim3(ls.nodeBlocks), dim3(kBlock), 0, s, t);
    hipLaunchKernelGGL(kInsertPoints, dim3(ls.pointBlocks), dim3(kBlock), 0, s, t, empty, 1u);
    hipLaunchKernelGGL(kInsertVoxels, dim3(ls.pointBlocks), dim3(kBlock), 0, s, t, empty, 0u);
}

void launchStatsSweep(const DevTree& t, const LaunchShape& ls, hipStream_t s) {
    const uint32_t blocks = std::min(ls.nodeBlocks, kStatsBlocks);
    hipLaunchKernelGGL(kStatsSweep, dim3(blocks), dim3(kBlock), 0, s, t);
}

}  // namespace remo_hip

// remolod_oracle.cpp -- CPU restatement of RemoLOD's progressive LOD path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker the HIP path is compared
// against; nothing in the product links, loads or calls it. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
//
// It is a serial transcription of the reference kernels, phase by phase, with
// IEEE-754 single/double semantics and no FMA contraction (built with
// -ffp-contract=off). Every function cites the reference lines it follows
// (paths relative to the RemoBench repository). Storage differs where storage is
// not observable: a leaf's 1000-point chunk list is a std::vector, but chunk and
// grid allocations are still counted with the reference's byte strides so that
// allocatedBytes_persistent and the memCapacityReached stop are reproduced.
//
// Pinning (what checks this oracle, see tests/test_oracle_*.py):
//   * node tree, counters, occupied cells  -- an independent numpy closed form
//     (oracle/closed_form.py) and hand-derived tiny cases;
//   * quantisation                         -- exact rational arithmetic (fractions);
//   * accumulate sums, Morton watermark    -- numpy textbook sums and a bit-by-bit
//     interleave; fixed point against float128-free integer re-derivation;
//   * persistent-byte accounting           -- the closed-form chunk/grid count.
// Divergences from the reference, applied identically in the HIP path:
//   * a split refused by the node pool does not spill its points (the reference
//     spills them and re-inserts duplicates, remolod_octree.cu:250-321);
//   * spilled points past 3,000,000 are flagged, never dropped
//     (remolod_octree.cu:872-875);
//   * a node at level MAX_DEPTH never splits (the reference's descent cannot
//     reach below it, remolod_octree.cu:182).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#include "../include/remolod.h"

namespace {

// kernels/remolod/remolod_structures.cuh:12-19
constexpr uint32_t kMaxPointsPerNode = 50000;
constexpr uint32_t kPointsPerChunk = 1000;
constexpr uint32_t kGridSize = 128;
constexpr uint32_t kGridWords = kGridSize * kGridSize * kGridSize / 32;
constexpr int kMaxDepth = 20;
constexpr float kMaxDepthGridSize = 268435456.0f;  // 2^28
// AllocatorGlobal::alloc rounds size to 16 * ((size + 16) / 16)
// (kernels/simlod/utils.h.cu:188-199): Chunk is 16,016 B, OccupancyGrid 262,144 B.
constexpr uint64_t kChunkAllocBytes = 16ull * ((16016ull + 16ull) / 16ull);    // 16,032
constexpr uint64_t kGridAllocBytes = 16ull * ((262144ull + 16ull) / 16ull);    // 262,160
constexpr uint64_t kSafetyMargin = 200000000ull;  // remolod_octree.cu:1195
constexpr uint32_t kSpillDropIndex = 3000000;     // remolod_octree.cu:872

// Saturating float -> u32 conversion, as the device's cvt does: NaN and
// negatives to 0, >= 2^32 to 0xFFFFFFFF.
inline uint32_t f2u_sat(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return uint32_t(f);
}

// remolod_octree.cu:163-169: X = u32(2^20 * (x - min) / size), same for 2^28.
inline void quantise(const remo_point& p, float size, uint32_t q20[3], uint32_t q28[3]) {
    const float fGridSize = 1048576.0f;  // pow(2.0f, float(MAX_DEPTH))
    const float c[3] = {p.x, p.y, p.z};
    for (int a = 0; a < 3; ++a) {
        const float d = c[a] - 0.0f;  // octreeMin = boxMin = 0
        volatile float num20 = fGridSize * d;
        volatile float num28 = kMaxDepthGridSize * d;
        q20[a] = f2u_sat(num20 / size);
        q28[a] = f2u_sat(num28 / size);
    }
}

struct ONode {
    int32_t children[8];
    uint32_t counter = 0, numPoints = 0, level = 0, X = 0, Y = 0, Z = 0;
    uint32_t countIteration = 0;
    int32_t grid = -1;
    uint32_t numVoxels = 0, numVoxelsStored = 0;
    uint32_t voxelChunks = 0;  // length of the voxel chunk list
    std::vector<remo_point> points;
    std::vector<remo_point> voxels;
    // Accumulate side array (include/remo/RemoAccum.h:18-30) plus the fixed point.
    remo_node_accum acc;
    ONode() {
        for (int& c : children) c = -1;
        std::memset(&acc, 0, sizeof(acc));
    }
    bool isLeaf() const {
        for (int c : children)
            if (c >= 0) return false;
        return true;
    }
};

struct Key {
    uint32_t level, X, Y, Z;
    bool operator<(const Key& o) const {
        if (level != o.level) return level < o.level;
        if (X != o.X) return X < o.X;
        if (Y != o.Y) return Y < o.Y;
        return Z < o.Z;
    }
};

}  // namespace

struct oracle_tree {
    float boxSize[3];
    float octreeSize;
    uint32_t maxNodes;
    uint64_t persistentCapacity;
    uint32_t batchesPerStep;

    std::vector<ONode> nodes;
    std::vector<std::vector<uint32_t>> grids;

    // Stats (kernels/simlod/HostDeviceInterface.h:51-75)
    uint32_t batchletIndex = 0;
    uint64_t numPointsProcessed = 0;
    int64_t numAllocatedChunks = 0;
    int64_t chunkPoolSize = 0;
    uint64_t persistentOffset = 0;
    bool memCapacityReached = false;
    bool nodePoolOverflow = false;
    bool spillOverflow = false;
    uint64_t stepBatches = 0, expandIters = 0, nodesSplit = 0, spilledPoints = 0, newVoxels = 0;
    uint64_t totBatches = 0, totExpandIters = 0, totNodesSplit = 0, totSpilledPoints = 0, totNewVoxels = 0;
    remo_stats lastStats;

    // AccumGlobals (include/remo/RemoAccum.h:33-40)
    uint64_t mortonWatermark = 0, numAccumulated = 0, sumLeafCounts = 0;
    uint32_t pointsFolded = 0, leavesFolded = 0, innerWithSums = 0;

    // per-batch scratch
    std::vector<remo_point> spilled;
    std::vector<uint32_t> spilling;
    std::vector<remo_point> backlogVoxels;
    std::vector<uint32_t> backlogTargets;

    // remolod_reset.cu:41-78
    void reset() {
        nodes.clear();
        grids.clear();
        nodes.reserve(std::min<uint32_t>(maxNodes, 1u << 20));
        persistentOffset = 16;
        batchletIndex = 0;
        numPointsProcessed = 0;
        numAllocatedChunks = 0;
        chunkPoolSize = 0;
        memCapacityReached = false;
        nodePoolOverflow = false;
        spillOverflow = false;
        mortonWatermark = numAccumulated = sumLeafCounts = 0;
        pointsFolded = leavesFolded = innerWithSums = 0;
        totBatches = totExpandIters = totNodesSplit = totSpilledPoints = totNewVoxels = 0;
        nodes.emplace_back();
        ONode& root = nodes[0];
        root.grid = allocGrid();
        std::memset(&lastStats, 0, sizeof(lastStats));
    }

    int32_t allocGrid() {
        persistentOffset += kGridAllocBytes;
        grids.emplace_back(kGridWords, 0u);
        return int32_t(grids.size() - 1);
    }

    // Root -> leaf descent (remolod_octree.cu:182-199); stops at a node without
    // the child, or after MAX_DEPTH levels.
    uint32_t descend(const uint32_t q20[3]) const {
        uint32_t cur = 0;
        for (int level = 0; level < kMaxDepth; ++level) {
            const uint32_t cx = (q20[0] >> (kMaxDepth - level - 1)) & 1u;
            const uint32_t cy = (q20[1] >> (kMaxDepth - level - 1)) & 1u;
            const uint32_t cz = (q20[2] >> (kMaxDepth - level - 1)) & 1u;
            const int32_t child = nodes[cur].children[(cx << 2) | (cy << 1) | cz];
            if (child < 0) break;
            cur = uint32_t(child);
        }
        return cur;
    }

    // doCounting (remolod_octree.cu:142-301). Serial: the warp-aggregated
    // atomicAdd becomes one increment per point, which crosses 50,000 exactly
    // once per leaf, as the grouped version does.
    bool doCounting(const remo_point* pts, uint32_t n, uint32_t countIteration) {
        spilling.clear();
        auto countPoint = [&](const remo_point& p) {
            uint32_t q20[3], q28[3];
            quantise(p, octreeSize, q20, q28);
            ONode& leaf = nodes[descend(q20)];
            if (leaf.countIteration < countIteration) {
                const uint32_t old = leaf.counter;
                leaf.counter = old + 1;
                if (old <= kMaxPointsPerNode && old + 1 > kMaxPointsPerNode)
                    spilling.push_back(uint32_t(&leaf - nodes.data()));
            }
        };
        for (uint32_t i = 0; i < n; ++i) countPoint(pts[i]);
        const size_t nSpilledBefore = spilled.size();
        for (size_t i = 0; i < nSpilledBefore; ++i) countPoint(spilled[i]);

        // Split candidates in canonical key order; the node pool accepts a prefix
        // (the reference accepts a prefix of its atomic order, :309-321).
        std::sort(spilling.begin(), spilling.end(), [&](uint32_t a, uint32_t b) {
            const ONode &na = nodes[a], &nb = nodes[b];
            return Key{na.level, na.X, na.Y, na.Z} < Key{nb.level, nb.X, nb.Y, nb.Z};
        });
        std::vector<uint32_t> accepted;
        uint64_t room = nodes.size();
        for (uint32_t id : spilling) {
            if (nodes[id].level >= uint32_t(kMaxDepth)) continue;
            if (room + 8 > maxNodes) {
                nodePoolOverflow = true;
                continue;
            }
            room += 8;
            accepted.push_back(id);
        }
        spilling.swap(accepted);

        // spill loop (:250-285): copy each spilling leaf's stored points
        for (uint32_t id : spilling) {
            const ONode& nd = nodes[id];
            spilled.insert(spilled.end(), nd.points.begin(), nd.points.end());
        }
        // stamp (:293-295)
        for (ONode& nd : nodes) nd.countIteration = countIteration;
        return spilling.empty();
    }

    // doSplitting (remolod_octree.cu:303-400)
    void doSplitting() {
        for (uint32_t id : spilling) {
            const uint32_t childOffset = uint32_t(nodes.size());
            for (int i = 0; i < 8; ++i) {
                ONode child;
                const ONode& parent = nodes[id];
                child.level = parent.level + 1;
                child.X = 2 * parent.X + ((i >> 2) & 1);
                child.Y = 2 * parent.Y + ((i >> 1) & 1);
                child.Z = 2 * parent.Z + ((i >> 0) & 1);
                nodes.push_back(child);
            }
            ONode& sn = nodes[id];
            for (int i = 0; i < 8; ++i) sn.children[i] = int32_t(childOffset + i);
            // chunks back to the queue (:349-360)
            const uint32_t nChunks = (sn.numPoints + kPointsPerChunk - 1) / kPointsPerChunk;
            numAllocatedChunks -= nChunks;
            sn.numPoints = 0;
            sn.points.clear();
            sn.points.shrink_to_fit();
            if (sn.grid < 0) sn.grid = allocGrid();
            ++nodesSplit;
            ++totNodesSplit;
        }
        // zero every spilling node's grid (:372-387), including a pre-existing one
        for (uint32_t id : spilling) {
            std::fill(grids[nodes[id].grid].begin(), grids[nodes[id].grid].end(), 0u);
        }
    }

    // expand (remolod_octree.cu:402-462)
    void expand(const remo_point* pts, uint32_t n, uint32_t batchIndex) {
        for (int i = 0; i < 20; ++i) {
            const bool finished = doCounting(pts, n, batchIndex + 1);
            ++expandIters;
            ++totExpandIters;
            if (finished) break;
            doSplitting();
        }
        spilledPoints += spilled.size();
        totSpilledPoints += spilled.size();
    }

    // sampleVoxel (remolod_octree.cu:95-140)
    void sampleVoxel(uint32_t id, const uint32_t q28[3], const remo_point& p) {
        ONode& nd = nodes[id];
        if (nd.grid < 0) return;
        const uint32_t div = 1u << ((kMaxDepth + 1) - nd.level);
        const uint32_t cx = (q28[0] / div) % kGridSize;
        const uint32_t cy = (q28[1] / div) % kGridSize;
        const uint32_t cz = (q28[2] / div) % kGridSize;
        const uint32_t vi = cx + cy * kGridSize + cz * kGridSize * kGridSize;
        uint32_t& word = grids[nd.grid][vi / 32];
        const uint32_t mask = 1u << (vi % 32);
        if (word & mask) return;
        word |= mask;
        nd.numVoxels += 1;
        // nodeSize = octreeSize / pow(2, level) is exact for integer levels
        volatile float nodeSize = octreeSize / std::ldexp(1.0f, int(nd.level));
        volatile float minx = (float(nd.X) + 0.0f) * nodeSize + 0.0f;
        volatile float miny = (float(nd.Y) + 0.0f) * nodeSize + 0.0f;
        volatile float minz = (float(nd.Z) + 0.0f) * nodeSize + 0.0f;
        remo_point v;
        volatile float tx = nodeSize * (float(cx) + 0.5f);
        volatile float ty = nodeSize * (float(cy) + 0.5f);
        volatile float tz = nodeSize * (float(cz) + 0.5f);
        v.x = minx + tx / float(kGridSize);
        v.y = miny + ty / float(kGridSize);
        v.z = minz + tz / float(kGridSize);
        v.color = p.color;
        backlogVoxels.push_back(v);
        backlogTargets.push_back(id);
        ++newVoxels;
        ++totNewVoxels;
    }

    // voxelSampling (remolod_octree.cu:677-740)
    void voxelSampling(const remo_point* pts, uint32_t n) {
        auto traverse = [&](const remo_point& p) {
            uint32_t q20[3], q28[3];
            quantise(p, octreeSize, q20, q28);
            uint32_t cur = 0;
            for (int level = 0; level < kMaxDepth; ++level) {
                const uint32_t cx = (q20[0] >> (kMaxDepth - level - 1)) & 1u;
                const uint32_t cy = (q20[1] >> (kMaxDepth - level - 1)) & 1u;
                const uint32_t cz = (q20[2] >> (kMaxDepth - level - 1)) & 1u;
                sampleVoxel(cur, q28, p);
                const int32_t child = nodes[cur].children[(cx << 2) | (cy << 1) | cz];
                if (child < 0) break;
                cur = uint32_t(child);
            }
        };
        for (uint32_t i = 0; i < n; ++i) traverse(pts[i]);
        for (size_t i = 0; i < spilled.size(); ++i) traverse(spilled[i]);
    }

    // allocatePointChunks (remolod_octree.cu:742-792)
    void allocatePointChunks() {
        for (ONode& nd : nodes) {
            if (!nd.isLeaf() || !(nd.numPoints < nd.counter)) continue;
            const int req = int((nd.counter + kPointsPerChunk - 1) / kPointsPerChunk);
            const int have = int((nd.numPoints + kPointsPerChunk - 1) / kPointsPerChunk);
            for (int i = 0; i < req - have; ++i) {
                const int64_t chunkIndex = numAllocatedChunks++;
                if (chunkIndex >= chunkPoolSize) persistentOffset += kChunkAllocBytes;
            }
        }
        chunkPoolSize = std::max(chunkPoolSize, numAllocatedChunks);
    }

    // allocateVoxelChunks (remolod_octree.cu:885-916)
    void allocateVoxelChunks() {
        for (ONode& nd : nodes) {
            const uint32_t req = (nd.numVoxels + kPointsPerChunk - 1) / kPointsPerChunk;
            if (req == 0) continue;
            while (nd.voxelChunks < req) {
                persistentOffset += kChunkAllocBytes;
                ++nd.voxelChunks;
            }
        }
    }

    // insertPoints (remolod_octree.cu:794-883)
    void insertPoints(const remo_point* pts, uint32_t n) {
        auto insert = [&](const remo_point& p) {
            uint32_t q20[3], q28[3];
            quantise(p, octreeSize, q20, q28);
            ONode& leaf = nodes[descend(q20)];
            leaf.points.push_back(p);
            leaf.numPoints += 1;
        };
        for (uint32_t i = 0; i < n; ++i) insert(pts[i]);
        if (spilled.size() > size_t(kSpillDropIndex) + 1) spillOverflow = true;
        for (size_t i = 0; i < spilled.size(); ++i) insert(spilled[i]);
    }

    // insertVoxels (remolod_octree.cu:918-938)
    void insertVoxels() {
        for (size_t i = 0; i < backlogVoxels.size(); ++i) {
            ONode& t = nodes[backlogTargets[i]];
            t.voxels.push_back(backlogVoxels[i]);
            t.numVoxelsStored += 1;
        }
    }

    // addBatch (remolod_octree.cu:940-1060)
    void addBatch(const remo_point* pts, uint32_t n, uint32_t batchIndex) {
        spilled.clear();
        backlogVoxels.clear();
        backlogTargets.clear();
        expand(pts, n, batchIndex);
        voxelSampling(pts, n);
        allocatePointChunks();
        allocateVoxelChunks();
        insertPoints(pts, n);
        insertVoxels();
    }

    // kernel_construct batch loop (remolod_octree.cu:1166-1241) and stats sweep
    // (:1248-1319). The 10 ms device-time budget is not modelled: a step folds
    // min(uploaded - batchletIndex, batchesPerStep) batches.
    uint32_t buildStep(const remo_point* slots, const uint32_t* batchSizes, uint32_t numSlots,
                       uint32_t slotCapacity, uint32_t numUploaded) {
        stepBatches = expandIters = nodesSplit = spilledPoints = newVoxels = 0;
        const uint32_t numBatches = std::min(numUploaded - batchletIndex, batchesPerStep);
        const uint32_t first = batchletIndex, last = first + numBatches;
        for (uint32_t b = first; b < last; ++b) {
            const uint32_t slot = b % numSlots;
            const uint32_t size = batchSizes[slot];
            const bool reached = persistentOffset + kSafetyMargin >= persistentCapacity;
            memCapacityReached = reached;
            if (reached) break;
            addBatch(slots + uint64_t(slot) * slotCapacity, size, batchletIndex);
            batchletIndex += 1;
            numPointsProcessed += size;
            ++stepBatches;
            ++totBatches;
        }
        sweepStats();
        return batchletIndex;
    }

    void sweepStats() {
        remo_stats& s = lastStats;
        std::memset(&s, 0, sizeof(s));
        s.num_nodes = uint32_t(nodes.size());
        for (const ONode& nd : nodes) {
            s.max_depth = std::max(s.max_depth, nd.level);
            if (nd.isLeaf()) {
                s.num_leaves++;
                s.num_points += nd.numPoints;
                s.num_chunks_points += (nd.numPoints + kPointsPerChunk - 1) / kPointsPerChunk;
                s.max_points_per_node = std::max(s.max_points_per_node, nd.numPoints);
                if (nd.counter < nd.numPoints) s.counter_underflows++;
                if (nd.numPoints > 0) s.num_nonempty_leaves++;
            } else {
                s.num_inner++;
                s.num_voxels += nd.numVoxels;
                s.num_chunks_voxels += (nd.numVoxels + kPointsPerChunk - 1) / kPointsPerChunk;
            }
        }
        s.num_points_processed = numPointsProcessed;
        s.batchlet_index = batchletIndex;
        s.num_allocated_chunks = uint64_t(numAllocatedChunks);
        s.chunk_pool_size = uint64_t(chunkPoolSize);
        s.allocated_bytes_persistent = persistentOffset;
        s.persistent_capacity = persistentCapacity;
        s.mem_capacity_reached = memCapacityReached;
        s.node_capacity_reached = s.num_nodes >= maxNodes;
        s.node_pool_overflow = nodePoolOverflow;
        s.spill_overflow = spillOverflow;
        s.batches = stepBatches;
        s.expand_iters = expandIters;
        s.nodes_split = nodesSplit;
        s.spilled_points = spilledPoints;
        s.new_voxels = newVoxels;
        s.total_batches = totBatches;
        s.total_expand_iters = totExpandIters;
        s.total_nodes_split = totNodesSplit;
        s.total_spilled_points = totSpilledPoints;
        s.total_new_voxels = totNewVoxels;
        s.morton_watermark = mortonWatermark;
        s.num_accumulated = numAccumulated;
        s.sum_leaf_counts = sumLeafCounts;
        s.points_folded = pointsFolded;
        s.leaves_folded = leavesFolded;
        s.inner_with_sums = innerWithSums;
    }

    // remoMortonSpread / remoMortonCode (kernels/remolod/remolod_accum.cu:24-36)
    static uint64_t mortonSpread(uint32_t v) {
        uint64_t x = uint64_t(v) & 0xFFFFFull;
        x = (x | (x << 32)) & 0x001F00000000FFFFull;
        x = (x | (x << 16)) & 0x001F0000FF0000FFull;
        x = (x | (x << 8)) & 0x100F00F00F00F00Full;
        x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
        x = (x | (x << 2)) & 0x1249249249249249ull;
        return x;
    }

    // kernel_accumulate (kernels/remolod/remolod_accum.cu:47-212) + the build's
    // node-centred int64 fixed point, and kernel_accum_verify (:214-233).
    void accumulate(uint32_t batchIndex) {
        pointsFolded = leavesFolded = 0;
        sumLeafCounts = 0;
        innerWithSums = 0;
        const double invSize = 1.0 / double(octreeSize);
        const double fGrid = double(1u << kMaxDepth);
        for (ONode& nd : nodes) {
            remo_node_accum& a = nd.acc;
            if (!nd.isLeaf()) {
                if (a.count != 0) std::memset(&a, 0, sizeof(a));
                continue;
            }
            const uint32_t from = a.count, to = nd.numPoints;
            if (to <= from) continue;
            const double levelScale = double(1ull << nd.level);
            uint64_t mmax = 0;
            for (uint32_t i = from; i < to; ++i) {
                const remo_point& p = nd.points[i];
                volatile double lx = (double(p.x) - 0.0) * invSize;
                volatile double ly = (double(p.y) - 0.0) * invSize;
                volatile double lz = (double(p.z) - 0.0) * invSize;
                const uint32_t maxCoord = (1u << kMaxDepth) - 1u;
                const uint32_t X = std::min(uint32_t(fGrid * lx), maxCoord);
                const uint32_t Y = std::min(uint32_t(fGrid * ly), maxCoord);
                const uint32_t Z = std::min(uint32_t(fGrid * lz), maxCoord);
                const uint64_t m = (mortonSpread(X) << 2) | (mortonSpread(Y) << 1) | mortonSpread(Z);
                mmax = std::max(mmax, m);
                double x = lx * levelScale, y = ly * levelScale, z = lz * levelScale;
                x = x - double(nd.X);
                y = y - double(nd.Y);
                z = z - double(nd.Z);
                x = std::fmin(std::fmax(x, 0.0), 1.0);
                y = std::fmin(std::fmax(y, 0.0), 1.0);
                z = std::fmin(std::fmax(z, 0.0), 1.0);
                a.g[0] += x; a.g[1] += y; a.g[2] += z;
                a.g[3] += x * x; a.g[4] += x * y; a.g[5] += x * z;
                a.g[6] += y * y; a.g[7] += y * z; a.g[8] += z * z;
                const float cr = float((p.color >> 0) & 0xFFu) * (1.0f / 255.0f);
                const float cg = float((p.color >> 8) & 0xFFu) * (1.0f / 255.0f);
                const float cb = float((p.color >> 16) & 0xFFu) * (1.0f / 255.0f);
                a.c[0] += cr; a.c[1] += cg; a.c[2] += cb;
                a.c[3] += cr * cr; a.c[4] += cg * cg; a.c[5] += cb * cb;
                // fixed point: q = floor(l * 2^20) - 2^19, l in [0,1]
                const int64_t qx = int64_t(std::floor(x * 1048576.0)) - 524288;
                const int64_t qy = int64_t(std::floor(y * 1048576.0)) - 524288;
                const int64_t qz = int64_t(std::floor(z * 1048576.0)) - 524288;
                a.q1[0] += qx; a.q1[1] += qy; a.q1[2] += qz;
                a.q2[0] += qx * qx; a.q2[1] += qx * qy; a.q2[2] += qx * qz;
                a.q2[3] += qy * qy; a.q2[4] += qy * qz; a.q2[5] += qz * qz;
            }
            a.count = to;
            a.last_touched_batch = batchIndex;
            a.morton_max = std::max(a.morton_max, mmax);
            mortonWatermark = std::max(mortonWatermark, mmax);
            numAccumulated += to - from;
            pointsFolded += to - from;
            leavesFolded += 1;
        }
        for (const ONode& nd : nodes) {
            if (nd.isLeaf())
                sumLeafCounts += nd.acc.count;
            else if (nd.acc.count != 0)
                innerWithSums += 1;
        }
        sweepStats();
    }
};

extern "C" {

oracle_tree* oracle_create(const float boxSize[3], uint32_t maxNodes, uint64_t persistentCapacity,
                           uint32_t batchesPerStep) {
    oracle_tree* t = new oracle_tree();
    for (int a = 0; a < 3; ++a) t->boxSize[a] = boxSize[a];
    t->octreeSize = std::max(std::max(boxSize[0], boxSize[1]), boxSize[2]);
    t->maxNodes = maxNodes ? maxNodes : REMO_DEFAULT_MAX_NODES;
    t->persistentCapacity = persistentCapacity;
    t->batchesPerStep = batchesPerStep ? batchesPerStep : 20;
    t->reset();
    return t;
}

void oracle_destroy(oracle_tree* t) { delete t; }
void oracle_reset(oracle_tree* t) { t->reset(); }

uint32_t oracle_build_step(oracle_tree* t, const remo_point* slots, const uint32_t* batchSizes,
                           uint32_t numSlots, uint32_t slotCapacity, uint32_t numUploaded) {
    return t->buildStep(slots, batchSizes, numSlots, slotCapacity, numUploaded);
}

void oracle_accumulate(oracle_tree* t, uint32_t batchIndex) { t->accumulate(batchIndex); }

void oracle_stats(const oracle_tree* t, remo_stats* out) { *out = t->lastStats; }

uint32_t oracle_num_nodes(const oracle_tree* t) { return uint32_t(t->nodes.size()); }

void oracle_dump_nodes(const oracle_tree* t, remo_node_record* out) {
    for (size_t i = 0; i < t->nodes.size(); ++i) {
        const ONode& nd = t->nodes[i];
        remo_node_record& r = out[i];
        std::memset(&r, 0, sizeof(r));
        r.level = nd.level;
        r.x = nd.X;
        r.y = nd.Y;
        r.z = nd.Z;
        r.first_child = nd.isLeaf() ? -1 : nd.children[0];
        r.counter = nd.counter;
        r.num_points = nd.numPoints;
        r.num_voxels = nd.numVoxels;
        r.num_voxels_stored = nd.numVoxelsStored;
        r.grid = nd.grid;
    }
}

uint32_t oracle_node_samples(const oracle_tree* t, uint32_t node, int voxels, remo_point* out,
                             uint32_t cap) {
    const std::vector<remo_point>& v = voxels ? t->nodes[node].voxels : t->nodes[node].points;
    const uint32_t n = uint32_t(std::min<size_t>(cap, v.size()));
    if (out && n) std::memcpy(out, v.data(), size_t(n) * sizeof(remo_point));
    return uint32_t(v.size());
}

int oracle_grid(const oracle_tree* t, uint32_t node, uint32_t* words) {
    const int32_t g = t->nodes[node].grid;
    if (g < 0) return -1;
    std::memcpy(words, t->grids[g].data(), kGridWords * 4);
    return 0;
}

void oracle_dump_accums(const oracle_tree* t, remo_node_accum* out) {
    for (size_t i = 0; i < t->nodes.size(); ++i) out[i] = t->nodes[i].acc;
}

// Exposed so tests can pin the quantisation against exact rational arithmetic.
void oracle_quantise(const remo_point* p, float size, uint32_t q20[3], uint32_t q28[3]) {
    quantise(*p, size, q20, q28);
}

uint64_t oracle_morton(uint32_t x, uint32_t y, uint32_t z) {
    return (oracle_tree::mortonSpread(x) << 2) | (oracle_tree::mortonSpread(y) << 1) |
           oracle_tree::mortonSpread(z);
}

}  // extern "C"

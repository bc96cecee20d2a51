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
//   * persistent-byte accounting           -- the closed-form chunk/grid count;
//   * refinement (colour filtering)        -- a numpy box filter of each node's
//     region by cell (tests/test_oracle_refine.py).
// Divergences from the reference, applied identically in the HIP path:
//   * a split refused by the node pool does not spill its points (the reference
//     spills them and re-inserts duplicates, remolod_octree.cu:250-321);
//   * spilled points past 3,000,000 are flagged, never dropped
//     (remolod_octree.cu:872-875);
//   * a node at level MAX_DEPTH never splits (the reference's descent cannot
//     reach below it, remolod_octree.cu:182).

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <thread>
#include <unordered_map>
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
// remolod_octree.cu:1191-1195: 200 MB, or 1 GB for the analytic-bucketing arm
constexpr uint64_t kSafetyMargin = 200000000ull;
constexpr uint64_t kSafetyMarginFixed = 1000000000ull;
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
    bool dead = false;         // a pooled slot of a block freed by a collapse: not in the tree
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
    uint32_t fixedDepth = 0;  // 0: iterative expand; else the analytic-bucketing arm
    // Threads for the per-point passes (descents, voxel sampling) and the
    // per-leaf folds; 1 runs everything on the calling thread. Every result is
    // the serial one: the passes keep point order wherever it is observable.
    uint32_t threads = 1;

    std::vector<ONode> nodes;
    std::vector<std::vector<uint32_t>> grids;
    std::vector<uint32_t> gridChunks;  // voxel chunks held by each grid (the chunk list stays with the grid)
    // Refinement's collapse returns what it frees to pools the splits draw from
    // first: 8-node child blocks (by first index) and grids.
    std::vector<uint32_t> freeBlocks;
    std::vector<int32_t> freeGrids;
    uint32_t maxLevel = 0;             // deepest level a split has created since reset
    uint64_t totCollapsed = 0;

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

    // per-batch scratch (kept from batch to batch, so its memory is faulted in once)
    std::vector<uint32_t> leafOf;  // leaf of each point of the batch + spill list
    struct Visit {
        uint32_t point, node, cell;  // a point's visit of a gridded node: the grid cell
    };
    std::vector<std::vector<Visit>> filed;             // threaded voxel sampling: [part][owner]
    std::vector<std::vector<uint32_t>> filedPoints;    // threaded descents: [part][owner] point indices
    std::vector<std::vector<remo_point>> freshVoxels;  // [owner] its new voxels, in point order
    std::vector<std::vector<uint32_t>> freshNodes;
    std::vector<remo_point> spilled;
    std::vector<uint32_t> spilling;
    std::vector<remo_point> backlogVoxels;
    std::vector<uint32_t> backlogTargets;

    // remolod_reset.cu:41-78
    void reset() {
        nodes.clear();
        grids.clear();
        gridChunks.clear();
        freeBlocks.clear();
        freeGrids.clear();
        maxLevel = 0;
        totCollapsed = 0;
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

    // A pooled grid first (no store bytes: its memory is already counted), then
    // a new one from the store (remolod_octree.cu:365-370).
    int32_t allocGrid() {
        if (!freeGrids.empty()) {
            const int32_t g = freeGrids.back();
            freeGrids.pop_back();
            return g;
        }
        persistentOffset += kGridAllocBytes;
        grids.emplace_back(kGridWords, 0u);
        gridChunks.push_back(0);
        return int32_t(grids.size() - 1);
    }

    // The node pool's acceptance of split candidates, given in canonical key
    // order (remolod_octree.cu:309-321, where the pool accepts a prefix of the
    // atomic order): pooled 8-node blocks first, then the unused pool while a
    // block of 8 fits below maxNodes; the rest are refused and flagged.
    void acceptByNodePool(std::vector<uint32_t>& cands) {
        std::vector<uint32_t> accepted;
        size_t pooled = freeBlocks.size();
        uint64_t room = nodes.size();
        for (uint32_t id : cands) {
            if (pooled > 0) {
                --pooled;
                accepted.push_back(id);
                continue;
            }
            if (room + 8 > maxNodes) {
                nodePoolOverflow = true;
                continue;
            }
            room += 8;
            accepted.push_back(id);
        }
        cands.swap(accepted);
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

    // Parts a pass over `n` items is cut into (one per thread; small passes stay
    // on the calling thread).
    uint32_t partsFor(size_t n) const { return n < 16384 ? 1u : std::max(threads, 1u); }

    // f(part) for part in [0, parts), a thread each.
    template <class F>
    static void eachPart(uint32_t parts, F&& f) {
        if (parts <= 1) {
            f(0u);
            return;
        }
        std::vector<std::thread> pool;
        pool.reserve(parts);
        for (uint32_t k = 0; k < parts; ++k) pool.emplace_back([&f, k] { f(k); });
        for (std::thread& th : pool) th.join();
    }

    // The part that owns node id in a threaded pass (a multiplicative hash: ids
    // modulo the part count would follow the octant, children being allocated
    // in blocks of 8, and a scene's points crowd into few octants).
    static uint32_t ownerOf(uint32_t id, uint32_t parts) { return (uint32_t(id * 0x9E3779B1u) >> 8) % parts; }

    // Item i of a batch pass: the batch's points, then the spill list.
    const remo_point& batchPoint(const remo_point* pts, uint32_t n, size_t i) const {
        return i < n ? pts[i] : spilled[i - n];
    }

    // leafOf[i] = the leaf item i descends to, for the first `total` items.
    // Threaded, each part also files its items under the part that owns their
    // leaf, so that an owner can visit its leaves' items in order.
    void descendAll(const remo_point* pts, uint32_t n, size_t total) {
        leafOf.resize(total);
        const uint32_t P = partsFor(total);
        if (P > 1) filedPoints.resize(size_t(P) * P);
        eachPart(P, [&](uint32_t part) {
            if (P > 1)
                for (uint32_t owner = 0; owner < P; ++owner) filedPoints[size_t(part) * P + owner].clear();
            for (size_t i = total * part / P, e = total * (part + 1) / P; i < e; ++i) {
                uint32_t q20[3], q28[3];
                quantise(batchPoint(pts, n, i), octreeSize, q20, q28);
                leafOf[i] = descend(q20);
                if (P > 1) filedPoints[size_t(part) * P + ownerOf(leafOf[i], P)].push_back(uint32_t(i));
            }
        });
    }

    // f(i) for every item the descent filed under `owner`, in item order.
    template <class F>
    void ownedItems(uint32_t P, uint32_t owner, F&& f) const {
        if (P == 1) {
            for (size_t i = 0; i < leafOf.size(); ++i) f(i);
            return;
        }
        for (uint32_t part = 0; part < P; ++part)
            for (uint32_t i : filedPoints[size_t(part) * P + owner]) f(size_t(i));
    }

    // doCounting (remolod_octree.cu:142-301). Serial: the warp-aggregated
    // atomicAdd becomes one increment per point, which crosses 50,000 exactly
    // once per leaf, as the grouped version does.
    bool doCounting(const remo_point* pts, uint32_t n, uint32_t countIteration) {
        spilling.clear();
        // the descents in parallel; then each part increments the counters of
        // the leaves it owns in point order
        const size_t total = size_t(n) + spilled.size();
        descendAll(pts, n, total);
        const uint32_t P = partsFor(total);
        std::vector<std::vector<uint32_t>> crossed(P);
        eachPart(P, [&](uint32_t part) {
            ownedItems(P, part, [&](size_t i) {
                const uint32_t id = leafOf[i];
                ONode& leaf = nodes[id];
                if (leaf.countIteration < countIteration) {
                    const uint32_t old = leaf.counter;
                    leaf.counter = old + 1;
                    if (old <= kMaxPointsPerNode && old + 1 > kMaxPointsPerNode) crossed[part].push_back(id);
                }
            });
        });
        // (the list is put in key order below)
        for (const std::vector<uint32_t>& c : crossed) spilling.insert(spilling.end(), c.begin(), c.end());

        // Split candidates in canonical key order; the node pool accepts a prefix
        // (the reference accepts a prefix of its atomic order, :309-321).
        std::sort(spilling.begin(), spilling.end(), [&](uint32_t a, uint32_t b) {
            const ONode &na = nodes[a], &nb = nodes[b];
            return Key{na.level, na.X, na.Y, na.Z} < Key{nb.level, nb.X, nb.Y, nb.Z};
        });
        spilling.erase(std::remove_if(spilling.begin(), spilling.end(),
                                      [&](uint32_t id) { return nodes[id].level >= uint32_t(kMaxDepth); }),
                       spilling.end());
        acceptByNodePool(spilling);

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
            // a pooled block first (the last one pooled), else the end of the pool
            uint32_t childOffset = uint32_t(nodes.size());
            const bool pooled = !freeBlocks.empty();
            if (pooled) {
                childOffset = freeBlocks.back();
                freeBlocks.pop_back();
            }
            for (int i = 0; i < 8; ++i) {
                ONode child;
                const ONode& parent = nodes[id];
                child.level = parent.level + 1;
                child.X = 2 * parent.X + ((i >> 2) & 1);
                child.Y = 2 * parent.Y + ((i >> 1) & 1);
                child.Z = 2 * parent.Z + ((i >> 0) & 1);
                if (pooled)
                    nodes[childOffset + i] = child;
                else
                    nodes.push_back(child);
            }
            maxLevel = std::max(maxLevel, nodes[id].level + 1);
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

    // The voxel position of cell (cx, cy, cz) of node nd (remolod_octree.cu:118-131).
    remo_point cellCentre(const ONode& nd, uint32_t cx, uint32_t cy, uint32_t cz) const {
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
        v.color = 0;
        return v;
    }

    // fixedMorton (remolod_octree.cu:469-481)
    static uint32_t fixedMorton(uint32_t x, uint32_t y, uint32_t z, uint32_t level) {
        uint32_t code = 0;
        for (uint32_t l = 0; l < level; ++l) {
            const uint32_t bit = level - 1 - l;
            code = (code << 3) | (((x >> bit) & 1u) << 2) | (((y >> bit) & 1u) << 1) | ((z >> bit) & 1u);
        }
        return code;
    }

    // expandFixed (remolod_octree.cu:464-675): bucket the batch at the fixed
    // depth D, sum the pyramid, split level by level every leaf above D whose
    // cell holds points, seed the leaves' counters. One expand iteration, no
    // spill. As in the HIP path, a leaf above D that holds points (its split was
    // refused earlier) is not split: the reference's arm would drop its points.
    void expandFixed(const remo_point* pts, uint32_t n) {
        const uint32_t D = fixedDepth;
        std::vector<std::vector<uint32_t>> cells(D + 1);
        for (uint32_t L = 0; L <= D; ++L) cells[L].assign(size_t(1) << (3 * L), 0u);
        const uint32_t shift = uint32_t(kMaxDepth) - D, maxCoord = (1u << D) - 1u;
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t q20[3], q28[3];
            quantise(pts[i], octreeSize, q20, q28);
            cells[D][fixedMorton((q20[0] >> shift) & maxCoord, (q20[1] >> shift) & maxCoord,
                                 (q20[2] >> shift) & maxCoord, D)] += 1;
        }
        for (int L = int(D) - 1; L >= 0; --L)
            for (size_t c = 0; c < cells[L].size(); ++c) {
                uint32_t sum = 0;
                for (int k = 0; k < 8; ++k) sum += cells[L + 1][8 * c + k];
                cells[L][c] = sum;
            }
        for (uint32_t L = 0; L < D; ++L) {
            spilling.clear();
            for (uint32_t id = 0; id < nodes.size(); ++id) {
                const ONode& nd = nodes[id];
                if (nd.dead || !nd.isLeaf() || nd.level != L || nd.numPoints != 0) continue;
                if (cells[L][fixedMorton(nd.X, nd.Y, nd.Z, L)] == 0) continue;
                spilling.push_back(id);
            }
            std::sort(spilling.begin(), spilling.end(), [&](uint32_t a, uint32_t b) {
                const ONode &na = nodes[a], &nb = nodes[b];
                return Key{na.level, na.X, na.Y, na.Z} < Key{nb.level, nb.X, nb.Y, nb.Z};
            });
            acceptByNodePool(spilling);
            if (!spilling.empty()) doSplitting();
        }
        spilling.clear();
        for (ONode& nd : nodes) {
            if (nd.dead || !nd.isLeaf() || nd.level > D) continue;
            nd.counter += cells[nd.level][fixedMorton(nd.X, nd.Y, nd.Z, nd.level)];
        }
        ++expandIters;
        ++totExpandIters;
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
        remo_point v = cellCentre(nd, cx, cy, cz);
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
        const size_t total = size_t(n) + spilled.size();
        const uint32_t P = partsFor(total);
        if (P == 1) {
            for (uint32_t i = 0; i < n; ++i) traverse(pts[i]);
            for (size_t i = 0; i < spilled.size(); ++i) traverse(spilled[i]);
            return;
        }
        // Threaded: each part descends a contiguous range of the points and files
        // every (point, node, cell) it visits under the node's owner part; each owner then replays its visits part by part, so a
        // node's cells are tested in point order and the first point to reach a
        // cell still gives the voxel its colour; the new voxels of a node reach
        // the backlog in point order, which is all the insertion reads.
        filed.resize(size_t(P) * P);
        freshVoxels.resize(P);
        freshNodes.resize(P);
        eachPart(P, [&](uint32_t part) {
            for (uint32_t owner = 0; owner < P; ++owner) filed[size_t(part) * P + owner].clear();
            for (size_t i = total * part / P, e = total * (part + 1) / P; i < e; ++i) {
                uint32_t q20[3], q28[3];
                quantise(batchPoint(pts, n, i), octreeSize, q20, q28);
                uint32_t cur = 0;
                for (int level = 0; level < kMaxDepth; ++level) {
                    const ONode& nd = nodes[cur];
                    if (nd.grid >= 0) {
                        const uint32_t div = 1u << ((kMaxDepth + 1) - nd.level);
                        const uint32_t cx = (q28[0] / div) % kGridSize;
                        const uint32_t cy = (q28[1] / div) % kGridSize;
                        const uint32_t cz = (q28[2] / div) % kGridSize;
                        filed[size_t(part) * P + ownerOf(cur, P)].push_back(
                            Visit{uint32_t(i), cur, cx + cy * kGridSize + cz * kGridSize * kGridSize});
                    }
                    const uint32_t cx = (q20[0] >> (kMaxDepth - level - 1)) & 1u;
                    const uint32_t cy = (q20[1] >> (kMaxDepth - level - 1)) & 1u;
                    const uint32_t cz = (q20[2] >> (kMaxDepth - level - 1)) & 1u;
                    const int32_t child = nd.children[(cx << 2) | (cy << 1) | cz];
                    if (child < 0) break;
                    cur = uint32_t(child);
                }
            }
        });
        eachPart(P, [&](uint32_t owner) {
            freshVoxels[owner].clear();
            freshNodes[owner].clear();
            for (uint32_t part = 0; part < P; ++part) {
                for (const Visit& v : filed[size_t(part) * P + owner]) {
                    ONode& nd = nodes[v.node];
                    uint32_t& word = grids[size_t(nd.grid)][v.cell / 32];
                    const uint32_t mask = 1u << (v.cell % 32);
                    if (word & mask) continue;
                    word |= mask;
                    nd.numVoxels += 1;
                    remo_point vox = cellCentre(nd, v.cell % kGridSize, (v.cell / kGridSize) % kGridSize,
                                                v.cell / (kGridSize * kGridSize));
                    vox.color = batchPoint(pts, n, v.point).color;
                    freshVoxels[owner].push_back(vox);
                    freshNodes[owner].push_back(v.node);
                }
            }
        });
        for (uint32_t owner = 0; owner < P; ++owner) {
            backlogVoxels.insert(backlogVoxels.end(), freshVoxels[owner].begin(), freshVoxels[owner].end());
            backlogTargets.insert(backlogTargets.end(), freshNodes[owner].begin(), freshNodes[owner].end());
            newVoxels += freshNodes[owner].size();
            totNewVoxels += freshNodes[owner].size();
        }
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
            if (req == 0 || nd.grid < 0) continue;
            uint32_t& have = gridChunks[size_t(nd.grid)];
            while (have < req) {
                persistentOffset += kChunkAllocBytes;
                ++have;
            }
        }
    }

    // insertPoints (remolod_octree.cu:794-883)
    void insertPoints(const remo_point* pts, uint32_t n) {
        // the descents in parallel; then each part appends to the leaves it
        // owns, in point order
        const size_t total = size_t(n) + spilled.size();
        descendAll(pts, n, total);
        if (spilled.size() > size_t(kSpillDropIndex) + 1) spillOverflow = true;
        const uint32_t P = partsFor(total);
        eachPart(P, [&](uint32_t part) {
            ownedItems(P, part, [&](size_t i) {
                ONode& leaf = nodes[leafOf[i]];
                leaf.points.push_back(batchPoint(pts, n, i));
                leaf.numPoints += 1;
            });
        });
    }

    // insertVoxels (remolod_octree.cu:918-938)
    void insertVoxels() {
        // each part appends to the nodes it owns, in backlog order
        const uint32_t P = partsFor(backlogVoxels.size());
        eachPart(P, [&](uint32_t part) {
            for (size_t i = 0; i < backlogVoxels.size(); ++i) {
                if (ownerOf(backlogTargets[i], P) != part) continue;
                ONode& t = nodes[backlogTargets[i]];
                t.voxels.push_back(backlogVoxels[i]);
                t.numVoxelsStored += 1;
            }
        });
    }

    // Refinement's local deepening (README.md:81, wiki/02_KernelPipeline.md:232-236;
    // no reference code), as the HIP path does it: the leaves holding at least
    // minPoints stored points (all counted points stored, level below
    // MAX_DEPTH), the densest first with ties by (level, X, Y, Z), up to
    // maxSplits; the node pool accepts them in key order as in a split round;
    // each accepted leaf's stored points go through the spill list to its
    // children (counters = the points per octant), with voxel sampling, chunk
    // acquisition and insertion as for a batch of no new points.
    // The chosen prefix also stops before its points would pass pointBudget
    // (the HIP path's spill list and voxel backlog capacities).
    void deepen(uint32_t maxSplits, uint32_t minPoints, uint64_t pointBudget, uint32_t* candidatesOut,
                uint32_t* splitsOut, uint64_t* movedOut, uint64_t* voxelsOut) {
        struct Cand {
            uint32_t np, level, X, Y, Z, id;
        };
        std::vector<Cand> cands;
        for (uint32_t id = 0; id < nodes.size(); ++id) {
            const ONode& nd = nodes[id];
            if (!nd.isLeaf() || nd.level >= uint32_t(kMaxDepth)) continue;
            if (nd.numPoints < std::max(1u, minPoints) || nd.counter != nd.numPoints) continue;
            cands.push_back({nd.numPoints, nd.level, nd.X, nd.Y, nd.Z, id});
        }
        *candidatesOut = uint32_t(cands.size());
        std::sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& b) {
            if (a.np != b.np) return a.np > b.np;
            return Key{a.level, a.X, a.Y, a.Z} < Key{b.level, b.X, b.Y, b.Z};
        });
        if (cands.size() > maxSplits) cands.resize(maxSplits);
        uint64_t points = 0;
        size_t fit = 0;
        while (fit < cands.size() && points + cands[fit].np <= pointBudget) points += cands[fit++].np;
        cands.resize(fit);
        spilling.clear();
        for (const Cand& c : cands) spilling.push_back(c.id);
        std::sort(spilling.begin(), spilling.end(), [&](uint32_t a, uint32_t b) {
            const ONode &na = nodes[a], &nb = nodes[b];
            return Key{na.level, na.X, na.Y, na.Z} < Key{nb.level, nb.X, nb.Y, nb.Z};
        });
        acceptByNodePool(spilling);
        spilled.clear();
        backlogVoxels.clear();
        backlogTargets.clear();
        std::vector<std::vector<remo_point>> moved;
        for (uint32_t id : spilling) {
            moved.push_back(nodes[id].points);
            spilled.insert(spilled.end(), nodes[id].points.begin(), nodes[id].points.end());
        }
        doSplitting();
        for (size_t k = 0; k < spilling.size(); ++k) {
            const ONode& nd = nodes[spilling[k]];
            for (const remo_point& p : moved[k]) {
                uint32_t q20[3], q28[3];
                quantise(p, octreeSize, q20, q28);
                const uint32_t sh = uint32_t(kMaxDepth) - nd.level - 1;
                const int o = int((((q20[0] >> sh) & 1u) << 2) | (((q20[1] >> sh) & 1u) << 1) | ((q20[2] >> sh) & 1u));
                nodes[size_t(nd.children[o])].counter += 1;
            }
        }
        voxelSampling(nullptr, 0);
        allocatePointChunks();
        allocateVoxelChunks();
        *voxelsOut = backlogVoxels.size();
        insertPoints(nullptr, 0);
        insertVoxels();
        *splitsOut = uint32_t(spilling.size());
        *movedOut = spilled.size();
        spilling.clear();
        spilled.clear();
        sweepStats();
    }

    // Refinement's collapsing (README.md:81, wiki/02_KernelPipeline.md:238: "merge
    // a node back up", which "needs a grid pool first"; no reference code), as the
    // HIP path does it. Candidates: inner nodes below the root whose 8 children
    // are all settled leaves (counter == numPoints) holding at most
    // min(maxPoints, 50,000) points together, so the merged leaf is one Update
    // would leave unsplit. The sparsest first, ties by (level, X, Y, Z), up to
    // maxCollapses, and a prefix whose points fit pointBudget (the HIP path's
    // spill list). Each chosen node:
    //   * takes its children's stored points back (they pass through the spill
    //     list and are inserted as for a batch of no new points: the node's
    //     counter becomes their count, its chunks come from the chunk queue);
    //   * returns its children's point chunks to the chunk queue, its grid to
    //     the grid pool and the grid's voxel chunks to the chunk queue (above its
    //     high-water mark, so later point chunks are taken from them first), and
    //     drops its voxels: as a leaf it draws its points;
    //   * frees its children's 8-node block to the node-block pool; the slots
    //     are dead (in no pass) until a split takes the block.
    // Accumulate then folds the merged leaf from scratch (its sums are cleared
    // as an inner node's are), so its fixed point is relative to its own extent.
    void collapse(uint32_t maxCollapses, uint32_t maxPoints, uint64_t pointBudget, uint32_t* candidatesOut,
                  uint32_t* collapsedOut, uint64_t* movedOut) {
        struct Cand {
            uint32_t np, level, X, Y, Z, id;
        };
        const uint32_t limit = std::min(maxPoints, kMaxPointsPerNode);
        std::vector<Cand> cands;
        for (uint32_t id = 0; id < nodes.size(); ++id) {
            const ONode& nd = nodes[id];
            if (nd.dead || nd.isLeaf() || nd.level == 0) continue;
            uint64_t total = 0;
            bool ok = true;
            for (int c : nd.children) {
                const ONode& ch = nodes[size_t(c)];
                if (!ch.isLeaf() || ch.counter != ch.numPoints || ch.grid >= 0) ok = false;
                total += ch.numPoints;
            }
            if (!ok || total > limit) continue;
            cands.push_back({uint32_t(total), nd.level, nd.X, nd.Y, nd.Z, id});
        }
        *candidatesOut = uint32_t(cands.size());
        std::sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& b) {
            if (a.np != b.np) return a.np < b.np;
            return Key{a.level, a.X, a.Y, a.Z} < Key{b.level, b.X, b.Y, b.Z};
        });
        if (cands.size() > maxCollapses) cands.resize(maxCollapses);
        uint64_t points = 0;
        size_t fit = 0;
        while (fit < cands.size() && points + cands[fit].np <= pointBudget) points += cands[fit++].np;
        cands.resize(fit);
        spilled.clear();
        backlogVoxels.clear();
        backlogTargets.clear();
        for (const Cand& cd : cands) {
            ONode& nd = nodes[cd.id];
            const uint32_t first = uint32_t(nd.children[0]);
            for (int i = 0; i < 8; ++i) {
                ONode& ch = nodes[first + uint32_t(i)];
                spilled.insert(spilled.end(), ch.points.begin(), ch.points.end());
                numAllocatedChunks -= (ch.numPoints + kPointsPerChunk - 1) / kPointsPerChunk;
                const uint32_t lvl = ch.level, X = ch.X, Y = ch.Y, Z = ch.Z;
                ch = ONode();
                ch.level = lvl;
                ch.X = X;
                ch.Y = Y;
                ch.Z = Z;
                ch.dead = true;
            }
            freeBlocks.push_back(first);
            for (int& c : nd.children) c = -1;
            if (nd.grid >= 0) {
                chunkPoolSize += gridChunks[size_t(nd.grid)];
                gridChunks[size_t(nd.grid)] = 0;
                freeGrids.push_back(nd.grid);
                nd.grid = -1;
            }
            nd.voxels.clear();
            nd.voxels.shrink_to_fit();
            nd.numVoxels = nd.numVoxelsStored = 0;
            nd.points.clear();
            nd.numPoints = 0;
            nd.counter = cd.np;
            std::memset(&nd.acc, 0, sizeof(nd.acc));
            ++totCollapsed;
        }
        allocatePointChunks();
        insertPoints(nullptr, 0);
        *collapsedOut = uint32_t(cands.size());
        *movedOut = spilled.size();
        spilled.clear();
        sweepStats();
    }

    // addBatch (remolod_octree.cu:940-1060)
    void addBatch(const remo_point* pts, uint32_t n, uint32_t batchIndex) {
        spilled.clear();
        backlogVoxels.clear();
        backlogTargets.clear();
        if (fixedDepth)
            expandFixed(pts, n);
        else
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
            const uint64_t margin = fixedDepth ? kSafetyMarginFixed : kSafetyMargin;
            const bool reached = persistentOffset + margin >= persistentCapacity;
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

    void analyse(const remo_shared_uniforms& u, uint32_t closureMargin,
                 std::vector<struct oracle_analysis_out>& out);

    // Refinement, colour filtering. The reference specifies it without code
    // (README.md:81, wiki/02_KernelPipeline.md:228-242); this restates the HIP
    // path's definition directly, without its hierarchy: for every gridded node at
    // level >= minLevel, walk every stored point of every leaf and add its colour
    // into each gridded node on its root path at that node's cell of the point;
    // then rewrite the node's voxel list in cell order (x fastest, then y, then
    // z), one voxel per set grid bit at the sampling cell centre, coloured
    // (sum + n/2) / n per RGBA channel, and set numVoxels = numVoxelsStored = the
    // set-bit count. A node is left alone when numVoxelsStored != numVoxels, when
    // numVoxels is below the set-bit count, when a point's cell bit is clear, or
    // when an inner child (refined in the same pass, deeper first) was left alone
    // or has no grid. Returns the number of nodes rewritten.
    uint32_t refine(uint32_t minLevel, uint32_t* skippedOut, uint32_t* emptyOut) {
        struct Sum {
            uint64_t c[4] = {0, 0, 0, 0};
            uint64_t n = 0;
        };
        const size_t N = nodes.size();
        std::vector<int64_t> parent(N, -1);
        for (size_t i = 0; i < N; ++i)
            for (int c : nodes[i].children)
                if (c >= 0) parent[size_t(c)] = int64_t(i);
        // cell sums per gridded node, keyed by the 21-bit cell index
        std::vector<std::unordered_map<uint32_t, Sum>> sums(N);
        std::vector<char> missing(N, 0);
        auto cellOf = [&](const uint32_t q28[3], uint32_t level) {
            const uint32_t div = 1u << ((kMaxDepth + 1) - level);
            return ((q28[0] / div) % kGridSize) + ((q28[1] / div) % kGridSize) * kGridSize +
                   ((q28[2] / div) % kGridSize) * kGridSize * kGridSize;
        };
        for (size_t i = 0; i < N; ++i) {
            const ONode& leaf = nodes[i];
            if (!leaf.isLeaf()) continue;
            for (const remo_point& p : leaf.points) {
                uint32_t q20[3], q28[3];
                quantise(p, octreeSize, q20, q28);
                for (int64_t x = int64_t(i); x >= 0; x = parent[size_t(x)]) {
                    const ONode& a = nodes[size_t(x)];
                    if (a.grid < 0 || a.level < minLevel) continue;
                    const uint32_t vi = cellOf(q28, a.level);
                    if (!(grids[a.grid][vi / 32] & (1u << (vi % 32)))) {
                        missing[size_t(x)] = 1;
                        continue;
                    }
                    Sum& s = sums[size_t(x)][vi];
                    for (int ch = 0; ch < 4; ++ch) s.c[ch] += (p.color >> (8 * ch)) & 0xFFu;
                    s.n += 1;
                }
            }
        }
        uint32_t deepest = 0;
        for (const ONode& nd : nodes) deepest = std::max(deepest, nd.level);
        std::vector<char> done(N, 0), skipped(N, 0);
        uint32_t refined = 0, nSkipped = 0, empty = 0;
        for (int L = int(deepest); L >= int(minLevel); --L) {
            for (size_t i = 0; i < N; ++i) {
                ONode& nd = nodes[i];
                if (nd.grid < 0 || int(nd.level) != L) continue;
                const std::vector<uint32_t>& g = grids[nd.grid];
                uint32_t cells = 0;
                for (uint32_t w : g) cells += uint32_t(__builtin_popcount(w));
                bool ok = nd.numVoxelsStored == nd.numVoxels && nd.numVoxels >= cells && !missing[i];
                for (int c : nd.children)
                    if (c >= 0 && !nodes[size_t(c)].isLeaf() &&
                        (nodes[size_t(c)].grid < 0 || skipped[size_t(c)] || !done[size_t(c)]))
                        ok = false;
                if (!ok) {
                    skipped[i] = 1;
                    ++nSkipped;
                    continue;
                }
                std::vector<remo_point> out;
                out.reserve(cells);
                for (uint32_t vi = 0; vi < kGridWords * 32; ++vi) {
                    if (!(g[vi / 32] & (1u << (vi % 32)))) continue;
                    remo_point v = cellCentre(nd, vi % kGridSize, (vi / kGridSize) % kGridSize,
                                              vi / (kGridSize * kGridSize));
                    const auto it = sums[i].find(vi);
                    const Sum* s = it == sums[i].end() ? nullptr : &it->second;
                    if (!s || s->n == 0) {
                        ++empty;
                    } else {
                        uint32_t color = 0;
                        for (int ch = 0; ch < 4; ++ch)
                            color |= uint32_t(((s->c[ch] + s->n / 2) / s->n) & 0xFFu) << (8 * ch);
                        v.color = color;
                    }
                    out.push_back(v);
                }
                nd.voxels.swap(out);
                nd.numVoxels = nd.numVoxelsStored = cells;
                done[i] = 1;
                ++refined;
            }
        }
        sweepStats();
        if (skippedOut) *skippedOut = nSkipped;
        if (emptyOut) *emptyOut = empty;
        return refined;
    }

    void render(const remo_shared_uniforms& u, uint64_t* fb, uint32_t* numItems, uint64_t* numSamples,
                uint32_t capacity);

    void sweepStats() {
        remo_stats& s = lastStats;
        std::memset(&s, 0, sizeof(s));
        s.num_nodes = uint32_t(nodes.size() - 8 * freeBlocks.size());
        for (const ONode& nd : nodes) {
            if (nd.dead) continue;
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
        s.node_capacity_reached = nodes.size() >= maxNodes && freeBlocks.empty();
        s.free_node_blocks = uint32_t(freeBlocks.size());
        s.pooled_grids = uint32_t(freeGrids.size());
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
    // Folds leaf nd's new points into its accumulator; returns how many it folded
    // (0: nothing new, or an inner node, whose sums are cleared) and raises mmax
    // to the largest Morton code among them.
    uint32_t foldNode(ONode& nd, uint32_t batchIndex, uint64_t& mmax) const {
        const double invSize = 1.0 / double(octreeSize);
        const double fGrid = double(1u << kMaxDepth);
        remo_node_accum& a = nd.acc;
        if (!nd.isLeaf()) {
            if (a.count != 0) std::memset(&a, 0, sizeof(a));
            return 0;
        }
        const uint32_t from = a.count, to = nd.numPoints;
        if (to <= from) return 0;
        const double levelScale = double(1ull << nd.level);
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
        return to - from;
    }

    void accumulate(uint32_t batchIndex) {
        pointsFolded = leavesFolded = 0;
        sumLeafCounts = 0;
        innerWithSums = 0;
        // leaves in parallel (a shared cursor over blocks of 16 nodes: a batch's
        // points crowd into few leaves), the totals combined after
        struct Part {
            uint64_t mmax = 0, points = 0;
            uint32_t leaves = 0;
        };
        const uint32_t P = partsFor(size_t(nodes.size()) * 1024u);
        std::vector<Part> parts(P);
        std::atomic<size_t> cursor{0};
        eachPart(P, [&](uint32_t part) {
            Part& r = parts[part];
            for (size_t b = cursor.fetch_add(16); b < nodes.size(); b = cursor.fetch_add(16)) {
                for (size_t k = b, e = std::min(b + 16, nodes.size()); k < e; ++k) {
                    uint64_t mmax = 0;
                    const uint32_t folded = foldNode(nodes[k], batchIndex, mmax);
                    if (!folded) continue;
                    nodes[k].acc.morton_max = std::max(nodes[k].acc.morton_max, mmax);
                    r.mmax = std::max(r.mmax, mmax);
                    r.points += folded;
                    r.leaves += 1;
                }
            }
        });
        for (const Part& r : parts) {
            mortonWatermark = std::max(mortonWatermark, r.mmax);
            numAccumulated += r.points;
            pointsFolded += uint32_t(r.points);
            leavesFolded += r.leaves;
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


// ---------------------------------------------------------------------------
// Analysis (spec only in the reference: README.md:80, wiki/02_KernelPipeline.md:205-224,
// include/remo/RemoAccum.h:18-30). The build's definitions, restated here:
//  * roll-up: a leaf's fixed-point sums (units 2^-(20+L) of the root extent,
//    centred on the leaf) are rescaled to units U = 2^-(20+D), D the deepest
//    node level, and re-centred on each ancestor with the exact child offset
//    +-2^(18+D-L) per axis; everything in __int128;
//  * covariance in node-extent units from num = n*S2 - S1*S1^T (exact int128 when
//    it fits, FP64 otherwise), cov = num / n^2 / 4^(20+D-L);
//  * eigenvalues: cyclic Jacobi on cov / trace in FP32 (the eigenvalues of a
//    symmetric matrix are perturbed by at most the norm of the perturbation, so
//    FP32 holds them to ~1e-7 of the trace at any eigen-gap); FP64 Jacobi is
//    computed beside it as the accuracy reference;
//  * surface variation l0/(l0+l1+l2), planarity (l1-l0)/l2, linearity (l2-l1)/l2;
//  * coverage: projected 8-corner screen extent dx*dy, clamped to the screen;
//    score = 3 * surface variation * coverage;
//  * closure: a leaf is closed when the current batch is >= lastTouchedBatch +
//    margin; an inner node when all its children are; closed nodes are Finalized.

namespace anal {

using i128 = __int128;

// Left shift of a possibly negative __int128 (two's complement, no UB).
inline i128 shl(i128 x, int k) { return i128((unsigned __int128)x << k); }

inline double i128ToDouble(i128 x) {
    const bool neg = x < 0;
    unsigned __int128 u = neg ? (unsigned __int128)(-x) : (unsigned __int128)x;
    const double hi = double(uint64_t(u >> 64));
    const double lo = double(uint64_t(u));
    const double d = hi * 18446744073709551616.0 + lo;
    return neg ? -d : d;
}

inline int bitsU128(unsigned __int128 u) {
    int b = 0;
    while (u) {
        ++b;
        u >>= 1;
    }
    return b;
}
inline int bitsI128(i128 x) { return bitsU128(x < 0 ? (unsigned __int128)(-x) : (unsigned __int128)x); }

template <typename T>
void jacobi3(T a[6], T ev[3]) {
    // a: a00 a01 a02 a11 a12 a22
    T m[3][3] = {{a[0], a[1], a[2]}, {a[1], a[3], a[4]}, {a[2], a[4], a[5]}};
    for (int sweep = 0; sweep < 10; ++sweep) {
        const T off = m[0][1] * m[0][1] + m[0][2] * m[0][2] + m[1][2] * m[1][2];
        if (!(off > T(0))) break;
        for (int pq = 0; pq < 3; ++pq) {
            const int p = pq == 2 ? 1 : 0;
            const int q = pq == 0 ? 1 : 2;
            const T apq = m[p][q];
            if (apq == T(0)) continue;
            const T theta = (m[q][q] - m[p][p]) / (T(2) * apq);
            const T at = theta < T(0) ? -theta : theta;
            T t;
            if (at > T(1e15))
                t = T(0.5) / theta;
            else
                t = (theta < T(0) ? T(-1) : T(1)) / (at + std::sqrt(theta * theta + T(1)));
            const T c = T(1) / std::sqrt(t * t + T(1));
            const T sn = t * c;
            const int r = 3 - p - q;
            const T arp = m[r][p], arq = m[r][q];
            m[p][p] = m[p][p] - t * apq;
            m[q][q] = m[q][q] + t * apq;
            m[p][q] = m[q][p] = T(0);
            const T nrp = c * arp - sn * arq;
            const T nrq = sn * arp + c * arq;
            m[r][p] = m[p][r] = nrp;
            m[r][q] = m[q][r] = nrq;
        }
    }
    ev[0] = m[0][0];
    ev[1] = m[1][1];
    ev[2] = m[2][2];
    // ascending
    if (ev[0] > ev[1]) std::swap(ev[0], ev[1]);
    if (ev[1] > ev[2]) std::swap(ev[1], ev[2]);
    if (ev[0] > ev[1]) std::swap(ev[0], ev[1]);
}

struct Sums {
    i128 s1[3], s2[6];
    uint64_t n;
};

}  // namespace anal

struct oracle_analysis_out {
    remo_node_analysis r;
    double lambda64[3];  // FP64 Jacobi on the same covariance (the reference)
    double cov[6];       // covariance in node-extent units
    double mean[3];      // mean offset from the node centre, node-extent units
    int64_t s1_hi[3], s1_lo[3], s2_hi[6], s2_lo[6];  // the rolled-up int128 sums
    uint64_t n;
    uint32_t e;          // 20 + D - L: log2 of the node extent in units U
    uint32_t exact;      // 1: covariance numerator computed exactly in int128
};

// Projected extent of a node's box (kernels/remolod/remolod_render.cu:30-70).
static void nodeExtent(const ONode& nd, float cubeSize, const remo_shared_uniforms& u, float& dx, float& dy) {
    const float size = cubeSize / std::ldexp(1.0f, int(nd.level));
    const float mn[3] = {0.0f + float(nd.X) * size, 0.0f + float(nd.Y) * size, 0.0f + float(nd.Z) * size};
    const float mx[3] = {mn[0] + size, mn[1] + size, mn[2] + size};
    float minX = 1e30f, minY = 1e30f, maxX = -1e30f, maxY = -1e30f;
    bool any = false;
    dx = dy = 0.0f;
    for (int corner = 0; corner < 8; ++corner) {
        const float x = (corner & 1) ? mx[0] : mn[0];
        const float y = (corner & 2) ? mx[1] : mn[1];
        const float z = (corner & 4) ? mx[2] : mn[2];
        const auto& m = u.transform_frozen.rows;
        const float cx = m[0][0] * x + m[0][1] * y + m[0][2] * z + m[0][3] * 1.0f;
        const float cy = m[1][0] * x + m[1][1] * y + m[1][2] * z + m[1][3] * 1.0f;
        const float cw = m[3][0] * x + m[3][1] * y + m[3][2] * z + m[3][3] * 1.0f;
        if (cw <= 0.0f) continue;
        const float invW = 1.0f / cw;
        const float sx = (cx * invW * 0.5f + 0.5f) * u.width;
        const float sy = (cy * invW * 0.5f + 0.5f) * u.height;
        minX = std::fmin(minX, sx);
        maxX = std::fmax(maxX, sx);
        minY = std::fmin(minY, sy);
        maxY = std::fmax(maxY, sy);
        any = true;
    }
    if (!any) return;
    dx = maxX - minX;
    dy = maxY - minY;
}

void oracle_tree::analyse(const remo_shared_uniforms& u, uint32_t closureMargin,
                          std::vector<oracle_analysis_out>& out) {
    using namespace anal;
    const size_t N = nodes.size();
    out.assign(N, oracle_analysis_out{});
    // units 2^-(20+D) with D the deepest level any split created (a collapse does
    // not lower it), as the HIP path keeps it
    const uint32_t D = maxLevel;
    std::vector<Sums> sums(N);
    std::vector<int> closed(N, 0);
    // bottom-up: children always have larger indices than their parent here
    std::vector<uint32_t> order(N);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return nodes[a].level > nodes[b].level; });
    const uint32_t current = batchletIndex;
    for (uint32_t id : order) {
        const ONode& nd = nodes[id];
        Sums& S = sums[id];
        std::memset(&S, 0, sizeof(S));
        if (nd.isLeaf()) {
            const remo_node_accum& a = nd.acc;
            const int sh = int(D - nd.level);
            for (int i = 0; i < 3; ++i) S.s1[i] = shl(i128(a.q1[i]), sh);
            for (int i = 0; i < 6; ++i) S.s2[i] = shl(i128(a.q2[i]), 2 * sh);
            S.n = a.count;
            closed[id] = (a.count == 0) || (current >= a.last_touched_batch + closureMargin);
        } else {
            const i128 h = shl(i128(1), int(18 + D - nd.level));
            closed[id] = 1;
            for (int ci = 0; ci < 8; ++ci) {
                const uint32_t c = uint32_t(nd.children[ci]);
                const Sums& C = sums[c];
                const i128 o[3] = {(ci & 4) ? h : -h, (ci & 2) ? h : -h, (ci & 1) ? h : -h};
                const i128 n = i128(C.n);
                for (int i = 0; i < 3; ++i) S.s1[i] += C.s1[i] + n * o[i];
                const int pi[6] = {0, 0, 0, 1, 1, 2}, pj[6] = {0, 1, 2, 1, 2, 2};
                for (int k = 0; k < 6; ++k) {
                    const int a_ = pi[k], b_ = pj[k];
                    S.s2[k] += C.s2[k] + o[a_] * C.s1[b_] + C.s1[a_] * o[b_] + n * o[a_] * o[b_];
                }
                S.n += C.n;
                closed[id] &= closed[c];
            }
        }
    }
    const float cubeSize = octreeSize;
    for (size_t id = 0; id < N; ++id) {
        const ONode& nd = nodes[id];
        const Sums& S = sums[id];
        oracle_analysis_out& o = out[id];
        o.n = S.n;
        o.e = 20 + D - nd.level;
        for (int i = 0; i < 3; ++i) {
            o.s1_hi[i] = int64_t(S.s1[i] >> 64);
            o.s1_lo[i] = int64_t(uint64_t(S.s1[i]));
        }
        for (int i = 0; i < 6; ++i) {
            o.s2_hi[i] = int64_t(S.s2[i] >> 64);
            o.s2_lo[i] = int64_t(uint64_t(S.s2[i]));
        }
        remo_node_analysis& r = o.r;
        std::memset(&r, 0, sizeof(r));
        r.n = S.n;
        float dx, dy;
        nodeExtent(nd, cubeSize, u, dx, dy);
        const float cw = std::fmin(dx, u.width), ch = std::fmin(dy, u.height);
        r.coverage = cw * ch;
        r.state = closed[id] ? 2u : 0u;
        if (S.n < 3) continue;
        // covariance numerator
        const int pi[6] = {0, 0, 0, 1, 1, 2}, pj[6] = {0, 1, 2, 1, 2, 2};
        const i128 n = i128(S.n);
        int maxS2 = 0, maxS1 = 0;
        for (int k = 0; k < 6; ++k) maxS2 = std::max(maxS2, bitsI128(S.s2[k]));
        for (int k = 0; k < 3; ++k) maxS1 = std::max(maxS1, bitsI128(S.s1[k]));
        const bool exact = bitsU128((unsigned __int128)S.n) + maxS2 <= 125 && 2 * maxS1 <= 125;
        o.exact = exact;
        const double scale = std::ldexp(1.0, -int(o.e));
        const double dn = double(S.n);
        double cov[6];
        for (int k = 0; k < 6; ++k) {
            double num;
            if (exact) {
                num = i128ToDouble(n * S.s2[k] - S.s1[pi[k]] * S.s1[pj[k]]);
            } else {
                num = i128ToDouble(S.s2[k]) * dn - i128ToDouble(S.s1[pi[k]]) * i128ToDouble(S.s1[pj[k]]);
            }
            cov[k] = num / (dn * dn) * scale * scale;
            o.cov[k] = cov[k];
        }
        for (int i = 0; i < 3; ++i) o.mean[i] = i128ToDouble(S.s1[i]) / dn * scale;
        const double trace = cov[0] + cov[3] + cov[5];
        if (!(trace > 0.0)) continue;
        double a64[6];
        float a32[6];
        for (int k = 0; k < 6; ++k) {
            a64[k] = cov[k] / trace;
            a32[k] = float(a64[k]);
        }
        double ev64[3];
        jacobi3<double>(a64, ev64);
        float ev32[3];
        jacobi3<float>(a32, ev32);
        double ev[3];
        for (int i = 0; i < 3; ++i) ev[i] = double(ev32[i]);
        for (int i = 0; i < 3; ++i) {
            o.lambda64[i] = ev64[i] * trace;
            r.lambda[i] = float(ev[i] * trace);
        }
        const double sum = ev[0] + ev[1] + ev[2];
        const float sv = sum > 0.0 ? float(ev[0] / sum) : 0.0f;
        r.surface_variation = sv;
        r.planarity = ev[2] > 0.0 ? float((ev[1] - ev[0]) / ev[2]) : 0.0f;
        r.linearity = ev[2] > 0.0 ? float((ev[2] - ev[1]) / ev[2]) : 0.0f;
        r.score = 3.0f * sv * r.coverage;
    }
}


// ---------------------------------------------------------------------------
// Rasterize: kernel_render's selection and the shared rasteriser
// (kernels/remolod/remolod_render.cu:72-181, kernels/shared/remo_draw.cuh:118-167,
// kernels/shared/remo_framebuffer.cuh:214-248, kernels/shared/remo_math.cuh:16-79).
// Order-independent: a pixel keeps the minimum of (float_bits(depth) << 32) | rgba.

namespace rast {

struct Plane {
    float nx, ny, nz, d;
};

inline void frustumPlanes(const remo_mat4& m, Plane P[6]) {
    const float(*r)[4] = m.rows;
    const float raw[6][4] = {
        {r[3][0] - r[0][0], r[3][1] - r[0][1], r[3][2] - r[0][2], r[3][3] - r[0][3]},
        {r[3][0] + r[0][0], r[3][1] + r[0][1], r[3][2] + r[0][2], r[3][3] + r[0][3]},
        {r[3][0] + r[1][0], r[3][1] + r[1][1], r[3][2] + r[1][2], r[3][3] + r[1][3]},
        {r[3][0] - r[1][0], r[3][1] - r[1][1], r[3][2] - r[1][2], r[3][3] - r[1][3]},
        {r[3][0] - r[2][0], r[3][1] - r[2][1], r[3][2] - r[2][2], r[3][3] - r[2][3]},
        {r[3][0] + r[2][0], r[3][1] + r[2][1], r[3][2] + r[2][2], r[3][3] + r[2][3]},
    };
    for (int i = 0; i < 6; ++i) {
        const float a = raw[i][0], b = raw[i][1], c = raw[i][2], d = raw[i][3];
        const float len = std::sqrt(a * a + b * b + c * c);
        const float inv = len > 0.0f ? 1.0f / len : 0.0f;
        P[i] = {a * inv, b * inv, c * inv, d * inv};
    }
}

inline int32_t f2i_sat(float f) {
    if (!(f == f)) return 0;
    if (f >= 2147483647.0f) return 2147483647;
    if (f <= -2147483648.0f) return int32_t(0x80000000u);
    return int32_t(f);
}

inline void drawPoint(uint64_t* fb, const remo_shared_uniforms& u, int32_t w, int32_t h, const remo_point& p) {
    const auto& M = u.transform.rows;
    const float cx = M[0][0] * p.x + M[0][1] * p.y + M[0][2] * p.z + M[0][3] * 1.0f;
    const float cy = M[1][0] * p.x + M[1][1] * p.y + M[1][2] * p.z + M[1][3] * 1.0f;
    const float cw = M[3][0] * p.x + M[3][1] * p.y + M[3][2] * p.z + M[3][3] * 1.0f;
    if (cw <= 0.0f) return;
    const float invW = 1.0f / cw;
    const float px = (cx * invW * 0.5f + 0.5f) * u.width;
    const float py = (cy * invW * 0.5f + 0.5f) * u.height;
    const int64_t half = u.point_size / 2;
    const int64_t ix = f2i_sat(px), iy = f2i_sat(py);
    if (ix + half < 0 || iy + half < 0 || ix - half >= w || iy - half >= h) return;
    uint32_t bits;
    std::memcpy(&bits, &cw, 4);
    const uint32_t color = u.color_mode == 3 ? 0xFFFFFFFFu : p.color;  // COLOR_WHITE or COLOR_RGB
    const uint64_t packed = (uint64_t(bits) << 32) | color;
    for (int64_t oy = -half; oy <= half; ++oy) {
        const int64_t sy = iy + oy;
        if (sy < 0 || sy >= h) continue;
        for (int64_t ox = -half; ox <= half; ++ox) {
            const int64_t sx = ix + ox;
            if (sx < 0 || sx >= w) continue;
            uint64_t& px64 = fb[sy * w + sx];
            if (packed < px64) px64 = packed;
        }
    }
}

}  // namespace rast

void oracle_tree::render(const remo_shared_uniforms& u, uint64_t* fb, uint32_t* numItems, uint64_t* numSamples,
                         uint32_t capacity) {
    const int32_t w = int32_t(u.width), h = int32_t(u.height);
    for (int64_t i = 0; i < int64_t(w) * h; ++i) fb[i] = 0x7f80000000332211ull;
    const uint32_t n = uint32_t(std::min<size_t>(nodes.size(), maxNodes));
    std::vector<uint8_t> visible(n), large(n);
    rast::Plane P[6];
    rast::frustumPlanes(u.transform_frozen, P);
    for (uint32_t i = 0; i < n; ++i) {
        const ONode& nd = nodes[i];
        const float size = octreeSize / std::ldexp(1.0f, int(nd.level));
        const float mn[3] = {0.0f + float(nd.X) * size, 0.0f + float(nd.Y) * size, 0.0f + float(nd.Z) * size};
        const float mx[3] = {mn[0] + size, mn[1] + size, mn[2] + size};
        bool inside = true;
        for (int k = 0; k < 6 && inside; ++k) {
            const rast::Plane& p = P[k];
            const float px = p.nx > 0.0f ? mx[0] : mn[0];
            const float py = p.ny > 0.0f ? mx[1] : mn[1];
            const float pz = p.nz > 0.0f ? mx[2] : mn[2];
            if (p.nx * px + p.ny * py + p.nz * pz + p.d < 0.0f) inside = false;
        }
        const bool hasSamples = nd.numPoints > 0 || nd.numVoxels > 0;
        visible[i] = hasSamples && inside;
        float dx, dy;
        nodeExtent(nd, octreeSize, u, dx, dy);
        large[i] = (dx > u.lod_pixel_budget || dy > u.lod_pixel_budget) ? 1 : 0;
    }
    std::vector<uint32_t> items;
    uint64_t samples = 0;
    auto emit = [&](uint32_t id) {
        const ONode& nd = nodes[id];
        if (nd.numPoints == 0 && nd.numVoxels == 0) return;
        if (items.size() >= capacity) return;
        items.push_back(id);
        samples += uint64_t(nd.numPoints) + nd.numVoxels;
    };
    for (uint32_t i = 0; i < n; ++i) {
        if (!large[i]) continue;
        const ONode& nd = nodes[i];
        if (!nd.isLeaf()) {
            for (int c = 0; c < 8; ++c) {
                const int32_t child = nd.children[c];
                if (child < 0 || uint32_t(child) >= n) continue;
                if (large[child] || !visible[child]) continue;
                emit(uint32_t(child));
            }
        } else if (visible[i]) {
            emit(i);
        }
    }
    if (u.show_points != 0) {
        for (uint32_t id : items) {
            const ONode& nd = nodes[id];
            for (uint32_t k = 0; k < nd.numPoints && k < nd.points.size(); ++k) rast::drawPoint(fb, u, w, h, nd.points[k]);
            for (uint32_t k = 0; k < nd.numVoxels && k < nd.voxels.size(); ++k) rast::drawPoint(fb, u, w, h, nd.voxels[k]);
        }
    }
    *numItems = uint32_t(items.size());
    *numSamples = samples;
}

extern "C" {

oracle_tree* oracle_create(const float boxSize[3], uint32_t maxNodes, uint64_t persistentCapacity,
                           uint32_t batchesPerStep, uint32_t fixedDepth) {
    oracle_tree* t = new oracle_tree();
    t->fixedDepth = fixedDepth;
    for (int a = 0; a < 3; ++a) t->boxSize[a] = boxSize[a];
    t->octreeSize = std::max(std::max(boxSize[0], boxSize[1]), boxSize[2]);
    t->maxNodes = maxNodes ? maxNodes : REMO_DEFAULT_MAX_NODES;
    t->persistentCapacity = persistentCapacity;
    t->batchesPerStep = batchesPerStep ? batchesPerStep : 20;
    t->reset();
    return t;
}

void oracle_destroy(oracle_tree* t) { delete t; }
// Threads for the per-point passes and the per-leaf folds (results unchanged).
void oracle_set_threads(oracle_tree* t, uint32_t threads) { t->threads = threads ? threads : 1u; }
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
        r.free_slot = nd.dead ? 1u : 0u;
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

// Analysis over the current tree; out has oracle_num_nodes() entries.
uint32_t oracle_analysis_record_size() { return uint32_t(sizeof(oracle_analysis_out)); }

void oracle_analyse(oracle_tree* t, const remo_shared_uniforms* u, uint32_t closureMargin,
                    oracle_analysis_out* out) {
    std::vector<oracle_analysis_out> v;
    t->analyse(*u, closureMargin, v);
    std::memcpy(out, v.data(), v.size() * sizeof(oracle_analysis_out));
}

// The flat pipeline's Rasterize (kernels/flat/flat_render.cu:9-76): every point,
// no LOD, the shared splat. COLOR_RGB and COLOR_WHITE.
void oracle_render_flat(const remo_point* pts, uint64_t n, const remo_shared_uniforms* u, uint64_t* fb) {
    const int32_t w = int32_t(u->width), h = int32_t(u->height);
    for (int64_t i = 0; i < int64_t(w) * h; ++i) fb[i] = 0x7f80000000332211ull;
    if (u->show_points == 0) return;
    for (uint64_t i = 0; i < n; ++i) rast::drawPoint(fb, *u, w, h, pts[i]);
}

// remoApplyEDL (kernels/shared/remo_framebuffer.cuh:78-122), with the C library's
// log2f / expf where the reference uses the fast intrinsics.
void oracle_apply_edl(const remo_shared_uniforms* u, uint64_t* fb) {
    if (u->enable_edl == 0) return;
    const int32_t w = int32_t(u->width), h = int32_t(u->height);
    auto depthOf = [](uint64_t px) {
        const uint32_t bits = uint32_t(px >> 32);
        float d;
        std::memcpy(&d, &bits, 4);
        return d;
    };
    for (int32_t y = 0; y < h; ++y)
        for (int32_t x = 0; x < w; ++x) {
            const uint64_t pixel = fb[int64_t(y) * w + x];
            const float depth = depthOf(pixel);
            if (!(depth < 3.0e38f)) continue;
            const float logDepth = std::log2(depth);
            float response = 0.0f;
            int taps = 0;
            const int32_t off[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
            for (int k = 0; k < 4; ++k) {
                const int32_t nx = x + off[k][0], ny = y + off[k][1];
                if (nx < 0 || ny < 0 || nx >= w || ny >= h) continue;
                const float nd = depthOf(fb[int64_t(ny) * w + nx]);
                if (!(nd < 3.0e38f)) continue;
                response += std::max(0.0f, logDepth - std::log2(nd));
                ++taps;
            }
            if (taps == 0) continue;
            response /= float(taps);
            const float shade = std::exp(-response * 24.0f * u->edl_strength);
            const uint32_t c = uint32_t(pixel);
            const uint32_t r = uint32_t(float(c & 0xFFu) * shade);
            const uint32_t g = uint32_t(float((c >> 8) & 0xFFu) * shade);
            const uint32_t b = uint32_t(float((c >> 16) & 0xFFu) * shade);
            fb[int64_t(y) * w + x] =
                (pixel & 0xFFFFFFFF00000000ull) | (r & 0xFFu) | ((g & 0xFFu) << 8) | ((b & 0xFFu) << 16) | (255u << 24);
        }
}

void oracle_deepen(oracle_tree* t, uint32_t maxSplits, uint32_t minPoints, uint64_t pointBudget, uint32_t* candidates,
                   uint32_t* splits, uint64_t* moved, uint64_t* voxels) {
    t->deepen(maxSplits, minPoints, pointBudget, candidates, splits, moved, voxels);
}

void oracle_collapse(oracle_tree* t, uint32_t maxCollapses, uint32_t maxPoints, uint64_t pointBudget,
                     uint32_t* candidates, uint32_t* collapsed, uint64_t* moved) {
    t->collapse(maxCollapses, maxPoints, pointBudget, candidates, collapsed, moved);
}

uint32_t oracle_refine(oracle_tree* t, uint32_t minLevel, uint32_t* skipped, uint32_t* empty) {
    return t->refine(minLevel, skipped, empty);
}

void oracle_render(oracle_tree* t, const remo_shared_uniforms* u, uint64_t* fb, uint32_t* numItems,
                   uint64_t* numSamples) {
    t->render(*u, fb, numItems, numSamples, 131072u);
}

// The FP32 and FP64 Jacobi solvers on one symmetric matrix (a00 a01 a02 a11 a12 a22).
void oracle_jacobi(const double a[6], double ev64[3], float ev32[3]) {
    double a64[6];
    float a32[6];
    for (int k = 0; k < 6; ++k) {
        a64[k] = a[k];
        a32[k] = float(a[k]);
    }
    anal::jacobi3<double>(a64, ev64);
    anal::jacobi3<float>(a32, ev32);
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

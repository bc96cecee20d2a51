// Per-frame passes over a built tree: Analysis (roll-up, 3x3 eigen, score) and
// Rasterize (frontier selection + 64-bit depth|colour splat).
#pragma once

#include <functional>
#include <string>

#include "remolod_internal.h"

namespace remo_hip {

// Exact roll-up record of one node: sums of point coordinates (integers in
// units 2^-(20 + D) of the root extent, D the deepest level) and of their
// products, over the points under the node; e = 20 + D - level, the node extent
// being 2^e units. Stored in the root frame (origin at the root's minimum
// corner, sums modulo 2^128), so a parent's record is the sum of its children's;
// read in the node's own frame (centred on the node), where the sums are exact
// __int128 values.
struct RollupSums {
    __int128 s1[3];
    __int128 s2[6];
    unsigned long long n;
    uint32_t e;
    uint32_t pad;
};

// Left shift of a possibly negative __int128 (two's complement, no UB).
__device__ __forceinline__ __int128 shl(__int128 x, int k) { return __int128((unsigned __int128)x << k); }

// a * b modulo 2^128 for a small unsigned factor (a 64x32 and a 32x32 product of
// limbs instead of the general 128x128 expansion).
__device__ __forceinline__ __int128 mulU32(__int128 a, uint32_t b) {
    const unsigned __int128 u = (unsigned __int128)a;
    const uint64_t lo = uint64_t(u), hi = uint64_t(u >> 64);
    const unsigned __int128 plo = (unsigned __int128)lo * b;
    return __int128(plo + ((unsigned __int128)(hi * b) << 64));
}

// A leaf's record from its accumulator. The accumulator's sums are in units
// 2^-(20+L) centred on the leaf (L its level); in the root frame (units
// 2^-(20+D), D the deepest level) a point is x = c + 2^(D-L) u with the leaf's
// centre c = odd * 2^(19+D-L), so S1 = 2^(D-L) Q1 + n c and
// S2_ab = 4^(D-L) Q2_ab + c_a 2^(D-L) Q1_b + 2^(D-L) Q1_a c_b + n c_a c_b.
struct LeafFrame {
    int sh, s;         // D - L and 19 + D - L
    uint32_t odd[3];   // 2 x + 1 per axis, x the leaf's integer coordinate
};

__device__ __forceinline__ LeafFrame leafFrame(uint32_t D, uint32_t level, uint32_t x, uint32_t y, uint32_t z) {
    return LeafFrame{int(D - level), int(19 + D - level), {2u * x + 1u, 2u * y + 1u, 2u * z + 1u}};
}

// Value f of the record: S1[f] for f < 3, S2[f - 3] (xx xy xz yy yz zz) after.
__device__ __forceinline__ __int128 leafRecordValue(const LeafFrame& F, const int64_t q1[3], int64_t q2, uint64_t n,
                                                    int f) {
    const __int128 nn = __int128(n);
    if (f < 3) return shl(__int128(q1[f]), F.sh) + shl(mulU32(nn, F.odd[f]), F.s);
    const int q = f - 3;
    const int a = q < 3 ? 0 : (q < 5 ? 1 : 2);
    const int b = q < 3 ? q : (q < 5 ? q - 2 : 2);
    const __int128 qa = shl(__int128(q1[a]), F.sh), qb = shl(__int128(q1[b]), F.sh);
    const __int128 cross = mulU32(qb, F.odd[a]) + mulU32(qa, F.odd[b]);
    return shl(__int128(q2), 2 * F.sh) + shl(cross, F.s) + shl(mulU32(mulU32(nn, F.odd[a]), F.odd[b]), 2 * F.s);
}

// Marks a node's record as changed since its shape was solved.
__device__ __forceinline__ void markRecordChanged(const DevTree& t, uint32_t node) { t.shapeStale[node] = 1u; }

// Nearest double to an __int128 (two halves, one rounding each).
__host__ __device__ inline double i128ToDouble(__int128 x) {
    const bool neg = x < 0;
    const unsigned __int128 u = neg ? (unsigned __int128)(-x) : (unsigned __int128)x;
    const double hi = double(uint64_t(u >> 64));
    const double lo = double(uint64_t(u));
    const double d = hi * 18446744073709551616.0 + lo;
    return neg ? -d : d;
}

// Per-frame parameters, read by the Analysis and Rasterize kernels from device
// memory so that a captured graph of those kernels stays valid from frame to frame.
struct FrameParams {
    remo_shared_uniforms u;
    float arrival;        // fraction of the cloud's batches ingested
    uint32_t pad[3];
};

// One unit of raster work: `count` samples of a draw item's points (pass 0) or
// voxels (pass 1), starting at sample `first`.
struct RasterUnit {
    uint32_t node, pass, first, count;
};

// Launch shape of the per-frame kernels (the autotuner's knobs).
struct FrameShape {
    uint32_t nodeBlocks;    // workgroups of the node-parallel Analysis / selection kernels
    uint32_t rasterBlocks;  // workgroups of the raster kernel
    uint32_t rasterTile;    // samples per raster unit
};

// One item of the frontier draw list (kernels/shared/remo_draw.cuh:13-22).
struct DrawItem {
    uint32_t node;
    uint32_t numPoints;
    uint32_t numVoxels;
    uint32_t level;
};

// A node's analytic shape, kept until its roll-up sums change.
struct ShapeCache {
    float lambda[3];
    float sv, planarity, linearity;
    uint32_t pad[2];
    unsigned long long n;
};

// The incremental roll-up marks a dirty node's ancestors in an LDS bitmap of
// this many node indices; larger node pools always take the full roll-up.
constexpr uint32_t kMarkBits = 262144;

struct AnalysisState {
    remo_node_analysis* results = nullptr;
    RollupSums* rollup = nullptr;
    uint8_t* closed = nullptr;        // closure flag per node (last Analysis)
    uint32_t* closeAt = nullptr;      // batch index from which the node is closed (max over its leaves)
    uint32_t* shapeStale = nullptr;   // 1: the node's record changed since its cached shape was solved
    ShapeCache* shapes = nullptr;     // [maxNodes] cached analytic shape per node
    uint4* markList = nullptr;        // [2 maxNodes] the incremental roll-up's marked nodes
    // the incremental roll-up's slots in device memory, for calls that mark more
    // nodes than its LDS holds ([min(maxNodes, kMarkBits)], child links x8)
    RollupSums* slotSums = nullptr;
    uint32_t* slotNode = nullptr;
    uint32_t* slotFc = nullptr;
    uint32_t* slotClose = nullptr;
    uint32_t* slotChild = nullptr;
    uint32_t* slotNk = nullptr;
    uint32_t* slotOf = nullptr;       // [maxNodes] slot of each marked node
    bool needFull = true;             // the next Analysis rolls up every node
    uint32_t rollLdsSlots = 384;      // incremental roll-up: marked nodes kept in LDS at most
    uint32_t lastD = 0;               // tree depth of the last roll-up
    uint32_t closureMargin = 8;       // batches untouched before a leaf is closed
    float arrival = 0.0f;             // fraction of the cloud's batches ingested (last Analysis)
    uint32_t* queue = nullptr;        // work queue of nodes (score > 0)
    // [0], [1] the counters of the queue's front (a launch appends to the one [2]
    // names, then flips [2]), [3] the workgroups of the launch that have read the
    // dirty list, [4], [5] the counts of the queue's back segment (the rolled-up
    // nodes, from the queue's end backwards) by the same parity
    uint32_t* queueCtl = nullptr;
    uint8_t* visible = nullptr;
    uint8_t* large = nullptr;
    DrawItem* drawList = nullptr;
    uint32_t* drawCounters = nullptr;  // [0] items, [1] overflow, [2..3] samples (u64), [4] units, [5] unit overflow
    uint32_t drawCapacity = 131072;    // REMOLOD_DRAWLIST_CAPACITY, remolod_render.cu:22
    RasterUnit* units = nullptr;
    uint32_t unitCapacity = 1u << 20;
    FrameParams* frame = nullptr;      // device copy of the current frame's parameters
    FrameParams lastFrame{};           // what it holds (host mirror), when frameSet
    bool frameSet = false;
    FrameShape shape{0, 0, 1000};
    // the captured Analysis + Rasterize graph and what it was captured for
    hipGraphExec_t frameExec = nullptr;
    uint64_t frameKey[6] = {0, 0, 0, 0, 0, 0};
    uint32_t maxNodes = 0;
    // host mirrors of the last render's diagnostics
    uint32_t drawItems = 0, drawListOverflow = 0;
    uint64_t drawSamples = 0;
    uint16_t* mlpWeights = nullptr;    // BF16 parameters of the learned score
    bool mlpLoaded = false;
};

using AllocFn = std::function<int(void**, uint64_t, const char*)>;

int analysisAllocate(AnalysisState& a, uint64_t maxNodes, const AllocFn& alloc);
void analysisReset(AnalysisState& a);
void analysisRelease(AnalysisState& a);
// maxLevel: the deepest node level (Globals::maxLevel as last read back);
// arrival: the fraction of the cloud's batches ingested (a learned-score input).
// numNodes: Globals::numNodes as last read back (picks the roll-up's launch shape).
int runAnalysis(AnalysisState& a, const DevTree& t, const remo_shared_uniforms& u, int mode, uint32_t maxLevel,
                uint32_t numNodes, float arrival, hipStream_t s, std::string* err);
// The learned score's 13 inputs of every node from the last Analysis roll-up.
int runFeatures(AnalysisState& a, const DevTree& t, float* out, uint32_t cap, hipStream_t s);
int runRender(AnalysisState& a, const DevTree& t, const remo_shared_uniforms& u, uint64_t* fb, hipStream_t s,
              std::string* err);
// Analysis + Rasterize as one captured graph (the incremental roll-up, the
// finalize and the Rasterize kernels; a full roll-up, when one is due, runs
// eagerly before it), re-captured when the mode, the framebuffer, its size, the
// launch shape or the roll-up's LDS limit change.
int runFrame(AnalysisState& a, const DevTree& t, const remo_shared_uniforms& u, int mode, uint32_t maxLevel,
             uint32_t numNodes, float arrival, uint64_t* fb, hipStream_t s, std::string* err);
void analysisReadDiagnostics(AnalysisState& a, hipStream_t s);
// The same in two halves: queue the copy of the draw counters into pinned host
// memory (6 words), then, after the caller's synchronisation, take them.
void analysisQueueDiagnostics(const AnalysisState& a, uint32_t* host, hipStream_t s);
void analysisTakeDiagnostics(AnalysisState& a, const uint32_t* host);
// The first n roll-up records, each in its own node's frame, into `out` (device).
void launchCentredRollups(const AnalysisState& a, const DevTree& t, RollupSums* out, uint32_t n, hipStream_t s);
// The flat (LOD-free) Rasterize of n resident points, and eye-dome lighting.
void launchRenderFlat(const remo_point* pts, uint64_t n, const remo_shared_uniforms& u, uint64_t* fb, uint32_t blocks,
                      hipStream_t s);
void launchApplyEdl(uint64_t* fb, const remo_shared_uniforms& u, hipStream_t s);

}  // namespace remo_hip

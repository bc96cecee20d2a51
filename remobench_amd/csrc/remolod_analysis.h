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
    uint32_t* rollStamp = nullptr;    // Analysis stamp of the node's last roll-up write
    uint32_t* shapeStamp = nullptr;   // stamp of the sums its cached shape was computed from
    ShapeCache* shapes = nullptr;     // [maxNodes] cached analytic shape per node
    uint32_t* markList = nullptr;     // [maxNodes] the incremental roll-up's marked nodes, then changed leaves
    // the incremental roll-up's slots in device memory, for calls that mark more
    // nodes than its LDS holds ([min(maxNodes, kMarkBits)], child links x8)
    RollupSums* slotSums = nullptr;
    uint32_t* slotNode = nullptr;
    uint32_t* slotFc = nullptr;
    uint32_t* slotClose = nullptr;
    uint32_t* slotChild = nullptr;
    uint32_t* slotNk = nullptr;
    uint32_t* slotOf = nullptr;       // [maxNodes] slot of each marked node
    uint32_t* stamp = nullptr;        // [1] Analysis counter (device; never reset)
    bool needFull = true;             // the next Analysis rolls up every node
    uint32_t rollLdsSlots = 384;      // incremental roll-up: marked nodes kept in LDS at most
    uint32_t lastD = 0;               // tree depth of the last roll-up
    uint32_t closureMargin = 8;       // batches untouched before a leaf is closed
    float arrival = 0.0f;             // fraction of the cloud's batches ingested (last Analysis)
    uint32_t* queue = nullptr;        // work queue of nodes (score > 0)
    uint32_t* queueCount = nullptr;
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

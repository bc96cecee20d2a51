// Per-frame passes over a built tree: Analysis (roll-up, 3x3 eigen, score) and
// Rasterize (frontier selection + 64-bit depth|colour splat).
#pragma once

#include <functional>
#include <string>

#include "remolod_internal.h"

namespace remo_hip {

// Exact roll-up record of one node, in the node's own frame: coordinates are
// integers in units of 2^-(20 + E) of the node extent, centred on the node
// centre, where E = deepest level in the tree - node level. __int128 keeps the
// sums exact (see analysis.hip for the bound).
struct RollupSums {
    __int128 s1[3];
    __int128 s2[6];
    unsigned long long n;
    uint32_t e;
    uint32_t pad;
};

// One item of the frontier draw list (kernels/shared/remo_draw.cuh:13-22).
struct DrawItem {
    uint32_t node;
    uint32_t numPoints;
    uint32_t numVoxels;
    uint32_t level;
};

struct MlpState;

struct AnalysisState {
    remo_node_analysis* results = nullptr;
    RollupSums* rollup = nullptr;
    uint32_t* arrived = nullptr;      // children arrived, per node (bottom-up roll-up)
    uint8_t* closed = nullptr;        // closure flag per node
    uint32_t closureMargin = 8;       // batches untouched before a leaf is closed
    float* learned = nullptr;         // [n * 2] MLP predictions
    uint32_t* queue = nullptr;        // work queue of nodes (score > 0)
    uint32_t* queueCount = nullptr;
    uint8_t* visible = nullptr;
    uint8_t* large = nullptr;
    DrawItem* drawList = nullptr;
    uint32_t* drawCounters = nullptr;  // [0] items, [1] overflow, [2..3] samples (u64)
    uint32_t drawCapacity = 131072;    // REMOLOD_DRAWLIST_CAPACITY, remolod_render.cu:22
    uint32_t maxNodes = 0;
    // host mirrors of the last render's diagnostics
    uint32_t drawItems = 0, drawListOverflow = 0;
    uint64_t drawSamples = 0;
    MlpState* mlp = nullptr;
};

using AllocFn = std::function<int(void**, uint64_t, const char*)>;

int analysisAllocate(AnalysisState& a, uint64_t maxNodes, const AllocFn& alloc);
void analysisReset(AnalysisState& a);
// maxLevel: the deepest node level (Globals::maxLevel as last read back).
int runAnalysis(AnalysisState& a, const DevTree& t, const remo_shared_uniforms& u, int mode, uint32_t maxLevel,
                const LaunchShape& ls, hipStream_t s, std::string* err);
int runRender(AnalysisState& a, const DevTree& t, const remo_shared_uniforms& u, uint64_t* fb,
              const LaunchShape& ls, hipStream_t s, std::string* err);
void analysisReadDiagnostics(AnalysisState& a, hipStream_t s);

}  // namespace remo_hip

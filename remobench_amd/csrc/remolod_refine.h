// Refinement, colour filtering: deterministic voxel colours from sparse per-grid
// cell sums (see refine.hip).
#pragma once

#include <functional>
#include <string>

#include "remolod_internal.h"

namespace remo_hip {

constexpr uint32_t kRefineBlockWords = 8;                             // words per rank block
constexpr uint32_t kRefineBlocks = kGridWords / kRefineBlockWords;   // 8,192 blocks per grid
constexpr uint32_t kRefineLevels = uint32_t(kMaxDepth) + 1;

// Device counters of one call (zeroed at its start, except arenaUsed).
struct RefineCounters {
    uint32_t levelCount[kRefineLevels];              // dirty gridded nodes per level
    uint32_t levelFill[kRefineLevels];               // list scatter cursors
    uint32_t levelKids[kRefineLevels];               // their inner children
    uint32_t levelKidFill[kRefineLevels];            // inner children listed by the prep pass
    unsigned long long levelPoints[kRefineLevels];   // their leaf points
    unsigned long long levelVoxels[kRefineLevels];   // their voxels
    unsigned long long pointsFolded, childSums, voxelsWritten;
    uint32_t refined, skipped, missingCells, emptyCells, arenaFull, cellOverflow;
    unsigned long long arenaUsed;                    // accumulator entries handed out (persistent)
};

// The list offsets of the levels being processed (kernel argument by value).
struct RefineLevelOffsets {
    uint32_t off[kRefineLevels];
    uint32_t lo, hi;  // levels [lo, hi] are listed
};

struct RefineState {
    bool allocated = false;
    // per node
    uint32_t* refPoints = nullptr;   // numPoints folded at the node's parent's last refinement
    uint32_t* refVoxels = nullptr;   // cells written by the node's last refinement, kNone: no valid sums
    uint32_t* dirty = nullptr;
    uint32_t* status = nullptr;      // 1 while a node's refinement in this call is sound
    uint32_t* cells = nullptr;       // occupied cells found by this call's prefix pass
    uint32_t* list = nullptr;        // dirty gridded nodes, grouped by level
    uint32_t* kids = nullptr;        // their inner children, grouped by level
    // per grid
    uint32_t* prefix = nullptr;      // [grid][kRefineBlocks] set bits before each block
    unsigned long long* accOff = nullptr;  // first accumulator entry of the grid's sums
    uint32_t* accCap = nullptr;      // entries reserved (0: none)
    // the accumulator store: three u64 per entry (r + g<<32, b + a<<32, n)
    unsigned long long* acc = nullptr;
    uint64_t accEntries = 0;
    RefineCounters* ctr = nullptr;
    RefineCounters* host = nullptr;  // pinned mirror
    hipEvent_t t0 = nullptr, t1 = nullptr;
    double modelScale = 1.0;         // measured / estimated device time of recent calls (running)
    double overheadMs = 0.3;         // a call's wall time beyond its timed kernels (a decaying peak)
};

using RefineAllocFn = std::function<int(void**, uint64_t, const char*)>;

struct RefineStatsOut {
    uint32_t levelsPending, levelsProcessed, nodesDirty, nodesRefined, nodesSkipped, complete;
    uint64_t pointsFolded, childSums, voxelsWritten;
    uint32_t missingCells, emptyCells, arenaFull, cellOverflow;
    float estimateMs, deviceMs;
};

// First use: the per-node and per-grid arrays and an accumulator store of
// `entries` entries (allocated through `alloc`, owned by the caller's list).
int refineAllocate(RefineState& r, const DevTree& t, uint64_t entries, const RefineAllocFn& alloc,
                   hipStream_t s);
// Forget every refinement (tree reset): all sums invalid, store empty.
void refineReset(RefineState& r, const DevTree& t, hipStream_t s);
// Replaces the store by one of `entries` entries and forgets every refinement
// (all sums invalid, so the next call recomputes every gridded node).
int refineGrow(RefineState& r, const DevTree& t, uint64_t entries, hipStream_t s, std::string* err);
// runRefine's answer when the dirty nodes' sums would not fit beside the live
// ones: replace the store (refineGrow) and call again. Nothing was changed.
constexpr int kRefineReplaceStore = 1;
// One Refinement call (see remo_refine). numNodes: Globals::numNodes as last read.
int runRefine(RefineState& r, const DevTree& t, uint32_t numNodes, float budgetMs, uint32_t maxLevels,
              const LaunchShape& ls, hipStream_t s, RefineStatsOut* out, std::string* err);
void refineRelease(RefineState& r);

}  // namespace remo_hip

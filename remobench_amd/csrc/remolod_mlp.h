// Learned score: a 13 -> 64 -> 64 -> 2 GELU MLP over a node's partial accumulators.
#pragma once

#include "remolod_analysis.h"

namespace remo_hip {

int mlpAllocate(AnalysisState& a, uint64_t maxNodes, const AllocFn& alloc);
int mlpPredict(AnalysisState& a, const DevTree& t, const LaunchShape& ls, hipStream_t s, std::string* err);

}  // namespace remo_hip

// Learned score (filled in below).
#include "remolod_mlp.h"

namespace remo_hip {

int mlpAllocate(AnalysisState& a, uint64_t maxNodes, const AllocFn& alloc) {
    (void)a;
    (void)maxNodes;
    (void)alloc;
    return REMO_OK;
}

int mlpPredict(AnalysisState& a, const DevTree& t, const LaunchShape& ls, hipStream_t s, std::string* err) {
    (void)a;
    (void)t;
    (void)ls;
    (void)s;
    if (err) *err = "no trained score model loaded";
    return REMO_ERR_STATE;
}

}  // namespace remo_hip

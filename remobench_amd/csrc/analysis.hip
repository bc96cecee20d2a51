// Analysis and Rasterize (filled in below).
#include "remolod_analysis.h"

namespace remo_hip {

int analysisAllocate(AnalysisState& a, uint64_t maxNodes, const AllocFn& alloc) {
    a.maxNodes = uint32_t(maxNodes);
    void* p = nullptr;
    int rc = alloc(&p, sizeof(remo_node_analysis) * maxNodes, "analysis results");
    a.results = static_cast<remo_node_analysis*>(p);
    return rc;
}

void analysisReset(AnalysisState& a) {
    a.drawItems = 0;
    a.drawListOverflow = 0;
    a.drawSamples = 0;
}

int runAnalysis(AnalysisState&, const DevTree&, const remo_shared_uniforms&, int, const LaunchShape&,
                hipStream_t, std::string* err) {
    if (err) *err = "analysis not built";
    return REMO_ERR_STATE;
}

int runRender(AnalysisState&, const DevTree&, const remo_shared_uniforms&, uint64_t*, const LaunchShape&,
              hipStream_t, std::string* err) {
    if (err) *err = "render not built";
    return REMO_ERR_STATE;
}

void analysisReadDiagnostics(AnalysisState&, hipStream_t) {}

}  // namespace remo_hip

// The `remolod` plugin on the HIP path: a from-scratch remo::ILodPipeline over
// the C ABI (remolod.h), with the lifecycle of the reference's RemolodPipeline
// (src/pipelines/RemolodPipeline.{h,cpp}).
#pragma once

#include <string>
#include <vector>

#include "remo/ILodPipeline.h"

namespace remo {

class RemolodHipPipeline final : public ILodPipeline {
public:
    // Sizing constants of RemolodPipeline::allocate (src/pipelines/RemolodPipeline.cpp:105-150).
    static constexpr uint64_t kMomentaryBytes = 512ull << 20;
    static constexpr double kBytesPerPoint = 48.0;
    static constexpr double kMinBytesPerPoint = 26.0;
    static constexpr uint64_t kMinPersistentBytes = 512ull << 20;
    static constexpr uint32_t kMaxNodes = REMO_DEFAULT_MAX_NODES;

    struct Options {
        bool accumulate = true;             // --remolod-no-accum turns it off
        uint32_t batchesPerStep = 20;       // kernel_construct's per-launch batch cap
        float stepBudgetMs = 10.0f;         // its %globaltimer budget; < 0: none
        int analysisMode = -1;              // -1: no Analysis in build(); else REMO_ANALYSIS_*
        uint32_t closureBatches = 8;
        uint32_t slotCapacity = kSlotCapacity;  // ring stride; the source must use the same
        uint32_t fixedDepth = 0;            // --remolod-fixed-depth: the analytic-bucketing arm
    };

    RemolodHipPipeline() = default;
    explicit RemolodHipPipeline(const Options& o) : m_opt(o) {}
    ~RemolodHipPipeline() override { release(); }

    PipelineInfo info() const override;
    bool initPrograms(std::string* err) override;
    bool allocate(const CloudMeta& meta, const DeviceBudget& budget, std::string* err) override;
    void release() override;
    void reset() override;
    bool build(PointSource& source, const FrameContext& frame) override;
    void render(const FrameContext& frame) override;
    const PipelineStats& stats() const override { return m_stats; }
    TimingScopes timingScopes() const override;
    std::vector<std::string> diagnostics() const override;
    std::vector<ScopeTotal> scopeTotals() const override;

    // The per-frame Analysis (no reference code; README.md:80): analytic or learned.
    bool analyse(const FrameContext& frame, int mode, std::string* err);
    // Refinement's colour filtering within budgetMs (no reference code; README.md:81):
    // runs only when budget remains, may change nothing.
    bool refine(float budgetMs, remo_refine_stats* out, std::string* err);
    // Refinement's local deepening: split up to maxSplits of the densest leaves.
    bool deepen(uint32_t maxSplits, uint32_t minPoints, remo_deepen_stats* out, std::string* err);
    // Refinement's collapsing: merge up to maxCollapses of the sparsest inner nodes
    // whose children are all leaves (at most maxPoints points together).
    bool collapse(uint32_t maxCollapses, uint32_t maxPoints, remo_collapse_stats* out, std::string* err);

    remo_ctx* context() const { return m_ctx; }
    const remo_stats& rawStats() const { return m_raw; }
    uint32_t batchesConsumed() const { return m_batchesConsumed; }
    bool complete() const { return m_complete; }
    const std::string& lastError() const { return m_lastError; }

private:
    void readStats();   // remo_get_stats (synchronous), then takeStats
    void takeStats();   // m_raw into m_stats
    void fail(const char* what, int rc);

    Options m_opt;
    remo_ctx* m_ctx = nullptr;
    uint64_t m_persistentBytes = 0;
    bool m_needsReset = true;
    bool m_complete = false;
    bool m_ringMismatchReported = false;
    bool m_profiling = false;
    uint32_t m_batchesConsumed = 0;
    uint32_t m_batchesTotal = 0;
    PipelineStats m_stats;
    remo_stats m_raw{};
    std::string m_lastError;
};

}  // namespace remo

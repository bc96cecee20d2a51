// RemolodHipPipeline: the remolod plugin over the C ABI.
#include "RemolodHipPipeline.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>

namespace remo {

PipelineInfo RemolodHipPipeline::info() const {
    PipelineInfo i;
    i.id = "remolod";
    i.displayName = "RemoLOD (HIP, gfx950)";
    i.progressive = true;
    i.needsWholeCloudResident = false;
    i.ringSlots = REMO_DEFAULT_RING_SLOTS;
    i.bytesPerPointEstimate = kBytesPerPoint;
    i.minBytesPerPointEstimate = kMinBytesPerPoint;
    i.minStoreBytes = kMinPersistentBytes;
    i.fixedBytesEstimate = kMomentaryBytes + uint64_t(kMaxNodes) * (152 + 112);
    return i;
}

bool RemolodHipPipeline::initPrograms(std::string* err) {
    // Code objects are linked into libremolod ahead of time: nothing to compile.
    // What can fail is the device itself.
    int n = 0;
    if (remo_device_count(&n) != REMO_OK || n < 1) {
        if (err) *err = "no HIP device visible";
        return false;
    }
    return true;
}

bool RemolodHipPipeline::allocate(const CloudMeta& meta, const DeviceBudget& budget, std::string* err) {
    // src/pipelines/RemolodPipeline.cpp:105-150: the persistent store takes what the
    // budget leaves after the fixed terms, capped at the per-point appetite.
    release();
    if (meta.numPoints == 0) {
        if (err) *err = "cloud is empty";
        return false;
    }
    // the ring's input term comes out of the shared budget first
    const uint64_t batches = (meta.numPoints + m_opt.slotCapacity - 1) / m_opt.slotCapacity;
    const uint64_t inputBytes =
        std::min<uint64_t>(REMO_DEFAULT_RING_SLOTS, batches) * m_opt.slotCapacity * sizeof(remo_point);
    const uint64_t fixed = uint64_t(kMaxNodes) * (152 + 112) + kMomentaryBytes;
    const uint64_t want = std::max<uint64_t>(uint64_t(kBytesPerPoint * double(meta.numPoints)), kMinPersistentBytes);
    const uint64_t available = budget.bytes > inputBytes ? budget.bytes - inputBytes : 0;
    if (available < fixed + kMinPersistentBytes) {
        if (err)
            *err = "not enough device memory: needs at least " +
                   std::to_string((fixed + kMinPersistentBytes) >> 20) + " MB, " +
                   std::to_string(available >> 20) + " MB available";
        return false;
    }
    // The fixed-depth arm pre-splits, and every inner node keeps a 256 KB grid, so
    // it is not sized by the per-point coefficient: it takes what the budget
    // leaves (src/pipelines/RemolodPipeline.cpp:137-147).
    m_persistentBytes = m_opt.fixedDepth ? available - fixed : std::min(want, available - fixed);

    remo_config cfg{};
    cfg.persistent_bytes = m_persistentBytes;
    cfg.max_nodes = kMaxNodes;
    cfg.max_batches_per_step = m_opt.batchesPerStep;
    cfg.step_budget_ms = m_opt.stepBudgetMs;
    cfg.slot_capacity = m_opt.slotCapacity;
    cfg.ring_slots = REMO_DEFAULT_RING_SLOTS;
    cfg.flags = m_opt.accumulate ? 0u : REMO_FLAG_NO_ACCUM;
    cfg.closure_batches = m_opt.closureBatches;
    cfg.fixed_depth = m_opt.fixedDepth;
    const remo_cloud_meta m = meta.toAbi();
    const int rc = remo_create(&cfg, &m, &m_ctx);
    if (rc != REMO_OK) {
        if (err) *err = std::string("remo_create failed (") + std::to_string(rc) + ")";
        m_ctx = nullptr;
        return false;
    }
    m_stats = PipelineStats{};
    m_stats.bytesCapacity = budget.bytes;
    m_needsReset = true;
    m_complete = false;
    return true;
}

void RemolodHipPipeline::release() {
    if (m_ctx) remo_destroy(m_ctx);
    m_ctx = nullptr;
    m_complete = false;
    m_needsReset = true;
}

void RemolodHipPipeline::reset() {
    m_needsReset = true;
    m_complete = false;
    m_batchesConsumed = 0;
}

void RemolodHipPipeline::fail(const char* what, int rc) {
    m_lastError = std::string(what) + ": " + (m_ctx ? remo_last_error(m_ctx) : "no context") + " (" +
                  std::to_string(rc) + ")";
    if (rc == REMO_ERR_DEVICE) {
        // reportDeadContextAndExit (src/cuda/CudaCheck.cpp:8-35): a sticky device
        // fault leaves nothing to recover; fail fast with the same exit status.
        std::fprintf(stderr, "remobench: device context is dead after %s\n", m_lastError.c_str());
        std::exit(70);
    }
}

bool RemolodHipPipeline::build(PointSource& source, const FrameContext& frame) {
    // src/pipelines/RemolodPipeline.cpp:296-438.
    if (!m_ctx) return false;
    if (frame.profile != m_profiling) {
        remo_profile_enable(m_ctx, frame.profile ? 0xFFFFFFFFu : 0u);
        m_profiling = frame.profile;
    }
    if (m_needsReset) {
        const int rc = remo_reset(m_ctx);
        if (rc != REMO_OK) {
            fail("reset", rc);
            return false;
        }
        if (m_profiling) remo_profile_clear(m_ctx);
        m_needsReset = false;
        m_complete = false;
        m_batchesConsumed = 0;
        // the device batch cursor is back at zero, so the producer goes back too
        source.rewind();
        return true;
    }
    if (m_complete) return false;

    const BatchView view = source.view();
    if (!view.slots || !view.batchSizes || !view.numBatchesUploaded) return true;
    m_batchesTotal = view.numBatchesTotal;
    const remo_batch_view v = view.toAbi();
    uint32_t consumed = 0;
    const int rc = remo_build_step(m_ctx, &v, &consumed);
    if (rc == REMO_ERR_RING) {
        if (!m_ringMismatchReported) {
            m_ringMismatchReported = true;
            std::fprintf(stderr, "remobench: RemoLOD was handed a %u-slot ring but the context wraps at %u; "
                                 "refusing to build from it.\n", view.numSlots, REMO_DEFAULT_RING_SLOTS);
        }
        return false;
    }
    if (rc < 0) {
        fail("build_step", rc);
        return false;
    }
    m_batchesConsumed = consumed;
    source.setBatchesConsumed(consumed);
    if (m_opt.accumulate) {
        const int ra = remo_accumulate(m_ctx, consumed);
        if (ra != REMO_OK) {
            fail("accumulate", ra);
            return false;
        }
    }
    // the stats of this build and Accumulate, read while Analysis runs
    const bool statsQueued = remo_request_stats(m_ctx) == REMO_OK;
    if (m_opt.analysisMode >= 0) {
        std::string e;
        if (!analyse(frame, m_opt.analysisMode, &e)) return false;
    }
    if (statsQueued && remo_collect_stats(m_ctx, &m_raw) == REMO_OK)
        takeStats();
    else
        readStats();
    if (m_batchesTotal > 0 && m_batchesConsumed >= m_batchesTotal) {
        m_complete = true;
        return false;
    }
    return true;
}

bool RemolodHipPipeline::analyse(const FrameContext& frame, int mode, std::string* err) {
    if (!m_ctx || m_needsReset) return false;
    const int rc = remo_analysis(m_ctx, &frame.uniforms, mode);
    if (rc != REMO_OK) {
        fail("analysis", rc);
        if (err) *err = m_lastError;
        return false;
    }
    return true;
}

bool RemolodHipPipeline::refine(float budgetMs, remo_refine_stats* out, std::string* err) {
    if (!m_ctx || m_needsReset || !out) return false;
    const int rc = remo_refine(m_ctx, budgetMs, 0, out);
    if (rc != REMO_OK) {
        fail("refine", rc);
        if (err) *err = m_lastError;
        return false;
    }
    return true;
}

bool RemolodHipPipeline::deepen(uint32_t maxSplits, uint32_t minPoints, remo_deepen_stats* out, std::string* err) {
    if (!m_ctx || m_needsReset || !out) return false;
    const int rc = remo_deepen(m_ctx, maxSplits, minPoints, out);
    if (rc != REMO_OK) {
        fail("deepen", rc);
        if (err) *err = m_lastError;
        return false;
    }
    readStats();
    return true;
}

bool RemolodHipPipeline::collapse(uint32_t maxCollapses, uint32_t maxPoints, remo_collapse_stats* out,
                                  std::string* err) {
    if (!m_ctx || m_needsReset || !out) return false;
    const int rc = remo_collapse(m_ctx, maxCollapses, maxPoints, out);
    if (rc != REMO_OK) {
        fail("collapse", rc);
        if (err) *err = m_lastError;
        return false;
    }
    readStats();
    return true;
}

void RemolodHipPipeline::render(const FrameContext& frame) {
    // src/pipelines/RemolodPipeline.cpp:559-601.
    if (!m_ctx || m_needsReset || !frame.targets.framebuffer) return;
    const int rc = remo_render(m_ctx, &frame.uniforms, frame.targets.framebuffer);
    if (rc != REMO_OK) {
        fail("render", rc);
        return;
    }
    if (frame.strictTiming) remo_sync(m_ctx);
    uint32_t items = 0, overflow = 0;
    uint64_t samples = 0;
    if (remo_get_render_diagnostics(m_ctx, &items, &samples, &overflow) == REMO_OK) {
        m_stats.numVisibleNodes = items;
        m_stats.numVisiblePoints = samples;
        if (overflow) m_stats.allocOverflow = true;
    }
}

void RemolodHipPipeline::readStats() {
    if (remo_get_stats(m_ctx, &m_raw) != REMO_OK) return;
    takeStats();
}

void RemolodHipPipeline::takeStats() {
    // RemolodPipeline::readStats (src/pipelines/RemolodPipeline.cpp:440-470); the
    // visible counts belong to render() and are left alone here.
    m_batchesConsumed = m_raw.batchlet_index;
    m_stats.numPoints = m_raw.num_points;
    m_stats.numVoxels = m_raw.num_voxels;
    m_stats.numPointsIngested = m_raw.num_points_processed;
    m_stats.numNodes = m_raw.num_nodes;
    m_stats.numInner = m_raw.num_inner;
    m_stats.numLeaves = m_raw.num_leaves;
    m_stats.maxDepth = m_raw.max_depth;
    m_stats.maxPointsPerNode = m_raw.max_points_per_node;
    m_stats.bytesAllocated = m_raw.allocated_bytes_persistent;
    m_stats.memCapacityReached = m_raw.mem_capacity_reached != 0;
    m_stats.nodeCapacityReached = m_raw.node_capacity_reached != 0 || m_raw.num_nodes >= kMaxNodes;
    m_stats.allocOverflow = m_stats.allocOverflow || m_raw.alloc_overflow != 0;
}

TimingScopes RemolodHipPipeline::timingScopes() const {
    TimingScopes t;
    t.build = {"remolod.construct.count", "remolod.construct.split", "remolod.construct.voxelSampling",
               "remolod.construct.allocChunks", "remolod.construct.insertPoints",
               "remolod.construct.insertVoxels", "remolod.accumulate"};
    t.render = "remolod.render";
    return t;
}

std::vector<ScopeTotal> RemolodHipPipeline::scopeTotals() const {
    std::vector<ScopeTotal> out;
    if (!m_ctx) return out;
    remo_profile_entry e[16];
    uint32_t n = 0;
    if (remo_profile_read(m_ctx, e, 16, &n) != REMO_OK) return out;
    for (uint32_t i = 0; i < n && i < 16; ++i) out.push_back({e[i].name, e[i].total_ms, e[i].launches});
    return out;
}

std::vector<std::string> RemolodHipPipeline::diagnostics() const {
    // RemolodPipeline::diagnostics (src/pipelines/RemolodPipeline.cpp:603-726):
    // tab-separated label/value lines, the node pool first.
    std::vector<std::string> out;
    char buf[160];
    if (m_raw.node_pool_overflow) out.push_back("node pool\tOVERFLOW (raise MAX_NODES_CAPACITY)");
    if (m_raw.counter_underflows == 0)
        out.push_back("leaf counter >= pts\tok");
    else {
        std::snprintf(buf, sizeof buf, "leaf counter >= pts\tVIOLATED in %u leaves (counter < numPoints)",
                      m_raw.counter_underflows);
        out.push_back(buf);
    }
    if (!m_opt.accumulate) {
        out.push_back("accumulator\toff (--remolod-no-accum)");
        return out;
    }
    const bool agree = m_raw.sum_leaf_counts == m_raw.num_points;
    std::snprintf(buf, sizeof buf, "accum sum/points\t%llu / %llu  %s",
                  (unsigned long long)m_raw.sum_leaf_counts, (unsigned long long)m_raw.num_points,
                  agree ? "ok" : "MISMATCH");
    out.push_back(buf);
    std::snprintf(buf, sizeof buf, "accum inner w/ sums\t%u  %s", m_raw.inner_with_sums,
                  m_raw.inner_with_sums == 0 ? "ok" : "MISMATCH");
    out.push_back(buf);
    std::snprintf(buf, sizeof buf, "accum folded total\t%llu", (unsigned long long)m_raw.num_accumulated);
    out.push_back(buf);
    std::snprintf(buf, sizeof buf, "accum watermark\t0x%llx", (unsigned long long)m_raw.morton_watermark);
    out.push_back(buf);
    return out;
}

}  // namespace remo

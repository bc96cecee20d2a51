// PipelineRegistry over ILodPipeline factories (src/shell/PipelineRegistry.cpp).
#include "remo/PipelineRegistry.h"

#include <algorithm>
#include <cstdio>

#include "RemolodHipPipeline.h"

namespace remo {
namespace {

// The persistent store's safety margin: ingest stops this far below capacity
// (kernels/remolod/remolod_octree.cu:1181-1210).
constexpr uint64_t kPersistentSafetyMargin = 200'000'000ull;

uint64_t inputBytesFor(const PipelineInfo& info, const CloudMeta& meta) {
    if (!info.needsWholeCloudResident && info.ringSlots > 0) {
        const uint64_t total = (meta.numPoints + kSlotCapacity - 1) / kSlotCapacity;
        return std::min<uint64_t>(info.ringSlots, total) * kSlotCapacity * sizeof(remo_point);
    }
    return meta.numPoints * sizeof(remo_point);
}

}  // namespace

void PipelineRegistry::add(PipelineFactory factory) {
    if (!factory) return;
    std::unique_ptr<ILodPipeline> probe = factory();
    if (!probe) {
        std::fprintf(stderr, "remobench: pipeline factory returned null at registration\n");
        return;
    }
    PipelineInfo info = probe->info();
    probe.reset();
    if (info.id.empty()) {
        std::fprintf(stderr, "remobench: pipeline has an empty id; not registered\n");
        return;
    }
    if (find(info.id)) {
        std::fprintf(stderr, "remobench: duplicate pipeline id '%s'; not registered\n", info.id.c_str());
        return;
    }
    m_infos.push_back(info);
    m_entries.push_back({std::move(info), std::move(factory)});
}

const PipelineInfo* PipelineRegistry::find(const std::string& id) const {
    for (const Entry& e : m_entries)
        if (e.info.id == id) return &e.info;
    return nullptr;
}

uint64_t PipelineRegistry::minBytesRequired(const PipelineInfo& info, const CloudMeta& meta) const {
    return uint64_t(info.minBytesPerPointEstimate * double(meta.numPoints)) + inputBytesFor(info, meta) +
           info.fixedBytesEstimate;
}

uint64_t PipelineRegistry::refusalFloor(const PipelineInfo& info, const CloudMeta& meta) const {
    if (info.minStoreBytes == 0) return minBytesRequired(info, meta);
    return info.minStoreBytes + inputBytesFor(info, meta) + info.fixedBytesEstimate;
}

bool PipelineRegistry::fits(const PipelineInfo& info, const CloudMeta& meta, const DeviceBudget& budget) const {
    return refusalFloor(info, meta) <= budget.bytes;
}

double PipelineRegistry::predictedCompletion(const PipelineInfo& info, const CloudMeta& meta,
                                             const DeviceBudget& budget) const {
    if (meta.numPoints == 0) return 1.0;
    if (info.minBytesPerPointEstimate <= 0.0) return fits(info, meta, budget) ? 1.0 : 0.0;
    const uint64_t overhead = inputBytesFor(info, meta) + info.fixedBytesEstimate;
    if (budget.bytes <= overhead) return 0.0;
    const uint64_t want = uint64_t(info.bytesPerPointEstimate * double(meta.numPoints));
    const uint64_t persistent = std::min<uint64_t>(want, budget.bytes - overhead);
    const uint64_t usable = persistent > kPersistentSafetyMargin ? persistent - kPersistentSafetyMargin : 0;
    return std::clamp(double(usable) / info.minBytesPerPointEstimate / double(meta.numPoints), 0.0, 1.0);
}

std::string PipelineRegistry::unsupportedReason(const PipelineInfo& info, const CloudMeta& meta,
                                                const DeviceBudget& budget) const {
    if (fits(info, meta, budget)) return {};
    char buf[384];
    std::snprintf(buf, sizeof buf,
                  "cannot run at all below %.1f GB (%.1f GB %s + %.2f GB fixed + %.2f GB smallest store); "
                  "the budget is %.1f GB",
                  double(refusalFloor(info, meta)) / 1e9, double(inputBytesFor(info, meta)) / 1e9,
                  info.needsWholeCloudResident ? "resident cloud" : "ring", double(info.fixedBytesEstimate) / 1e9,
                  double(info.minStoreBytes) / 1e9, double(budget.bytes) / 1e9);
    return buf;
}

void PipelineRegistry::releaseActive() {
    if (m_active) m_active->release();
    m_active.reset();
    m_activeId.clear();
}

bool PipelineRegistry::switchTo(const std::string& id, PointSource* source, const CloudMeta& meta,
                                const DeviceBudget& budget, std::string* err) {
    const Entry* entry = nullptr;
    for (const Entry& e : m_entries)
        if (e.info.id == id) entry = &e;
    if (!entry) {
        if (err) *err = "unknown pipeline: " + id;
        return false;
    }
    if (const std::string reason = unsupportedReason(entry->info, meta, budget); !reason.empty()) {
        if (err) *err = entry->info.displayName + ": " + reason;
        return false;
    }
    releaseActive();
    // the ingest shape is the incoming pipeline's requirement, decided before it
    // allocates from the same budget
    if (source) {
        const PointSource::Mode mode =
            entry->info.needsWholeCloudResident ? PointSource::Mode::Whole : PointSource::Mode::Stream;
        if (!source->start(mode, entry->info.ringSlots, budget.bytes, err)) return false;
        source->rewind();
    }
    std::unique_ptr<ILodPipeline> pipeline = entry->factory();
    if (!pipeline) {
        if (err) *err = "factory returned null for " + id;
        return false;
    }
    if (!pipeline->initPrograms(err)) return false;
    if (!pipeline->allocate(meta, budget, err)) return false;
    pipeline->reset();
    m_active = std::move(pipeline);
    m_activeId = id;
    return true;
}

void registerHipPipelines(PipelineRegistry& registry) {
    registry.add([] { return std::make_unique<RemolodHipPipeline>(); });
}

}  // namespace remo

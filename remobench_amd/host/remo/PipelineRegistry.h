// Registry of pipeline factories and the switch lifecycle, with the method set of
// the reference's remo::PipelineRegistry (include/remo/PipelineRegistry.h:13-80,
// src/shell/PipelineRegistry.cpp): add(factory) keyed by info().id, the budget
// checks (minBytesRequired / refusalFloor / fits / predictedCompletion /
// unsupportedReason) and switchTo: release the old pipeline -> source.start(mode,
// ringSlots, budget) -> factory() -> initPrograms -> allocate -> reset.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "remo/ILodPipeline.h"

namespace remo {

class PipelineRegistry {
public:
    // Refuses (with a stderr line) a null factory, an empty id or a duplicate id.
    void add(PipelineFactory factory);

    const std::vector<PipelineInfo>& list() const { return m_infos; }
    ILodPipeline* active() const { return m_active.get(); }
    const std::string& activeId() const { return m_activeId; }
    const PipelineInfo* find(const std::string& id) const;

    uint64_t minBytesRequired(const PipelineInfo& info, const CloudMeta& meta) const;
    uint64_t refusalFloor(const PipelineInfo& info, const CloudMeta& meta) const;
    bool fits(const PipelineInfo& info, const CloudMeta& meta, const DeviceBudget& budget) const;
    double predictedCompletion(const PipelineInfo& info, const CloudMeta& meta, const DeviceBudget& budget) const;
    std::string unsupportedReason(const PipelineInfo& info, const CloudMeta& meta,
                                  const DeviceBudget& budget) const;

    // On failure the registry holds no pipeline and *err says why.
    bool switchTo(const std::string& id, PointSource* source, const CloudMeta& meta, const DeviceBudget& budget,
                  std::string* err);
    void releaseActive();

private:
    struct Entry {
        PipelineInfo info;
        PipelineFactory factory;
    };
    std::vector<Entry> m_entries;
    std::vector<PipelineInfo> m_infos;
    std::unique_ptr<ILodPipeline> m_active;
    std::string m_activeId;
};

// Registers the HIP `remolod` pipeline.
void registerHipPipelines(PipelineRegistry& registry);

}  // namespace remo

// The set of pipelines a shell can switch between, keyed by PipelineInfo::id.
//
// Interface kept from the reference (include/remo/PipelineRegistry.h): add a
// factory, look pipelines up, switch the active one. The switch follows the
// plugin lifecycle the reference documents (wiki/01_Overview.md:135-173): the
// old pipeline is released before anything of the new one exists, the point
// source is configured for the new pipeline's ingest shape, then the new
// pipeline is built, initialised, allocated and reset. Whether a pipeline fits
// the device budget is decided by its own allocate(), which refuses with a
// reason; the reference's separate budget estimates are not part of this path.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "remo/ILodPipeline.h"

namespace remo {

class PipelineRegistry {
public:
    // Probes the factory once for its PipelineInfo. Returns false (and says why on
    // stderr) when the factory yields nothing, the id is empty or already taken.
    bool add(PipelineFactory factory);

    // Registered pipelines, in registration order.
    const std::vector<PipelineInfo>& list() const { return m_order; }
    const PipelineInfo* find(const std::string& id) const;

    ILodPipeline* active() const { return m_active.get(); }
    const std::string& activeId() const { return m_activeId; }

    // Makes `id` the active pipeline for `meta` within `budget`. An unknown id
    // changes nothing; a step that fails after the old pipeline was released
    // leaves no pipeline active. *err names the step that failed.
    bool switchTo(const std::string& id, PointSource* source, const CloudMeta& meta, const DeviceBudget& budget,
                  std::string* err);
    void releaseActive();

private:
    struct Slot {
        PipelineInfo info;
        PipelineFactory make;
    };
    std::map<std::string, Slot> m_slots;
    std::vector<PipelineInfo> m_order;
    std::unique_ptr<ILodPipeline> m_active;
    std::string m_activeId;
};

// Registers the HIP `remolod` pipeline.
void registerHipPipelines(PipelineRegistry& registry);

}  // namespace remo

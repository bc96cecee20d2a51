// The pipeline registry: factories keyed by id and the switch between them.
#include "remo/PipelineRegistry.h"

#include <cstdio>

#include "RemolodHipPipeline.h"

namespace remo {

bool PipelineRegistry::add(PipelineFactory factory) {
    std::unique_ptr<ILodPipeline> sample = factory ? factory() : nullptr;
    if (!sample) {
        std::fprintf(stderr, "pipeline registry: a factory produced no pipeline; skipped\n");
        return false;
    }
    PipelineInfo info = sample->info();
    sample.reset();  // nothing of it was initialised or allocated
    if (info.id.empty() || m_slots.count(info.id) != 0) {
        std::fprintf(stderr, "pipeline registry: id \"%s\" is %s; skipped\n", info.id.c_str(),
                     info.id.empty() ? "empty" : "already registered");
        return false;
    }
    m_order.push_back(info);
    std::string key = info.id;  // not info.id itself: the Slot below takes info apart
    m_slots.emplace(std::move(key), Slot{std::move(info), std::move(factory)});
    return true;
}

const PipelineInfo* PipelineRegistry::find(const std::string& id) const {
    const auto it = m_slots.find(id);
    return it == m_slots.end() ? nullptr : &it->second.info;
}

void PipelineRegistry::releaseActive() {
    if (m_active) m_active->release();
    m_active.reset();
    m_activeId.clear();
}

bool PipelineRegistry::switchTo(const std::string& id, PointSource* source, const CloudMeta& meta,
                                const DeviceBudget& budget, std::string* err) {
    std::string why;
    auto refuse = [&](const std::string& step) {
        if (err) *err = "switch to \"" + id + "\": " + step + (why.empty() ? "" : ": " + why);
        return false;
    };
    const auto it = m_slots.find(id);
    if (it == m_slots.end()) return refuse("no such pipeline");
    const Slot& slot = it->second;

    // the outgoing pipeline gives its device memory back before the incoming one
    // takes any from the same budget
    releaseActive();
    if (source) {
        const PointSource::Mode mode =
            slot.info.needsWholeCloudResident ? PointSource::Mode::Whole : PointSource::Mode::Stream;
        if (!source->start(mode, slot.info.ringSlots, budget.bytes, &why)) return refuse("point source");
        source->rewind();
    }
    std::unique_ptr<ILodPipeline> next = slot.make();
    if (!next) return refuse("factory");
    if (!next->initPrograms(&why)) return refuse("initPrograms");
    if (!next->allocate(meta, budget, &why)) return refuse("allocate");
    next->reset();
    m_active = std::move(next);
    m_activeId = id;
    return true;
}

void registerHipPipelines(PipelineRegistry& registry) {
    registry.add([] { return std::make_unique<RemolodHipPipeline>(); });
}

}  // namespace remo

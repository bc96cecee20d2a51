// Test driver for the C++ pipeline registry (no device): registers fake
// pipelines and a fake point source that log every lifecycle call, runs the
// switches the test names, and prints one JSON line with the log.
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "remo/PipelineRegistry.h"

namespace {

std::vector<std::string> g_log;

class FakePipeline final : public remo::ILodPipeline {
public:
    FakePipeline(std::string id, bool failAllocate) : m_id(std::move(id)), m_failAllocate(failAllocate) {
        g_log.push_back("make " + m_id);
    }
    ~FakePipeline() override { g_log.push_back("destroy " + m_id); }
    remo::PipelineInfo info() const override {
        remo::PipelineInfo i;
        i.id = m_id;
        i.displayName = "fake " + m_id;
        i.ringSlots = 7;
        i.needsWholeCloudResident = m_id == "whole";
        return i;
    }
    bool initPrograms(std::string*) override {
        g_log.push_back("init " + m_id);
        return true;
    }
    bool allocate(const remo::CloudMeta& meta, const remo::DeviceBudget& budget, std::string* err) override {
        g_log.push_back("allocate " + m_id + " " + std::to_string(meta.numPoints) + " " + std::to_string(budget.bytes));
        if (m_failAllocate && err) *err = "budget too small";
        return !m_failAllocate;
    }
    void release() override { g_log.push_back("release " + m_id); }
    void reset() override { g_log.push_back("reset " + m_id); }
    bool build(remo::PointSource&, const remo::FrameContext&) override { return false; }
    void render(const remo::FrameContext&) override {}
    const remo::PipelineStats& stats() const override { return m_stats; }
    remo::TimingScopes timingScopes() const override { return {}; }

private:
    std::string m_id;
    bool m_failAllocate;
    remo::PipelineStats m_stats;
};

class FakeSource final : public remo::PointSource {
public:
    const remo::CloudMeta& meta() const override { return m_meta; }
    bool start(Mode mode, uint32_t ringSlots, size_t maxDeviceBytes, std::string*) override {
        g_log.push_back(std::string("source start ") + (mode == Mode::Whole ? "whole " : "stream ") +
                        std::to_string(ringSlots) + " " + std::to_string(maxDeviceBytes));
        return true;
    }
    void rewind() override { g_log.push_back("source rewind"); }
    void stop() override {}
    void pump() override {}
    remo::BatchView view() const override { return {}; }
    uint64_t numPointsUploaded() const override { return 0; }
    bool isFullyResident() const override { return false; }
    const remo_point* residentPoints() const override { return nullptr; }
    void setBatchesConsumed(uint32_t) override {}
    bool overranConsumer() const override { return false; }

private:
    remo::CloudMeta m_meta;
};

std::string quoted(const std::string& s) {
    std::string q = "\"";
    for (char ch : s) {
        if (ch == '"' || ch == '\\') q += '\\';
        q += ch;
    }
    return q + "\"";
}

}  // namespace

int main() {
    remo::PipelineRegistry reg;
    std::vector<std::string> results;
    auto add = [&](remo::PipelineFactory f) { results.push_back(reg.add(std::move(f)) ? "added" : "refused"); };
    add([] { return std::make_unique<FakePipeline>("stream", false); });
    add([] { return std::make_unique<FakePipeline>("whole", false); });
    add([] { return std::make_unique<FakePipeline>("stream", false); });  // duplicate id
    add([] { return std::make_unique<FakePipeline>("", false); });        // empty id
    add([] { return std::unique_ptr<remo::ILodPipeline>(); });            // nothing
    add(remo::PipelineFactory());                                         // no factory
    add([] { return std::make_unique<FakePipeline>("broken", true); });

    FakeSource src;
    remo::CloudMeta meta;
    meta.numPoints = 1234;
    remo::DeviceBudget budget;
    budget.bytes = 5678;
    std::vector<std::string> switches;
    for (const char* id : {"stream", "whole", "nope", "broken"}) {
        g_log.push_back(std::string("-- switch ") + id);
        std::string err;
        const bool ok = reg.switchTo(id, &src, meta, budget, &err);
        switches.push_back(ok ? "ok:" + reg.activeId() : "error:" + err + (reg.active() ? ":active" : ":none"));
    }
    g_log.push_back("-- release");
    reg.releaseActive();

    std::string out = "{\"add\": [";
    for (size_t i = 0; i < results.size(); ++i) out += (i ? ", " : "") + quoted(results[i]);
    out += "], \"list\": [";
    for (size_t i = 0; i < reg.list().size(); ++i) out += (i ? ", " : "") + quoted(reg.list()[i].id);
    out += "], \"find_whole\": " + std::string(reg.find("whole") ? "true" : "false");
    out += ", \"find_nope\": " + std::string(reg.find("nope") ? "true" : "false");
    out += ", \"switches\": [";
    for (size_t i = 0; i < switches.size(); ++i) out += (i ? ", " : "") + quoted(switches[i]);
    out += "], \"log\": [";
    for (size_t i = 0; i < g_log.size(); ++i) out += (i ? ", " : "") + quoted(g_log[i]);
    out += "]}";
    std::printf("%s\n", out.c_str());
    return 0;
}

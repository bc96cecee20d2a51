// The pipeline-plugin interface of the host side.
//
// Same types, method set and lifecycle as the reference's remo::ILodPipeline
// (include/remo/ILodPipeline.h:34-167): info -> initPrograms -> allocate -> reset,
// then per frame render(frame), source.pump(), build(source, frame); release or the
// destructor frees every device buffer. The GL surface of RenderTargets becomes a
// device pointer to the u64 depth|colour buffer, and the CUstream a hipStream_t
// passed as void* (the HIP path runs on its context's own stream).
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "remo/PointSource.h"
#include "remolod.h"

namespace remo {

using SharedUniforms = remo_shared_uniforms;

// include/remo/ILodPipeline.h:34-63.
struct PipelineInfo {
    std::string id;
    std::string displayName;
    bool progressive = false;
    bool needsWholeCloudResident = false;
    uint32_t ringSlots = 0;
    double bytesPerPointEstimate = 0.0;
    double minBytesPerPointEstimate = 0.0;
    uint64_t minStoreBytes = 0;
    uint64_t fixedBytesEstimate = 0;
};

// include/remo/ILodPipeline.h:65-69.
struct DeviceBudget {
    size_t vramTotal = 0;
    size_t vramFreeAtStartup = 0;
    size_t bytes = 0;
};

// include/remo/ILodPipeline.h:71-75: the output surface, here a device buffer of
// width * height u64 values (float_bits(depth) << 32 | rgba).
struct RenderTargets {
    uint64_t* framebuffer = nullptr;
    int width = 0;
    int height = 0;
};

// GpuProfiler's role (include/remo/GpuProfiler.h:62-158): per-scope totals.
struct ScopeTotal {
    std::string name;
    double ms = 0.0;
    uint64_t launches = 0;
};

// include/remo/ILodPipeline.h:77-86.
struct FrameContext {
    SharedUniforms uniforms{};
    RenderTargets targets;
    void* stream = nullptr;
    int numSMs = 0;
    bool strictTiming = false;
    bool profile = false;  // time every scope with HIP events
};

// include/remo/ILodPipeline.h:88-111.
struct PipelineStats {
    uint64_t numPoints = 0;
    uint64_t numVoxels = 0;
    uint64_t numPointsIngested = 0;
    uint32_t numNodes = 0;
    uint32_t numInner = 0;
    uint32_t numLeaves = 0;
    uint32_t maxDepth = 0;
    uint32_t maxPointsPerNode = 0;
    uint32_t numVisibleNodes = 0;
    uint64_t numVisiblePoints = 0;
    uint64_t numVisibleVoxels = 0;
    uint64_t bytesAllocated = 0;
    uint64_t bytesCapacity = 0;
    uint64_t samplesPerLevel[24] = {};
    bool memCapacityReached = false;
    bool nodeCapacityReached = false;
    bool allocOverflow = false;
};

// include/remo/ILodPipeline.h:113-116.
struct TimingScopes {
    std::vector<std::string> build;
    std::string render;
};

// include/remo/ILodPipeline.h:118-145.
class ILodPipeline {
public:
    virtual ~ILodPipeline() = default;
    virtual PipelineInfo info() const = 0;
    virtual bool initPrograms(std::string* err) = 0;
    virtual bool allocate(const CloudMeta& meta, const DeviceBudget& budget, std::string* err) = 0;
    virtual void release() = 0;
    virtual void reset() = 0;
    virtual bool build(PointSource& source, const FrameContext& frame) = 0;
    virtual void render(const FrameContext& frame) = 0;
    virtual const PipelineStats& stats() const = 0;
    virtual TimingScopes timingScopes() const = 0;
    virtual std::vector<std::string> diagnostics() const { return {}; }
    virtual std::vector<ScopeTotal> scopeTotals() const { return {}; }
};

using PipelineFactory = std::function<std::unique_ptr<ILodPipeline>()>;

}  // namespace remo

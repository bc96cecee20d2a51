// PointSource / CloudMeta / BatchView of the host side, over the C ABI in
// remolod.h.
//
// Same roles and method set as the reference's remo::PointSource
// (include/remo/PointSource.h:16-115): a producer that fills a device ring of
// batch slots behind a consumer, with the backpressure invariant
// "batches produced - batches consumed <= ring depth". Device pointers are plain
// `const void*` / typed pointers from the C ABI instead of CUdeviceptr.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "remo/SimlodReader.h"
#include "remolod.h"

namespace remo {

// include/remo/PointSource.h:16-32 (the fields the path reads).
struct CloudMeta {
    uint64_t numPoints = 0;
    double translation[3] = {0, 0, 0};
    float boxSize[3] = {0, 0, 0};

    remo_cloud_meta toAbi() const;
};

// include/remo/PointSource.h:41: points per ring slot.
inline constexpr uint32_t kSlotCapacity = REMO_DEFAULT_SLOT_CAPACITY;

// include/remo/PointSource.h:43-67. Batch B is in slot B % numSlots.
struct BatchView {
    const remo_point* slots = nullptr;
    const uint32_t* batchSizes = nullptr;
    const uint32_t* numBatchesUploaded = nullptr;
    uint32_t slotCapacity = kSlotCapacity;
    uint32_t numSlots = 0;
    uint32_t numBatchesTotal = 0;
    bool wrapping = false;

    remo_batch_view toAbi() const;
};

// include/remo/PointSource.h:69-115.
class PointSource {
public:
    enum class Mode { Stream, Whole };

    virtual ~PointSource() = default;
    virtual const CloudMeta& meta() const = 0;
    virtual bool start(Mode mode, uint32_t ringSlots, size_t maxDeviceBytes, std::string* err) = 0;
    virtual void rewind() = 0;
    virtual void stop() = 0;
    virtual void pump() = 0;
    virtual BatchView view() const = 0;
    virtual uint64_t numPointsUploaded() const = 0;
    virtual bool isFullyResident() const = 0;
    virtual const remo_point* residentPoints() const = 0;
    virtual void setBatchesConsumed(uint32_t numBatches) = 0;
    virtual bool overranConsumer() const = 0;
};

// The ring bookkeeping every source shares: the producer fills ring slots ahead
// of the consumer within the window "produced - consumed <= ring depth".
// Mode::Whole sizes the ring to hold every batch (nothing wraps); Mode::Stream
// uses `ringSlots` slots. A source supplies produce(): batch `batch` (cloud
// points [first, first + count)) into its ring slot.
class RingSource : public PointSource {
public:
    ~RingSource() override;

    const CloudMeta& meta() const override { return m_meta; }
    bool start(Mode mode, uint32_t ringSlots, size_t maxDeviceBytes, std::string* err) override;
    void rewind() override;
    void stop() override;
    void pump() override;
    BatchView view() const override;
    uint64_t numPointsUploaded() const override { return m_pointsUploaded; }
    bool isFullyResident() const override;
    const remo_point* residentPoints() const override;
    void setBatchesConsumed(uint32_t numBatches) override;
    bool overranConsumer() const override { return m_overran; }

    uint32_t numBatchesTotal() const { return m_numBatches; }
    uint32_t batchesProduced() const { return m_produced; }
    // Why the last pump stopped short (a failed read or copy); empty when it did not.
    const std::string& lastError() const { return m_error; }

protected:
    RingSource(const CloudMeta& meta, uint32_t slotCapacity);
    virtual bool produce(uint32_t batch, uint64_t first, uint32_t count, std::string* err) = 0;
    // After a pump that produced batches (ordering an asynchronous consumer).
    virtual void afterPump() {}

    CloudMeta m_meta;
    uint32_t m_slotCapacity;
    uint32_t m_numBatches = 0;
    uint32_t m_numSlots = 0;
    Mode m_mode = Mode::Stream;
    remo_ring* m_ring = nullptr;
    uint32_t m_produced = 0;
    uint32_t m_consumed = 0;
    uint64_t m_pointsUploaded = 0;
    bool m_overran = false;
    std::string m_error;
};

// A host-resident cloud streamed through a remo_ring.
// makeSyntheticSource's role (src/io/PointSource.cpp:535-584) is played by
// makeSceneSource, which fills the cloud with remo_scene_generate.
class HostCloudSource final : public RingSource {
public:
    HostCloudSource(std::vector<remo_point> points, const CloudMeta& meta,
                    uint32_t slotCapacity = kSlotCapacity);
    ~HostCloudSource() override;

    // Asynchronous ingest (no reference counterpart: its pump copies
    // synchronously): the cloud moves to pinned host memory, pump() queues each
    // batch on the ring's copy stream (remo_ring_push; a slot is refilled only
    // after `consumer` has finished the build step that read it), and orders
    // `consumer`'s next build step after the copies of the next `batchesPerStep`
    // batches only. Call before start(); false when pinned memory is refused.
    bool enableAsync(remo_ctx* consumer, uint32_t batchesPerStep);

private:
    bool produce(uint32_t batch, uint64_t first, uint32_t count, std::string* err) override;
    void afterPump() override;
    const remo_point* hostPoints() const { return m_pinned ? m_pinned : m_points.data(); }

    std::vector<remo_point> m_points;
    remo_point* m_pinned = nullptr;  // the cloud in pinned host memory (asynchronous ingest)
    remo_ctx* m_consumer = nullptr;
    uint32_t m_perStep = 1;
};

// A .simlod file streamed batch by batch: batch k is read from the file when the
// ring has room for it, so the host holds one slot of points, never the cloud
// (the reference's streaming .simlod producer, src/io/PointSource.cpp:34-55,
// :510-532). The copy is synchronous, like the reference's pump.
class SimlodFileSource final : public RingSource {
public:
    SimlodFileSource(std::string path, const SimlodInfo& info, const CloudMeta& meta,
                     uint32_t slotCapacity = kSlotCapacity);

    const std::string& path() const { return m_path; }

private:
    bool produce(uint32_t batch, uint64_t first, uint32_t count, std::string* err) override;

    std::string m_path;
    SimlodInfo m_info;
    std::vector<remo_point> m_staging;  // one slot
};

// Opens `path`: the record count from the header, the observed bounds, and the
// translation that puts the box minimum at 0 (deriveTranslation,
// src/io/PointSource.cpp:360-366: translation = -min, box size = float(max +
// translation)). nullptr and a message on failure.
std::unique_ptr<SimlodFileSource> makeSimlodSource(const std::string& path, uint32_t slotCapacity,
                                                   std::string* err);

// Seed-generated lidar-like scene (remo_scene_generate), `numPoints` points.
std::unique_ptr<HostCloudSource> makeSceneSource(uint64_t numPoints, uint64_t seed,
                                                 const remo_scene_spec* spec = nullptr,
                                                 uint32_t slotCapacity = kSlotCapacity);

}  // namespace remo

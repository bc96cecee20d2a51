// Point sources over a device ring (the reference's PointSource contract,
// include/remo/PointSource.h:69-115): a host-resident cloud and a .simlod file.
#include <algorithm>
#include <cstdio>

#include "remo/PointSource.h"

namespace remo {

remo_cloud_meta CloudMeta::toAbi() const {
    remo_cloud_meta m{};
    m.num_points = numPoints;
    for (int i = 0; i < 3; ++i) {
        m.translation[i] = translation[i];
        m.box_size[i] = boxSize[i];
    }
    return m;
}

remo_batch_view BatchView::toAbi() const {
    remo_batch_view v{};
    v.slots = slots;
    v.batch_sizes = batchSizes;
    v.num_batches_uploaded = numBatchesUploaded;
    v.slot_capacity = slotCapacity;
    v.num_slots = numSlots;
    v.num_batches_total = numBatchesTotal;
    return v;
}

RingSource::RingSource(const CloudMeta& meta, uint32_t slotCapacity)
    : m_meta(meta), m_slotCapacity(slotCapacity) {
    m_numBatches = uint32_t((m_meta.numPoints + slotCapacity - 1) / slotCapacity);
}

RingSource::~RingSource() { RingSource::stop(); }

bool RingSource::start(Mode mode, uint32_t ringSlots, size_t maxDeviceBytes, std::string* err) {
    stop();
    m_mode = mode;
    m_numSlots = mode == Mode::Whole ? std::max(1u, m_numBatches) : std::max(1u, ringSlots);
    const uint64_t bytes = uint64_t(m_numSlots) * m_slotCapacity * sizeof(remo_point);
    if (maxDeviceBytes && bytes > maxDeviceBytes) {
        if (err) *err = "the point ring needs " + std::to_string(bytes >> 20) + " MB, over the budget";
        return false;
    }
    const int rc = remo_ring_create(m_slotCapacity, m_numSlots, &m_ring);
    if (rc != REMO_OK) {
        if (err) *err = "remo_ring_create failed (" + std::to_string(rc) + ")";
        m_ring = nullptr;
        return false;
    }
    rewind();
    if (mode == Mode::Whole) {
        pump();
        if (!m_error.empty()) {
            if (err) *err = m_error;
            return false;
        }
    }
    return true;
}

void RingSource::rewind() {
    if (m_ring) remo_ring_rewind(m_ring);  // waits for queued copies first
    m_produced = 0;
    m_consumed = 0;
    m_pointsUploaded = 0;
    m_error.clear();
}

void RingSource::stop() {
    if (m_ring) {
        remo_ring_sync(m_ring);
        remo_ring_destroy(m_ring);
    }
    m_ring = nullptr;
}

void RingSource::pump() {
    // Fill every slot the consumer window allows: producer - consumer <= depth.
    if (!m_ring) return;
    const uint32_t limit = std::min(m_numBatches, m_consumed + m_numSlots);
    while (m_produced < limit) {
        const uint64_t first = uint64_t(m_produced) * m_slotCapacity;
        const uint32_t n = uint32_t(std::min<uint64_t>(m_slotCapacity, m_meta.numPoints - first));
        if (!produce(m_produced, first, n, &m_error)) {
            std::fprintf(stderr, "remobench: batch %u: %s\n", m_produced, m_error.c_str());
            return;
        }
        m_pointsUploaded += n;
        ++m_produced;
    }
    afterPump();
}

BatchView RingSource::view() const {
    BatchView v;
    if (!m_ring) return v;
    remo_batch_view a{};
    if (remo_ring_view(m_ring, m_numBatches, &a) != REMO_OK) return v;
    v.slots = a.slots;
    v.batchSizes = a.batch_sizes;
    v.numBatchesUploaded = a.num_batches_uploaded;
    v.slotCapacity = a.slot_capacity;
    v.numSlots = a.num_slots;
    v.numBatchesTotal = a.num_batches_total;
    v.wrapping = a.num_batches_total > a.num_slots;
    return v;
}

bool RingSource::isFullyResident() const { return m_ring && m_numSlots >= m_numBatches && m_produced == m_numBatches; }

const remo_point* RingSource::residentPoints() const {
    if (!isFullyResident()) return nullptr;
    return view().slots;
}

void RingSource::setBatchesConsumed(uint32_t numBatches) {
    // The producer never runs ahead of the window, so an overrun here means the
    // consumer reported going backwards without a rewind.
    if (numBatches < m_consumed) m_overran = true;
    m_consumed = numBatches;
}

// ---- a host-resident cloud

HostCloudSource::HostCloudSource(std::vector<remo_point> points, const CloudMeta& meta, uint32_t slotCapacity)
    : RingSource([&] {
          CloudMeta m = meta;
          m.numPoints = points.size();
          return m;
      }(), slotCapacity),
      m_points(std::move(points)) {}

HostCloudSource::~HostCloudSource() {
    stop();  // no copy may still read the pinned cloud
    if (m_pinned) remo_host_free(m_pinned);
    m_pinned = nullptr;
}

bool HostCloudSource::enableAsync(remo_ctx* consumer, uint32_t batchesPerStep) {
    if (!consumer) return false;
    if (!m_pinned && !m_points.empty()) {
        void* p = nullptr;
        if (remo_host_alloc(m_points.size() * sizeof(remo_point), &p) != REMO_OK) return false;
        m_pinned = static_cast<remo_point*>(p);
        std::copy(m_points.begin(), m_points.end(), m_pinned);
        std::vector<remo_point>().swap(m_points);  // one host copy of the cloud
    }
    m_consumer = consumer;
    m_perStep = std::max(1u, batchesPerStep);
    return true;
}

bool HostCloudSource::produce(uint32_t batch, uint64_t first, uint32_t count, std::string* err) {
    const int rc = m_consumer ? remo_ring_push(m_ring, batch, hostPoints() + first, count, m_consumer)
                              : remo_ring_upload(m_ring, batch, hostPoints() + first, count);
    if (rc != REMO_OK && err) *err = "ring upload failed (" + std::to_string(rc) + ")";
    return rc == REMO_OK;
}

void HostCloudSource::afterPump() {
    // the consumer's next step reads at most m_perStep batches from m_consumed on
    if (m_consumer && m_produced > m_consumed)
        remo_ring_order_before(m_ring, m_consumer, m_consumed, std::min(m_perStep, m_produced - m_consumed));
}

// ---- a .simlod file

SimlodFileSource::SimlodFileSource(std::string path, const SimlodInfo& info, const CloudMeta& meta,
                                   uint32_t slotCapacity)
    : RingSource(meta, slotCapacity), m_path(std::move(path)), m_info(info) {
    m_staging.resize(std::min<uint64_t>(slotCapacity, info.numPoints));
}

bool SimlodFileSource::produce(uint32_t batch, uint64_t first, uint32_t count, std::string* err) {
    if (!readSimlodRange(m_path, m_info, m_meta.translation, first, count, m_staging.data(), err)) return false;
    const int rc = remo_ring_upload(m_ring, batch, m_staging.data(), count);  // synchronous: the slot is reusable
    if (rc != REMO_OK && err) *err = "ring upload failed (" + std::to_string(rc) + ")";
    return rc == REMO_OK;
}

std::unique_ptr<SimlodFileSource> makeSimlodSource(const std::string& path, uint32_t slotCapacity,
                                                   std::string* err) {
    SimlodInfo info;
    if (!readSimlodHeader(path, info, err)) return nullptr;
    double bounds[6];
    if (!readSimlodBounds(path, info, bounds, err)) return nullptr;
    CloudMeta meta;
    meta.numPoints = info.numPoints;
    for (int k = 0; k < 3; ++k) {
        meta.translation[k] = -bounds[k];
        meta.boxSize[k] = float(bounds[3 + k] + meta.translation[k]);
    }
    return std::make_unique<SimlodFileSource>(path, info, meta, slotCapacity);
}

std::unique_ptr<HostCloudSource> makeSceneSource(uint64_t numPoints, uint64_t seed, const remo_scene_spec* spec,
                                                 uint32_t slotCapacity) {
    remo_scene_spec s{};
    if (spec)
        s = *spec;
    else
        remo_scene_default(&s);
    std::vector<remo_point> pts(numPoints);
    remo_cloud_meta m{};
    if (remo_scene_generate(&s, seed, numPoints, pts.data(), &m) != REMO_OK) return nullptr;
    CloudMeta meta;
    meta.numPoints = m.num_points;
    for (int i = 0; i < 3; ++i) {
        meta.translation[i] = m.translation[i];
        meta.boxSize[i] = m.box_size[i];
    }
    return std::make_unique<HostCloudSource>(std::move(pts), meta, slotCapacity);
}

}  // namespace remo

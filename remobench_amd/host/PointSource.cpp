// HostCloudSource: a host-resident cloud streamed through a device ring
// (the reference's PointSource contract, include/remo/PointSource.h:69-115).
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

HostCloudSource::HostCloudSource(std::vector<remo_point> points, const CloudMeta& meta, uint32_t slotCapacity)
    : m_points(std::move(points)), m_meta(meta), m_slotCapacity(slotCapacity) {
    m_meta.numPoints = m_points.size();
    m_hostCount = m_points.size();
    m_numBatches = uint32_t((m_points.size() + slotCapacity - 1) / slotCapacity);
}

HostCloudSource::~HostCloudSource() {
    stop();
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

bool HostCloudSource::start(Mode mode, uint32_t ringSlots, size_t maxDeviceBytes, std::string* err) {
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
    if (mode == Mode::Whole) pump();
    return true;
}

void HostCloudSource::rewind() {
    if (m_ring) remo_ring_rewind(m_ring);  // waits for queued copies first
    m_produced = 0;
    m_consumed = 0;
    m_pointsUploaded = 0;
}

void HostCloudSource::stop() {
    if (m_ring) {
        remo_ring_sync(m_ring);
        remo_ring_destroy(m_ring);
    }
    m_ring = nullptr;
}

void HostCloudSource::pump() {
    // Fill every slot the consumer window allows: producer - consumer <= depth.
    if (!m_ring) return;
    const uint32_t limit = std::min(m_numBatches, m_consumed + m_numSlots);
    while (m_produced < limit) {
        const uint64_t first = uint64_t(m_produced) * m_slotCapacity;
        const uint32_t n = uint32_t(std::min<uint64_t>(m_slotCapacity, m_hostCount - first));
        const int rc = m_consumer ? remo_ring_push(m_ring, m_produced, hostPoints() + first, n, m_consumer)
                                  : remo_ring_upload(m_ring, m_produced, hostPoints() + first, n);
        if (rc != REMO_OK) {
            std::fprintf(stderr, "remobench: ring upload of batch %u failed\n", m_produced);
            return;
        }
        m_pointsUploaded += n;
        ++m_produced;
    }
    // the consumer's next step reads at most m_perStep batches from m_consumed on
    if (m_consumer && m_produced > m_consumed)
        remo_ring_order_before(m_ring, m_consumer, m_consumed, std::min(m_perStep, m_produced - m_consumed));
}

BatchView HostCloudSource::view() const {
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

bool HostCloudSource::isFullyResident() const { return m_ring && m_numSlots >= m_numBatches && m_produced == m_numBatches; }

const remo_point* HostCloudSource::residentPoints() const {
    if (!isFullyResident()) return nullptr;
    return view().slots;
}

void HostCloudSource::setBatchesConsumed(uint32_t numBatches) {
    // The producer never runs ahead of the window, so an overrun here means the
    // consumer reported going backwards without a rewind.
    if (numBatches < m_consumed) m_overran = true;
    m_consumed = numBatches;
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

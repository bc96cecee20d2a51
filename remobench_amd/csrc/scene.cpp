// Seed-generated lidar-like scenes: the synthetic source of the path.
//
// Plays the role of makeSyntheticSource (src/io/PointSource.cpp:535-584) for the
// scenes this path is measured on: a ground plane with Gaussian height noise,
// axis-aligned boxes (5 faces, no bottom), vertical cylinders (lateral surface)
// and optional Gaussian clusters. Points are shared between surfaces in
// proportion to area, ordered along a simulated sensor sweep (scan line =
// floor(y / pitch), then x), shuffled within fixed windows, and translated so the
// box minimum is 0 exactly as deriveTranslation/applyTranslation do
// (src/io/PointSource.cpp:360-379).
//
// Every random draw comes from SplitMix64 keyed by (seed, stream, index), and
// only IEEE-exact operations (+ - * / sqrt) touch the values, so a scene is the
// same bytes on every host and for every thread count.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/remolod.h"

namespace {

struct SplitMix {
    uint64_t s;
    SplitMix(uint64_t a, uint64_t b, uint64_t c) {
        s = a * 0x9E3779B97F4A7C15ull ^ (b + 0x632BE59BD9B4E019ull) * 0xD1B54A32D192ED03ull ^
            (c + 0x8CB92BA72F3D8DD7ull) * 0xC2B2AE3D27D4EB4Full;
        next();
    }
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // [0, 1) with 53 random bits.
    double uni() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
    // Irwin-Hall(12) - 6: zero mean, unit variance, built from exact adds only.
    double normal() {
        double acc = 0.0;
        for (int i = 0; i < 12; ++i) acc += uni();
        return acc - 6.0;
    }
    // Integer in [lo, hi].
    int range(int lo, int hi) { return lo + int(next() % uint64_t(hi - lo + 1)); }
};

enum Stream : uint64_t { kPrims = 1, kPoint = 2, kShuffle = 3 };

struct Box {
    double x0, y0, sx, sy, sz;
    uint32_t rgb[3];
    double faceArea[5];  // -x, +x, -y, +y, top
    double area;
};
struct Cylinder {
    double cx, cy;
    uint32_t rgb[3];
};
struct Cluster {
    double c[3];
    uint32_t rgb[3];
};

constexpr double kCylR = 0.15;
constexpr double kCylH = 6.0;
constexpr double kPi = 3.14159265358979323846;

uint32_t packColor(SplitMix& r, const uint32_t base[3]) {
    uint32_t out = 255u << 24;
    for (int ch = 0; ch < 3; ++ch) {
        int v = int(base[ch]) + r.range(-10, 10);
        v = v < 0 ? 0 : (v > 255 ? 255 : v);
        out |= uint32_t(v) << (8 * ch);
    }
    return out;
}

template <typename F>
void parallelFor(uint64_t n, unsigned threads, F&& f) {
    if (threads <= 1 || n < 65536) {
        f(0, n);
        return;
    }
    std::vector<std::thread> pool;
    const uint64_t per = (n + threads - 1) / threads;
    for (unsigned t = 0; t < threads; ++t) {
        const uint64_t b = uint64_t(t) * per, e = std::min<uint64_t>(n, b + per);
        if (b >= e) break;
        pool.emplace_back([&f, b, e] { f(b, e); });
    }
    for (auto& th : pool) th.join();
}

// Stable LSD radix sort of (key, index) pairs by key, 8 bits a pass, skipping
// passes whose byte is constant over all keys.
void radixSort(std::vector<uint64_t>& keys, std::vector<uint32_t>& idx, unsigned threads) {
    const uint64_t n = keys.size();
    std::vector<uint64_t> k2(n);
    std::vector<uint32_t> i2(n);
    uint64_t orAll = 0, andAll = ~0ull;
    for (uint64_t k : keys) {
        orAll |= k;
        andAll &= k;
    }
    const unsigned T = std::max(1u, threads);
    const uint64_t per = (n + T - 1) / T;
    for (int pass = 0; pass < 8; ++pass) {
        const int sh = pass * 8;
        if (((orAll ^ andAll) >> sh & 0xFF) == 0) continue;
        std::vector<uint64_t> hist(size_t(T) * 256, 0);
        parallelFor(T, T, [&](uint64_t tb, uint64_t te) {
            for (uint64_t t = tb; t < te; ++t) {
                const uint64_t b = t * per, e = std::min(n, b + per);
                uint64_t* h = &hist[t * 256];
                for (uint64_t i = b; i < e; ++i) h[(keys[i] >> sh) & 0xFF]++;
            }
        });
        uint64_t sum = 0;
        for (int d = 0; d < 256; ++d)
            for (unsigned t = 0; t < T; ++t) {
                const uint64_t c = hist[size_t(t) * 256 + d];
                hist[size_t(t) * 256 + d] = sum;
                sum += c;
            }
        parallelFor(T, T, [&](uint64_t tb, uint64_t te) {
            for (uint64_t t = tb; t < te; ++t) {
                const uint64_t b = t * per, e = std::min(n, b + per);
                uint64_t* h = &hist[t * 256];
                for (uint64_t i = b; i < e; ++i) {
                    const uint64_t pos = h[(keys[i] >> sh) & 0xFF]++;
                    k2[pos] = keys[i];
                    i2[pos] = idx[i];
                }
            }
        });
        keys.swap(k2);
        idx.swap(i2);
    }
}

}  // namespace

extern "C" void remo_scene_default(remo_scene_spec* s) {
    std::memset(s, 0, sizeof(*s));
    s->extent = 40.0;
    s->num_boxes = 20;
    s->num_cylinders = 50;
    s->cluster_fraction = 0.0;
    s->cluster_volume_fraction = 0.05;
    s->cluster_sigma = 0.5;
    s->ground_sigma = 0.02;
    s->scanline = 0.05;
    s->shuffle_window = 4096;
    s->threads = 0;
}

extern "C" int remo_scene_generate(const remo_scene_spec* spec, uint64_t seed,
                                   uint64_t numPoints, remo_point* out,
                                   remo_cloud_meta* meta) {
    if (!spec || (!out && numPoints) || !meta) return REMO_ERR_INVALID;
    if (numPoints > 0xFFFFFFFFull) return REMO_ERR_INVALID;
    if (!(spec->extent > 2 * kCylR + 10.0) || spec->scanline <= 0.0) return REMO_ERR_INVALID;
    const double E = spec->extent;
    unsigned threads = spec->threads ? spec->threads : std::thread::hardware_concurrency();
    threads = std::max(1u, std::min(threads, 64u));

    // ---- primitives ----
    SplitMix pr(seed, kPrims, 0);
    uint32_t groundRgb[3] = {uint32_t(pr.range(90, 130)), uint32_t(pr.range(90, 120)),
                             uint32_t(pr.range(60, 90))};
    std::vector<Box> boxes(spec->num_boxes);
    for (Box& b : boxes) {
        b.sx = 2.0 + 8.0 * pr.uni();
        b.sy = 2.0 + 8.0 * pr.uni();
        b.sz = 2.0 + 8.0 * pr.uni();
        b.x0 = (E - b.sx) * pr.uni();
        b.y0 = (E - b.sy) * pr.uni();
        for (uint32_t& c : b.rgb) c = uint32_t(pr.range(40, 220));
        b.faceArea[0] = b.faceArea[1] = b.sy * b.sz;
        b.faceArea[2] = b.faceArea[3] = b.sx * b.sz;
        b.faceArea[4] = b.sx * b.sy;
        b.area = 0;
        for (double a : b.faceArea) b.area += a;
    }
    std::vector<Cylinder> cyls(spec->num_cylinders);
    for (Cylinder& c : cyls) {
        c.cx = kCylR + (E - 2 * kCylR) * pr.uni();
        c.cy = kCylR + (E - 2 * kCylR) * pr.uni();
        uint32_t g = uint32_t(pr.range(120, 200));
        c.rgb[0] = g;
        c.rgb[1] = g;
        c.rgb[2] = uint32_t(pr.range(120, 200));
    }
    std::vector<Cluster> clusters;
    const double cf = std::min(1.0, std::max(0.0, spec->cluster_fraction));
    if (cf > 0.0) {
        // Enough spheres of radius 2 sigma to cover the requested share of the
        // scene volume E x E x 10 m.
        const double r = 2.0 * spec->cluster_sigma;
        const double vol = E * E * 10.0 * spec->cluster_volume_fraction;
        const size_t k = std::max<size_t>(1, size_t(std::ceil(vol / (4.0 / 3.0 * kPi * r * r * r))));
        clusters.resize(k);
        for (Cluster& c : clusters) {
            c.c[0] = E * pr.uni();
            c.c[1] = E * pr.uni();
            c.c[2] = 10.0 * pr.uni();
            for (uint32_t& v : c.rgb) v = uint32_t(pr.range(60, 250));
        }
    }

    // ---- point budget per primitive: clusters first, surfaces by area ----
    const uint64_t nCluster = clusters.empty() ? 0 : uint64_t(std::floor(double(numPoints) * cf));
    const uint64_t nSurface = numPoints - nCluster;
    const size_t nPrim = 1 + boxes.size() + cyls.size();
    std::vector<double> area(nPrim);
    area[0] = E * E;
    for (size_t i = 0; i < boxes.size(); ++i) area[1 + i] = boxes[i].area;
    for (size_t i = 0; i < cyls.size(); ++i) area[1 + boxes.size() + i] = 2.0 * kPi * kCylR * kCylH;
    double total = 0;
    for (double a : area) total += a;
    std::vector<uint64_t> cnt(nPrim);
    std::vector<std::pair<double, size_t>> rem(nPrim);
    uint64_t given = 0;
    for (size_t i = 0; i < nPrim; ++i) {
        const double exact = double(nSurface) * area[i] / total;
        cnt[i] = uint64_t(std::floor(exact));
        given += cnt[i];
        rem[i] = {exact - double(cnt[i]), i};
    }
    std::stable_sort(rem.begin(), rem.end(),
                     [](const auto& a, const auto& b) { return a.first > b.first; });
    for (size_t i = 0; given < nSurface; ++i, ++given) cnt[rem[i % nPrim].second]++;
    // prefix over [surfaces..., clusters...]
    const size_t nAll = nPrim + clusters.size();
    std::vector<uint64_t> start(nAll + 1, 0);
    for (size_t i = 0; i < nPrim; ++i) start[i + 1] = start[i] + cnt[i];
    for (size_t i = 0; i < clusters.size(); ++i) {
        const uint64_t c = nCluster / clusters.size() + (i < nCluster % clusters.size() ? 1 : 0);
        start[nPrim + i + 1] = start[nPrim + i] + c;
    }

    // ---- points ----
    std::vector<remo_point> raw(numPoints);
    parallelFor(numPoints, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) {
            const size_t prim = size_t(std::upper_bound(start.begin(), start.end(), i) - start.begin()) - 1;
            SplitMix r(seed, kPoint, i);
            double x, y, z;
            uint32_t color;
            if (prim == 0) {
                x = E * r.uni();
                y = E * r.uni();
                z = spec->ground_sigma * r.normal();
                color = packColor(r, groundRgb);
            } else if (prim < 1 + boxes.size()) {
                const Box& bx = boxes[prim - 1];
                double pick = bx.area * r.uni();
                int face = 0;
                while (face < 4 && pick >= bx.faceArea[face]) pick -= bx.faceArea[face++];
                const double u = r.uni(), v = r.uni();
                switch (face) {
                    case 0: x = bx.x0; y = bx.y0 + u * bx.sy; z = v * bx.sz; break;
                    case 1: x = bx.x0 + bx.sx; y = bx.y0 + u * bx.sy; z = v * bx.sz; break;
                    case 2: x = bx.x0 + u * bx.sx; y = bx.y0; z = v * bx.sz; break;
                    case 3: x = bx.x0 + u * bx.sx; y = bx.y0 + bx.sy; z = v * bx.sz; break;
                    default: x = bx.x0 + u * bx.sx; y = bx.y0 + v * bx.sy; z = bx.sz; break;
                }
                color = packColor(r, bx.rgb);
            } else if (prim < nPrim) {
                const Cylinder& c = cyls[prim - 1 - boxes.size()];
                double u, v, rr;
                do {  // uniform direction by rejection; only exact operations
                    u = 2.0 * r.uni() - 1.0;
                    v = 2.0 * r.uni() - 1.0;
                    rr = u * u + v * v;
                } while (rr > 1.0 || rr < 0.0625);
                const double inv = 1.0 / std::sqrt(rr);
                x = c.cx + kCylR * u * inv;
                y = c.cy + kCylR * v * inv;
                z = kCylH * r.uni();
                color = packColor(r, c.rgb);
            } else {
                const Cluster& c = clusters[prim - nPrim];
                x = c.c[0] + spec->cluster_sigma * r.normal();
                y = c.c[1] + spec->cluster_sigma * r.normal();
                z = c.c[2] + spec->cluster_sigma * r.normal();
                color = packColor(r, c.rgb);
            }
            raw[i] = {float(x), float(y), float(z), color};
        }
    });

    // ---- bounds and translation (deriveTranslation / applyTranslation) ----
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (const remo_point& p : raw) {
        const float c[3] = {p.x, p.y, p.z};
        for (int a = 0; a < 3; ++a) {
            mn[a] = std::min(mn[a], double(c[a]));
            mx[a] = std::max(mx[a], double(c[a]));
        }
    }
    if (numPoints == 0) {
        for (int a = 0; a < 3; ++a) mn[a] = mx[a] = 0.0;
    }
    meta->num_points = numPoints;
    float t[3];
    for (int a = 0; a < 3; ++a) {
        meta->translation[a] = -mn[a];
        meta->box_size[a] = float(mx[a] + meta->translation[a]);
        t[a] = float(meta->translation[a]);
    }
    meta->pad0 = 0;
    const bool shift = !(t[0] == 0.0f && t[1] == 0.0f && t[2] == 0.0f);
    if (shift) {
        parallelFor(numPoints, threads, [&](uint64_t b, uint64_t e) {
            for (uint64_t i = b; i < e; ++i) {
                raw[i].x += t[0];
                raw[i].y += t[1];
                raw[i].z += t[2];
            }
        });
    }

    // ---- scan-line order: (floor(y / pitch), x), stable ----
    std::vector<uint64_t> keys(numPoints);
    std::vector<uint32_t> idx(numPoints);
    parallelFor(numPoints, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) {
            const double line = std::floor(double(raw[i].y) / spec->scanline);
            uint32_t xb;
            std::memcpy(&xb, &raw[i].x, 4);  // x >= 0: float bits order = value order
            keys[i] = (uint64_t(line) << 32) | xb;
            idx[i] = uint32_t(i);
        }
    });
    radixSort(keys, idx, threads);
    parallelFor(numPoints, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) out[i] = raw[idx[i]];
    });

    // ---- shuffle within windows ----
    const uint64_t W = spec->shuffle_window;
    if (W > 1) {
        const uint64_t nw = (numPoints + W - 1) / W;
        parallelFor(nw, threads, [&](uint64_t b, uint64_t e) {
            for (uint64_t w = b; w < e; ++w) {
                const uint64_t lo = w * W, n = std::min(W, numPoints - lo);
                SplitMix r(seed, kShuffle, w);
                for (uint64_t i = n - 1; i > 0; --i) {
                    const uint64_t j = r.next() % (i + 1);
                    std::swap(out[lo + i], out[lo + j]);
                }
            }
        });
    }
    return REMO_OK;
}

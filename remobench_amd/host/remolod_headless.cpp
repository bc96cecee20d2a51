// remolod_headless: the reference shell's frame loop without a window
// (src/shell/App.cpp:420-450): per frame render(frame), source.pump(),
// build(source, frame), until build returns false. The camera is
// OrbitControls::frameBox's overview (src/shell/OrbitControls.h:68-93) with the
// shell's projection (fovy 60 degrees, near = extent * 1e-4, far = extent * 100;
// src/shell/App.cpp:200-248). Prints one JSON object of stats and timings.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "RemolodHipPipeline.h"
#include "remo/PipelineRegistry.h"

namespace {

struct M4 {
    double m[4][4];
};

M4 identity() {
    M4 r{};
    for (int i = 0; i < 4; ++i) r.m[i][i] = 1.0;
    return r;
}

M4 mul(const M4& a, const M4& b) {
    M4 r{};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            for (int k = 0; k < 4; ++k) r.m[i][j] += a.m[i][k] * b.m[k][j];
    return r;
}

M4 translate(double x, double y, double z) {
    M4 r = identity();
    r.m[0][3] = x;
    r.m[1][3] = y;
    r.m[2][3] = z;
    return r;
}

// Rotation about a unit axis (glm::rotate's matrix).
M4 rotate(double a, double x, double y, double z) {
    const double c = std::cos(a), s = std::sin(a), t = 1.0 - c;
    M4 r = identity();
    r.m[0][0] = c + x * x * t;     r.m[0][1] = x * y * t - z * s; r.m[0][2] = x * z * t + y * s;
    r.m[1][0] = y * x * t + z * s; r.m[1][1] = c + y * y * t;     r.m[1][2] = y * z * t - x * s;
    r.m[2][0] = z * x * t - y * s; r.m[2][1] = z * y * t + x * s; r.m[2][2] = c + z * z * t;
    return r;
}

// Inverse of a rigid transform [R t; 0 1].
M4 rigidInverse(const M4& a) {
    M4 r = identity();
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
    for (int i = 0; i < 3; ++i) r.m[i][3] = -(r.m[i][0] * a.m[0][3] + r.m[i][1] * a.m[1][3] + r.m[i][2] * a.m[2][3]);
    return r;
}

// General 4x4 inverse by Gauss-Jordan with partial pivoting.
M4 inverse(const M4& a) {
    double w[4][8];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 8; ++j) w[i][j] = j < 4 ? a.m[i][j] : (j - 4 == i ? 1.0 : 0.0);
    for (int c = 0; c < 4; ++c) {
        int p = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(w[r][c]) > std::fabs(w[p][c])) p = r;
        for (int j = 0; j < 8; ++j) std::swap(w[c][j], w[p][j]);
        const double d = w[c][c];
        for (int j = 0; j < 8; ++j) w[c][j] /= d;
        for (int r = 0; r < 4; ++r)
            if (r != c) {
                const double f = w[r][c];
                for (int j = 0; j < 8; ++j) w[r][j] -= f * w[c][j];
            }
    }
    M4 r{};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.m[i][j] = w[i][j + 4];
    return r;
}

void store(remo_mat4& dst, const M4& m) {
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) dst.rows[r][c] = float(m.m[r][c]);
}

remo_shared_uniforms overviewUniforms(const float box[3], int width, int height) {
    const double kPi = 3.14159265358979323846;
    const double fovy = 60.0 * kPi / 180.0;
    double bx = box[0], by = box[1], bz = box[2];
    const double diag = std::sqrt(bx * bx + by * by + bz * bz);
    const double radius = (diag * 0.5) / std::tan(fovy * 0.5) * 1.1;
    M4 flip{};
    flip.m[0][0] = 1; flip.m[1][2] = -1; flip.m[2][1] = 1; flip.m[3][3] = 1;
    const M4 world = mul(mul(mul(mul(translate(bx * 0.5, by * 0.5, bz * 0.5), rotate(0.35, 0, 0, 1)),
                                 rotate(-0.6, 1, 0, 0)), flip), translate(0, 0, radius));
    const M4 view = rigidInverse(world);
    const double extent = std::max(std::max(bx, by), std::max(bz, 1.0));
    const double n = extent * 1e-4, f = extent * 100.0, t = 1.0 / std::tan(fovy / 2.0);
    M4 proj{};
    proj.m[0][0] = t / (double(width) / double(height));
    proj.m[1][1] = t;
    proj.m[2][2] = -(f + n) / (f - n);
    proj.m[2][3] = -(2.0 * f * n) / (f - n);
    proj.m[3][2] = -1.0;
    const M4 transform = mul(proj, view);
    remo_shared_uniforms u{};
    u.width = float(width);
    u.height = float(height);
    u.fovy_rad = float(fovy);
    store(u.view, view);
    store(u.proj, proj);
    store(u.transform, transform);
    store(u.transform_frozen, transform);
    store(u.transform_frozen_inv, inverse(transform));
    u.box_max = {box[0], box[1], box[2]};
    u.lod_pixel_budget = 128.0f;
    u.min_node_size = 64.0f;
    u.lod_scale = 0.5f;
    u.point_size = 1;
    u.edl_strength = 0.4f;
    u.do_update_visibility = 1;
    u.enable_edl = 1;
    u.show_points = 1;
    return u;
}

void usage() {
    std::fprintf(stderr,
                 "usage: remolod_headless [--points N] [--seed S] [--slot N] [--batches-per-step N]\n"
                 "                        [--budget-ms MS] [--analysis none|analytic] [--width W --height H]\n"
                 "                        [--white] [--no-accum] [--profile] [--device-budget-mb MB]\n"
                 "                        [--dump-uniforms FILE] [--dump-fb FILE]\n"
                 "                        [--remolod-fixed-depth D] [--async-upload] [--simlod FILE]\n"
                 "                        [--refine-budget-ms MS] (Refinement each frame; then one\n"
                 "                        unbudgeted pass before the final frame)\n");
}

}  // namespace

int main(int argc, char** argv) {
    uint64_t points = 2'000'000, seed = 1;
    uint32_t slot = remo::kSlotCapacity, perStep = 20;
    float budgetMs = 10.0f;
    int width = 1920, height = 1080;
    bool white = false, accumulate = true, profile = false;
    std::string analysis = "analytic", dumpUniforms, dumpFb, simlodPath;
    float refineBudgetMs = -2.0f;  // below -1: no Refinement
    uint32_t fixedDepth = 0;
    bool asyncUpload = false;
    uint64_t deviceBudgetMb = 0;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> const char* {
            if (i + 1 >= argc) {
                usage();
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--points") points = std::strtoull(next(), nullptr, 10);
        else if (a == "--seed") seed = std::strtoull(next(), nullptr, 10);
        else if (a == "--slot") slot = uint32_t(std::strtoul(next(), nullptr, 10));
        else if (a == "--batches-per-step") perStep = uint32_t(std::strtoul(next(), nullptr, 10));
        else if (a == "--budget-ms") budgetMs = std::strtof(next(), nullptr);
        else if (a == "--analysis") analysis = next();
        else if (a == "--width") width = std::atoi(next());
        else if (a == "--height") height = std::atoi(next());
        else if (a == "--white") white = true;
        else if (a == "--no-accum") accumulate = false;
        else if (a == "--profile") profile = true;
        else if (a == "--device-budget-mb") deviceBudgetMb = std::strtoull(next(), nullptr, 10);
        else if (a == "--dump-uniforms") dumpUniforms = next();
        else if (a == "--dump-fb") dumpFb = next();
        else if (a == "--refine-budget-ms") refineBudgetMs = std::strtof(next(), nullptr);
        else if (a == "--remolod-fixed-depth") fixedDepth = uint32_t(std::strtoul(next(), nullptr, 10));
        else if (a == "--async-upload") asyncUpload = true;
        else if (a == "--simlod") simlodPath = next();
        else if (a == "--help" || a == "-h") {
            usage();
            return 0;
        } else {
            std::fprintf(stderr, "unknown argument %s\n", a.c_str());
            usage();
            return 2;
        }
    }
    if (analysis != "none" && analysis != "analytic") {
        std::fprintf(stderr, "--analysis must be none or analytic\n");
        return 2;
    }

    int devices = 0;
    if (remo_device_count(&devices) != REMO_OK || devices < 1) {
        std::fprintf(stderr, "remolod_headless: no HIP device visible\n");
        return 1;
    }
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    // the cloud: a .simlod file read batch by batch, or the seed-generated scene
    std::unique_ptr<remo::RingSource> source;
    remo::HostCloudSource* hostCloud = nullptr;
    if (!simlodPath.empty()) {
        if (asyncUpload) {
            std::fprintf(stderr, "--async-upload needs the cloud in host memory, not --simlod\n");
            return 2;
        }
        std::string openErr;
        source = remo::makeSimlodSource(simlodPath, slot, &openErr);
        if (!source) {
            std::fprintf(stderr, "remolod_headless: %s\n", openErr.c_str());
            return 1;
        }
    } else {
        auto scene = remo::makeSceneSource(points, seed, nullptr, slot);
        if (!scene) {
            std::fprintf(stderr, "scene generation failed\n");
            return 1;
        }
        hostCloud = scene.get();
        source = std::move(scene);
    }
    points = source->meta().numPoints;
    const double genS = std::chrono::duration<double>(clock::now() - t0).count();

    // App::computeBudget (src/shell/App.cpp:97-114)
    remo::DeviceBudget budget;
    remo_device_mem_info(&budget.vramFreeAtStartup, &budget.vramTotal);
    if (deviceBudgetMb)
        budget.bytes = size_t(deviceBudgetMb) << 20;
    else {
        const size_t reserve = size_t(512) << 20;
        budget.bytes = budget.vramFreeAtStartup > reserve ? size_t(double(budget.vramFreeAtStartup - reserve) * 0.85) : 0;
    }

    remo::RemolodHipPipeline::Options opt;
    opt.accumulate = accumulate;
    opt.batchesPerStep = perStep;
    opt.stepBudgetMs = budgetMs;
    opt.slotCapacity = slot;
    opt.fixedDepth = fixedDepth;
    remo::PipelineRegistry registry;
    registry.add([opt] { return std::make_unique<remo::RemolodHipPipeline>(opt); });
    std::string err;
    if (!registry.switchTo("remolod", source.get(), source->meta(), budget, &err)) {
        std::fprintf(stderr, "remolod_headless: %s\n", err.c_str());
        return 1;
    }
    auto* pipe = static_cast<remo::RemolodHipPipeline*>(registry.active());
    if (asyncUpload && !hostCloud->enableAsync(pipe->context(), perStep)) {
        std::fprintf(stderr, "remolod_headless: pinned host memory for the cloud was refused\n");
        return 1;
    }

    uint64_t* fb = nullptr;
    if (remo_dev_alloc(size_t(width) * height * 8, reinterpret_cast<void**>(&fb)) != REMO_OK) {
        std::fprintf(stderr, "framebuffer allocation failed\n");
        return 1;
    }
    remo::FrameContext frame;
    frame.uniforms = overviewUniforms(source->meta().boxSize, width, height);
    if (white) frame.uniforms.color_mode = 3;
    frame.targets = {fb, width, height};
    frame.profile = profile;

    const auto t1 = clock::now();
    uint32_t frames = 0, refineCalls = 0;
    uint64_t refineNodes = 0;
    double refineDeviceMs = 0.0;
    for (;;) {
        frame.uniforms.frame_counter = frames;
        pipe->render(frame);
        source->pump();
        if (!source->lastError().empty()) {
            std::fprintf(stderr, "remolod_headless: %s\n", source->lastError().c_str());
            return 1;
        }
        const bool more = pipe->build(*source, frame);
        if (more && analysis == "analytic" && pipe->batchesConsumed() > 0) pipe->analyse(frame, REMO_ANALYSIS_ANALYTIC, &err);
        if (more && refineBudgetMs >= 0.0f && pipe->batchesConsumed() > 0) {
            remo_refine_stats rs{};
            if (pipe->refine(refineBudgetMs, &rs, &err)) {
                ++refineCalls;
                refineNodes += rs.nodes_refined;
                refineDeviceMs += rs.device_ms;
            }
        }
        ++frames;
        if (!more) break;
        if (frames > 1'000'000) break;
    }
    if (!pipe->lastError().empty()) {
        std::fprintf(stderr, "remolod_headless: %s\n", pipe->lastError().c_str());
        return 1;
    }
    if (analysis == "analytic") pipe->analyse(frame, REMO_ANALYSIS_ANALYTIC, &err);
    remo_refine_stats finalRefine{};
    if (refineBudgetMs >= 0.0f && !pipe->refine(-1.0f, &finalRefine, &err)) {
        std::fprintf(stderr, "remolod_headless: %s\n", err.c_str());
        return 1;
    }
    pipe->render(frame);
    remo_sync(pipe->context());
    const double loopS = std::chrono::duration<double>(clock::now() - t1).count();
    const remo_stats& s = pipe->rawStats();
    const remo::PipelineStats& ps = pipe->stats();

    if (!dumpUniforms.empty()) {
        FILE* f = std::fopen(dumpUniforms.c_str(), "wb");
        if (f) {
            std::fwrite(&frame.uniforms, sizeof(frame.uniforms), 1, f);
            std::fclose(f);
        }
    }
    if (!dumpFb.empty()) {
        std::vector<uint64_t> host(size_t(width) * height);
        remo_memcpy_d2h(host.data(), fb, host.size() * 8);
        FILE* f = std::fopen(dumpFb.c_str(), "wb");
        if (f) {
            std::fwrite(host.data(), 8, host.size(), f);
            std::fclose(f);
        }
    }
    remo_dev_free(fb);

    std::printf("{\"pipeline\": \"%s\", \"points\": %llu, \"seed\": %llu, \"frames\": %u, "
                "\"generate_s\": %.3f, \"loop_s\": %.4f, \"mpts_per_s\": %.2f, "
                "\"num_nodes\": %u, \"num_inner\": %u, \"num_leaves\": %u, \"num_points\": %llu, "
                "\"num_voxels\": %llu, \"num_points_ingested\": %llu, \"max_depth\": %u, "
                "\"batchlet_index\": %u, \"sum_leaf_counts\": %llu, \"inner_with_sums\": %u, "
                "\"visible_nodes\": %u, \"visible_samples\": %llu, \"mem_capacity_reached\": %d, "
                "\"box_size\": [%.9g, %.9g, %.9g]",
                registry.activeId().c_str(), (unsigned long long)points, (unsigned long long)seed, frames, genS, loopS,
                double(ps.numPointsIngested) / loopS / 1e6, ps.numNodes, ps.numInner, ps.numLeaves,
                (unsigned long long)ps.numPoints, (unsigned long long)ps.numVoxels,
                (unsigned long long)ps.numPointsIngested, ps.maxDepth, s.batchlet_index,
                (unsigned long long)s.sum_leaf_counts, s.inner_with_sums, ps.numVisibleNodes,
                (unsigned long long)ps.numVisiblePoints, int(ps.memCapacityReached), double(source->meta().boxSize[0]),
                double(source->meta().boxSize[1]), double(source->meta().boxSize[2]));
    if (refineBudgetMs >= 0.0f)
        std::printf(", \"refine\": {\"budget_ms\": %.3f, \"calls\": %u, \"nodes_refined\": %llu, "
                    "\"device_ms\": %.3f, \"final_nodes_refined\": %u, \"final_complete\": %u, "
                    "\"final_skipped\": %u}",
                    refineBudgetMs, refineCalls, (unsigned long long)refineNodes, refineDeviceMs,
                    finalRefine.nodes_refined, finalRefine.complete, finalRefine.nodes_skipped);
    std::printf(", \"diagnostics\": [");
    const auto diag = pipe->diagnostics();
    for (size_t i = 0; i < diag.size(); ++i) {
        std::string d = diag[i];
        for (auto& ch : d)
            if (ch == '\t') ch = ' ';
        std::printf("%s\"%s\"", i ? ", " : "", d.c_str());
    }
    std::printf("]");
    if (profile) {
        std::printf(", \"scopes\": {");
        const auto tot = pipe->scopeTotals();
        for (size_t i = 0; i < tot.size(); ++i)
            std::printf("%s\"%s\": [%.4f, %llu]", i ? ", " : "", tot[i].name.c_str(), tot[i].ms,
                        (unsigned long long)tot[i].launches);
        std::printf("}");
    }
    std::printf("}\n");
    registry.releaseActive();
    return 0;
}

// Learned score: training and inference kernels of the 13 -> 64 -> 64 -> 2 MLP,
// and its C ABI (remo_mlp_*). No reference code exists for this part
// (plans/04_NeuralPass.md is empty); the specification is a 13-input MLP with
// GELU, BF16 weights and FP32 accumulation, trained with Adam on (partial,
// final) pairs the analytic pipeline logs.
//
// Training keeps an FP32 master copy of the parameters and Adam moments; every
// step rounds the master to BF16 (round to nearest even), and the forward and
// backward passes run on the BF16 values with FP32 activations and sums:
//   kMlpGrad   one wave per 32 samples: forward, loss, backward; the per-sample
//              vectors go to LDS and the wave reduces them into this group's
//              partial gradient (5,186 values) in sample order;
//   kMlpAdam   one thread per parameter: sums the groups' partials in group order
//              (deterministic), Adam update of the master, BF16 copy;
//   kMlpInfer  inference on the matrix cores (BF16 MFMA), 16 samples per wave.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>

#include "../../include/remolod.h"
#include "remolod_mlp.h"

namespace remo_hip {

int mlpAllocate(AnalysisState& a, const AllocFn& alloc) {
    void* p = nullptr;
    const int rc = alloc(&p, sizeof(uint16_t) * kMlpParams, "learned-score parameters");
    a.mlpWeights = static_cast<uint16_t*>(p);
    a.mlpLoaded = false;
    return rc;
}

namespace {

constexpr uint32_t kGroup = 32;      // samples per gradient workgroup
constexpr uint32_t kGradThreads = 64;   // one wave: lanes < kGroup run a sample, all reduce
constexpr uint32_t kRow = kMlpHid + 1;  // padded LDS row: lanes write their own rows
constexpr uint32_t kPartial = kMlpParams + 1;  // a group's gradient + its loss sum

// BF16 copy of the master parameters and zeroed Adam moments.
__global__ void kMlpReset(const float* master, uint16_t* bf16, float* m, float* v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < kMlpParams; i += gridDim.x * blockDim.x) {
        bf16[i] = f32ToBf16(master[i]);
        m[i] = 0.0f;
        v[i] = 0.0f;
    }
}

// Inference: rows of x -> rows of y on the matrix cores, 16 rows per wave. The
// row count comes from nPtr (device, e.g. the node count) when given, else n.
__global__ __launch_bounds__(kMlpInferThreads) void kMlpInfer(const uint16_t* wb, const float* x, const uint32_t* nPtr,
                                                              uint32_t n, float* y) {
    __shared__ MlpWaveShared sh[kMlpWavesPerBlock];
    const uint32_t rows = nPtr ? *nPtr : n;
    const uint32_t wave = threadIdx.x / 64;
    if ((blockIdx.x * kMlpWavesPerBlock + wave) * 16 >= rows) return;  // the wave has no tile (no barrier below)
    MlpFrag f;
    mlpLoadFrag(wb, f);
    for (uint32_t tile = blockIdx.x * kMlpWavesPerBlock + wave; tile * 16 < rows;
         tile += gridDim.x * kMlpWavesPerBlock) {
        mlpLoadRows(x, rows, tile * 16, sh[wave].in);
        mlpInferTile(f, sh[wave].in, sh[wave].act, rows, tile * 16, y + uint64_t(tile) * 16 * kMlpOut);
    }
}

// Loss L = 1 / (n * outputs) * sum (y - t)^2 over the batch.
__global__ __launch_bounds__(kGradThreads) void kMlpGrad(const uint16_t* wb, const float* x, const float* tgt,
                                                         uint32_t n, float* partials) {
    __shared__ float w[kMlpParams];
    __shared__ float sx[kGroup][kMlpIn];
    __shared__ float sh1[kGroup][kRow];
    __shared__ float sh2[kGroup][kRow];
    __shared__ float sd1[kGroup][kRow];  // z1, then dL/dz1
    __shared__ float sd2[kGroup][kRow];  // z2, then dL/dz2
    __shared__ float sdy[kGroup][kMlpOut];
    __shared__ float sloss[kGroup];
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < kMlpParams; i += kGradThreads) w[i] = bf16ToF32(wb[i]);
    const uint32_t lane = tid < kGroup ? tid : 0;
    const uint32_t s = blockIdx.x * kGroup + lane;
    const bool active = tid < kGroup && s < n;
    if (tid < kGroup)
        for (uint32_t k = 0; k < kMlpIn; ++k) sx[lane][k] = active ? x[uint64_t(s) * kMlpIn + k] : 0.0f;
    __syncthreads();

    // forward, keeping the pre-activations in the dz rows
    float y[kMlpOut];
    if (tid < kGroup) mlpForward(w, sx[lane], y, sd1[lane], sd2[lane], sh1[lane], sh2[lane]);
    if (tid < kGroup) {
        // loss and dL/dy (zero for the padding samples past n)
        float loss = 0.0f;
        const float inv = 1.0f / (float(n) * float(kMlpOut));
        for (uint32_t o = 0; o < kMlpOut; ++o) {
            const float d = active ? y[o] - tgt[uint64_t(s) * kMlpOut + o] : 0.0f;
            loss += d * d;
            sdy[lane][o] = 2.0f * d * inv;
        }
        sloss[lane] = loss;
        // backward through layer 3 and the second GELU
        for (uint32_t j = 0; j < kMlpHid; ++j) {
            float dh = 0.0f;
            for (uint32_t o = 0; o < kMlpOut; ++o) dh += w[kW3 + o * kMlpHid + j] * sdy[lane][o];
            sd2[lane][j] = dh * geluGrad(sd2[lane][j]);
        }
        // through layer 2 and the first GELU
        for (uint32_t j = 0; j < kMlpHid; ++j) {
            float dh = 0.0f;
            for (uint32_t i = 0; i < kMlpHid; ++i) dh += w[kW2 + i * kMlpHid + j] * sd2[lane][i];
            sd1[lane][j] = dh * geluGrad(sd1[lane][j]);
        }
    }
    __syncthreads();

    // this group's gradient: every parameter sums its per-sample terms in order
    float* out = partials + uint64_t(blockIdx.x) * kPartial;
    for (uint32_t p = tid; p < kMlpParams; p += kGradThreads) {
        float acc = 0.0f;
        if (p < kB1) {
            const uint32_t i = p / kMlpIn, j = p % kMlpIn;
            for (uint32_t q = 0; q < kGroup; ++q) acc += sd1[q][i] * sx[q][j];
        } else if (p < kW2) {
            const uint32_t i = p - kB1;
            for (uint32_t q = 0; q < kGroup; ++q) acc += sd1[q][i];
        } else if (p < kB2) {
            const uint32_t i = (p - kW2) / kMlpHid, j = (p - kW2) % kMlpHid;
            for (uint32_t q = 0; q < kGroup; ++q) acc += sd2[q][i] * sh1[q][j];
        } else if (p < kW3) {
            const uint32_t i = p - kB2;
            for (uint32_t q = 0; q < kGroup; ++q) acc += sd2[q][i];
        } else if (p < kB3) {
            const uint32_t o = (p - kW3) / kMlpHid, j = (p - kW3) % kMlpHid;
            for (uint32_t q = 0; q < kGroup; ++q) acc += sdy[q][o] * sh2[q][j];
        } else {
            const uint32_t o = p - kB3;
            for (uint32_t q = 0; q < kGroup; ++q) acc += sdy[q][o];
        }
        out[p] = acc;
    }
    if (tid == 0) {
        float l = 0.0f;
        for (uint32_t q = 0; q < kGroup; ++q) l += sloss[q];
        out[kMlpParams] = l;
    }
}

struct AdamArgs {
    float lr, beta1, beta2, eps;
    float corr1, corr2;  // 1 - beta1^t, 1 - beta2^t
    uint32_t apply;      // 0: gradient only
    uint32_t groups;
    uint32_t n;
};

__global__ __launch_bounds__(256) void kMlpAdam(const float* partials, AdamArgs a, float* master, float* m, float* v,
                                                uint16_t* bf16, float* grad, float* loss) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p <= kMlpParams; p += gridDim.x * blockDim.x) {
        float g = 0.0f;
        for (uint32_t k = 0; k < a.groups; ++k) g += partials[uint64_t(k) * kPartial + p];
        if (p == kMlpParams) {
            *loss = g / (float(a.n) * float(kMlpOut));
            continue;
        }
        grad[p] = g;
        if (!a.apply) continue;
        const float mk = a.beta1 * m[p] + (1.0f - a.beta1) * g;
        const float vk = a.beta2 * v[p] + (1.0f - a.beta2) * g * g;
        m[p] = mk;
        v[p] = vk;
        const float mh = mk / a.corr1, vh = vk / a.corr2;
        const float th = master[p] - a.lr * mh / (sqrtf(vh) + a.eps);
        master[p] = th;
        bf16[p] = f32ToBf16(th);
    }
}

// SplitMix64 (Steele, Lea, Flood 2014): the parameter initialiser's generator.
uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

void launchMlpInfer(const uint16_t* wb, const float* x, const uint32_t* nPtr, uint32_t n, float* y, uint32_t blocks,
                    hipStream_t s) {
    hipLaunchKernelGGL(kMlpInfer, dim3(std::max(1u, blocks)), dim3(kMlpInferThreads), 0, s, wb, x, nPtr, n, y);
}

}  // namespace remo_hip

using namespace remo_hip;

struct remo_mlp {
    hipStream_t stream = nullptr;
    float* master = nullptr;
    uint16_t* bf16 = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    float* grad = nullptr;
    float* loss = nullptr;
    float* partials = nullptr;
    uint32_t maxBatch = 0;
    uint64_t step = 0;
    bool dead = false;
};

namespace {

void mlpFree(remo_mlp* p) {
    for (void* q : {(void*)p->master, (void*)p->bf16, (void*)p->m, (void*)p->v, (void*)p->grad, (void*)p->loss,
                    (void*)p->partials})
        if (q) (void)hipFree(q);
    if (p->stream) (void)hipStreamDestroy(p->stream);
}

int mlpSync(remo_mlp* p) {
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(p->stream) != hipSuccess) {
        p->dead = true;
        return REMO_ERR_DEVICE;
    }
    return REMO_OK;
}

int runGrad(remo_mlp* p, const float* x, const float* t, uint32_t n, float lr, bool apply, float* lossOut) {
    if (!p || !x || !t || n == 0 || n > p->maxBatch) return REMO_ERR_INVALID;
    if (p->dead) return REMO_ERR_DEVICE;
    const uint32_t groups = (n + kGroup - 1) / kGroup;
    hipLaunchKernelGGL(kMlpGrad, dim3(groups), dim3(kGradThreads), 0, p->stream, p->bf16, x, t, n, p->partials);
    AdamArgs a{};
    a.lr = lr;
    a.beta1 = 0.9f;
    a.beta2 = 0.999f;
    a.eps = 1e-8f;
    const uint64_t step = apply ? p->step + 1 : p->step;
    a.corr1 = 1.0f - float(std::pow(0.9, double(std::max<uint64_t>(step, 1))));
    a.corr2 = 1.0f - float(std::pow(0.999, double(std::max<uint64_t>(step, 1))));
    a.apply = apply ? 1u : 0u;
    a.groups = groups;
    a.n = n;
    hipLaunchKernelGGL(kMlpAdam, dim3((kMlpParams + 256) / 256), dim3(256), 0, p->stream, p->partials, a, p->master,
                       p->m, p->v, p->bf16, p->grad, p->loss);
    if (apply) p->step = step;
    if (lossOut) {
        if (hipMemcpyAsync(lossOut, p->loss, 4, hipMemcpyDeviceToHost, p->stream) != hipSuccess) return REMO_ERR_DEVICE;
    }
    return mlpSync(p);
}

}  // namespace

extern "C" {

int remo_mlp_create(uint64_t seed, uint32_t max_batch, remo_mlp** out) {
    if (!out || max_batch == 0) return REMO_ERR_INVALID;
    *out = nullptr;
    remo_mlp* p = new (std::nothrow) remo_mlp();
    if (!p) return REMO_ERR_OOM;
    p->maxBatch = max_batch;
    const uint64_t groups = (uint64_t(max_batch) + kGroup - 1) / kGroup;
    bool ok = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipMalloc(&p->master, 4 * kMlpParams) == hipSuccess;
    ok = ok && hipMalloc(&p->bf16, 2 * kMlpParams) == hipSuccess;
    ok = ok && hipMalloc(&p->m, 4 * kMlpParams) == hipSuccess;
    ok = ok && hipMalloc(&p->v, 4 * kMlpParams) == hipSuccess;
    ok = ok && hipMalloc(&p->grad, 4 * kMlpParams) == hipSuccess;
    ok = ok && hipMalloc(&p->loss, 16) == hipSuccess;
    ok = ok && hipMalloc(&p->partials, 4 * groups * kPartial) == hipSuccess;
    if (!ok) {
        mlpFree(p);
        delete p;
        return REMO_ERR_OOM;
    }
    // Glorot-uniform weights from the seed, zero biases
    float host[kMlpParams];
    uint64_t st = seed;
    auto fill = [&](uint32_t off, uint32_t rows, uint32_t cols) {
        const double a = std::sqrt(6.0 / double(rows + cols));
        for (uint32_t i = 0; i < rows * cols; ++i) {
            const double u = double(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
            host[off + i] = float((2.0 * u - 1.0) * a);
        }
    };
    std::memset(host, 0, sizeof(host));
    fill(kW1, kMlpHid, kMlpIn);
    fill(kW2, kMlpHid, kMlpHid);
    fill(kW3, kMlpOut, kMlpHid);
    const int rc = remo_mlp_set_params(p, host, kMlpParams);
    if (rc != REMO_OK) {
        mlpFree(p);
        delete p;
        return rc;
    }
    *out = p;
    return REMO_OK;
}

void remo_mlp_destroy(remo_mlp* p) {
    if (!p) return;
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    mlpFree(p);
    delete p;
}

int remo_mlp_get_params(remo_mlp* p, float* host, uint32_t n) {
    if (!p || !host || n != kMlpParams) return REMO_ERR_INVALID;
    if (p->dead) return REMO_ERR_DEVICE;
    if (hipMemcpyAsync(host, p->master, 4 * kMlpParams, hipMemcpyDeviceToHost, p->stream) != hipSuccess)
        return REMO_ERR_DEVICE;
    return mlpSync(p);
}

int remo_mlp_set_params(remo_mlp* p, const float* host, uint32_t n) {
    if (!p || !host || n != kMlpParams) return REMO_ERR_INVALID;
    if (p->dead) return REMO_ERR_DEVICE;
    for (uint32_t i = 0; i < n; ++i)
        if (!std::isfinite(host[i])) return REMO_ERR_INVALID;
    if (hipMemcpyAsync(p->master, host, 4 * kMlpParams, hipMemcpyHostToDevice, p->stream) != hipSuccess)
        return REMO_ERR_DEVICE;
    // BF16 copy and a fresh optimiser state
    hipLaunchKernelGGL(kMlpReset, dim3(32), dim3(256), 0, p->stream, p->master, p->bf16, p->m, p->v);
    p->step = 0;
    return mlpSync(p);
}

int remo_mlp_forward(remo_mlp* p, const float* x, uint32_t n, float* y) {
    if (!p || !x || !y) return REMO_ERR_INVALID;
    if (p->dead) return REMO_ERR_DEVICE;
    if (n == 0) return REMO_OK;
    launchMlpInfer(p->bf16, x, nullptr, n, y, std::min<uint32_t>((n + 63) / 64, 2048), p->stream);
    return mlpSync(p);
}

int remo_mlp_gradient(remo_mlp* p, const float* x, const float* t, uint32_t n, float* loss, float* grad_host) {
    const int rc = runGrad(p, x, t, n, 0.0f, false, loss);
    if (rc != REMO_OK || !grad_host) return rc;
    if (hipMemcpyAsync(grad_host, p->grad, 4 * kMlpParams, hipMemcpyDeviceToHost, p->stream) != hipSuccess)
        return REMO_ERR_DEVICE;
    return mlpSync(p);
}

int remo_mlp_train_step(remo_mlp* p, const float* x, const float* t, uint32_t n, float lr, float* loss) {
    if (!(lr > 0.0f)) return REMO_ERR_INVALID;
    return runGrad(p, x, t, n, lr, true, loss);
}

}  // extern "C"

int remo_hip::mlpCopyWeights(const remo_mlp* p, uint16_t* dst) {
    if (!p || !dst) return REMO_ERR_INVALID;
    if (p->dead) return REMO_ERR_DEVICE;
    if (hipMemcpyAsync(dst, p->bf16, 2 * kMlpParams, hipMemcpyDeviceToDevice, p->stream) != hipSuccess ||
        hipStreamSynchronize(p->stream) != hipSuccess)
        return REMO_ERR_DEVICE;
    return REMO_OK;
}

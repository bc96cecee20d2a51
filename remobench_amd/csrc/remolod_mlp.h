// Learned score: a 13 -> 64 -> 64 -> 2 MLP with exact (erf) GELU that predicts a
// node's final surface variation and planarity from its partial accumulators.
// BF16 weights, FP32 activations and accumulation; trained with hand-written
// forward / backward / Adam kernels (mlp.hip) on (partial, final) pairs logged
// from the analytic pipeline.
#pragma once

#include <hip/hip_runtime.h>

#include "remolod_analysis.h"

namespace remo_hip {

constexpr uint32_t kMlpIn = REMO_MLP_INPUTS;     // 13
constexpr uint32_t kMlpHid = REMO_MLP_HIDDEN;    // 64
constexpr uint32_t kMlpOut = REMO_MLP_OUTPUTS;   // 2
// parameter layout (row-major W[out][in], then biases)
constexpr uint32_t kW1 = 0;
constexpr uint32_t kB1 = kW1 + kMlpHid * kMlpIn;
constexpr uint32_t kW2 = kB1 + kMlpHid;
constexpr uint32_t kB2 = kW2 + kMlpHid * kMlpHid;
constexpr uint32_t kW3 = kB2 + kMlpHid;
constexpr uint32_t kB3 = kW3 + kMlpOut * kMlpHid;
constexpr uint32_t kMlpParams = kB3 + kMlpOut;
static_assert(kMlpParams == REMO_MLP_PARAMS, "parameter count");

// The fallback threshold: a predicted surface variation above this sends the
// node to the analytic eigen-solve.
constexpr float kLearnedFallbackSv = 0.05f;

// ---- BF16 (round to nearest even) ------------------------------------------------

__host__ __device__ inline uint16_t f32ToBf16(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return uint16_t((u >> 16) | 0x40u);  // quiet NaN
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7FFFu + lsb;
    return uint16_t(u >> 16);
}

__host__ __device__ inline float bf16ToF32(uint16_t h) {
    const uint32_t u = uint32_t(h) << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// ---- activation ------------------------------------------------------------------

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

// d gelu / dz = Phi(z) + z phi(z)
__device__ __forceinline__ float geluGrad(float z) {
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

// Forward pass of one sample with the weights already expanded to FP32 (in LDS).
// Training passes rows for the pre-activations z1 / z2 and activations h1 / h2.
// Each first-layer unit is scattered into the 64 second-layer sums as soon as it
// is computed, so only those 64 sums live in registers (every lane reads the
// same weight, an LDS broadcast).
__device__ __forceinline__ void mlpForward(const float* w, const float x[kMlpIn], float y[kMlpOut],
                                           float* z1 = nullptr, float* z2 = nullptr, float* h1out = nullptr,
                                           float* h2out = nullptr) {
    float acc2[kMlpHid];
#pragma unroll
    for (uint32_t i = 0; i < kMlpHid; ++i) acc2[i] = w[kB2 + i];
    for (uint32_t j = 0; j < kMlpHid; ++j) {
        float a = w[kB1 + j];
#pragma unroll
        for (uint32_t k = 0; k < kMlpIn; ++k) a += w[kW1 + j * kMlpIn + k] * x[k];
        if (z1) z1[j] = a;
        const float h = gelu(a);
        if (h1out) h1out[j] = h;
#pragma unroll
        for (uint32_t i = 0; i < kMlpHid; ++i) acc2[i] += w[kW2 + i * kMlpHid + j] * h;
    }
    float y0 = w[kB3], y1 = w[kB3 + 1];
#pragma unroll
    for (uint32_t i = 0; i < kMlpHid; ++i) {
        if (z2) z2[i] = acc2[i];
        const float h = gelu(acc2[i]);
        if (h2out) h2out[i] = h;
        y0 += w[kW3 + i] * h;
        y1 += w[kW3 + kMlpHid + i] * h;
    }
    y[0] = y0;
    y[1] = y1;
}
static_assert(kMlpOut == 2, "mlpForward writes two outputs");

// ---- features --------------------------------------------------------------------

// The 13 inputs of a node, from its rolled-up exact sums (units 2^-(20+e) of the
// node extent, centred on the node):
//   0      log(1 + n) / log(1 + 50,000)
//   1..6   covariance / trace (xx xy xz yy yz zz), node-extent units
//   7..9   2 * mean offset from the node centre (in [-1, 1])
//   10     level / 20
//   11     occupancy fill: inner nodes numVoxels / 128^3, leaves min(1, n / 50,000)
//   12     arrival: the fraction of the cloud's batches ingested so far
__device__ inline void nodeFeatures(const DevTree& t, uint32_t node, const RollupSums& S, float arrival,
                                    float f[kMlpIn]) {
    for (uint32_t k = 0; k < kMlpIn; ++k) f[k] = 0.0f;
    const double dn = double(S.n);
    f[0] = float(log1p(dn) / log1p(50000.0));
    f[10] = float(t.level[node]) / 20.0f;
    f[11] = t.firstChild[node] >= 0 ? float(t.numVoxels[node]) / 2097152.0f : float(fmin(1.0, dn / 50000.0));
    f[12] = arrival;
    if (S.n == 0) return;
    const double scale = ldexp(1.0, -int(S.e));
    double mean[3];
    for (int i = 0; i < 3; ++i) {
        mean[i] = i128ToDouble(S.s1[i]) / dn * scale;
        f[7 + i] = float(2.0 * mean[i]);
    }
    if (S.n < 2) return;
    const int pi[6] = {0, 0, 0, 1, 1, 2}, pj[6] = {0, 1, 2, 1, 2, 2};
    double cov[6];
    for (int k = 0; k < 6; ++k)
        cov[k] = i128ToDouble(S.s2[k]) / dn * scale * scale - mean[pi[k]] * mean[pj[k]];
    const double trace = cov[0] + cov[3] + cov[5];
    if (!(trace > 0.0)) return;
    for (int k = 0; k < 6; ++k) f[1 + k] = float(cov[k] / trace);
}

// The context's copy of a trained model (BF16 parameters).
int mlpAllocate(AnalysisState& a, const AllocFn& alloc);
// Copies a training object's BF16 parameters to `dst` (device) and waits.
int mlpCopyWeights(const remo_mlp* p, uint16_t* dst);

}  // namespace remo_hip

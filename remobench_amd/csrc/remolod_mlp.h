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

// ---- batched inference on the matrix cores ------------------------------------------
//
// One wave computes 16 rows with v_mfma_f32_16x16x4_f32 (FP32 operands: the BF16
// weights expanded, FP32 activations and accumulation). Operand layout of that
// instruction for lane l: A[row l % 16][k l / 16], B[k l / 16][col l % 16], and
// the 16x16 result D[row 4 (l / 16) + r][col l % 16] in element r. Layer outputs
// go through LDS to become the next layer's A operand.
typedef float mlpF4 __attribute__((ext_vector_type(4)));
constexpr uint32_t kMlpWavesPerBlock = 4;
constexpr uint32_t kMlpInferThreads = 64 * kMlpWavesPerBlock;
// LDS rows padded to 68 floats: lane (r16, k4) of an operand read touches bank
// 4 r16 + k4, and a result write (row 4 k4 + r, column r16) bank 16 k4 + r16, so
// neither conflicts.
constexpr uint32_t kMlpActRow = kMlpHid + 4;
constexpr uint32_t kMlpW2Row = kMlpHid + 4;
// the parameters in LDS: the global layout with W2's rows padded to kMlpW2Row
constexpr uint32_t kMlpShW2 = kW2;
constexpr uint32_t kMlpShB2 = kMlpShW2 + kMlpHid * kMlpW2Row;
constexpr uint32_t kMlpShW3 = kMlpShB2 + kMlpHid;
constexpr uint32_t kMlpShB3 = kMlpShW3 + kMlpOut * kMlpHid;
constexpr uint32_t kMlpShParams = kMlpShB3 + kMlpOut;

struct MlpInferShared {
    float w[kMlpShParams];
    float act[kMlpWavesPerBlock][16][kMlpActRow];
};

__device__ __forceinline__ uint32_t mlpShIndex(uint32_t p) {
    if (p < kW2) return p;
    if (p < kB2) return kMlpShW2 + (p - kW2) / kMlpHid * kMlpW2Row + (p - kW2) % kMlpHid;
    return p + kMlpHid * (kMlpW2Row - kMlpHid);
}

// The BF16 parameters (16-byte aligned, as every device allocation is) expanded
// into the LDS layout by the whole block of kMlpInferThreads: 16-byte loads, all
// issued before the first store.
__device__ inline void mlpStageWeights(const uint16_t* wb, float* w) {
    constexpr uint32_t kGroups = kMlpParams / 8;
    constexpr uint32_t kPer = (kGroups + kMlpInferThreads - 1) / kMlpInferThreads;
    const uint4* v = reinterpret_cast<const uint4*>(wb);
    uint4 r[kPer];
#pragma unroll
    for (uint32_t i = 0; i < kPer; ++i) {
        const uint32_t g = threadIdx.x + i * kMlpInferThreads;
        if (g < kGroups) r[i] = v[g];
    }
#pragma unroll
    for (uint32_t i = 0; i < kPer; ++i) {
        const uint32_t g = threadIdx.x + i * kMlpInferThreads;
        if (g >= kGroups) continue;
        const uint32_t h[4] = {r[i].x, r[i].y, r[i].z, r[i].w};
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) {
            const uint32_t p = g * 8 + 2 * q;
            w[mlpShIndex(p)] = bf16ToF32(uint16_t(h[q] & 0xFFFFu));
            w[mlpShIndex(p + 1)] = bf16ToF32(uint16_t(h[q] >> 16));
        }
    }
    for (uint32_t p = kGroups * 8 + threadIdx.x; p < kMlpParams; p += kMlpInferThreads) w[mlpShIndex(p)] = bf16ToF32(wb[p]);
}

// Rows [row0, row0 + 16) of x (n rows of kMlpIn) into act[0..16)[0..16), zero
// past the inputs and past n.
__device__ inline void mlpLoadRows(const float* x, uint32_t n, uint32_t row0, float (*act)[kMlpActRow]) {
    const uint32_t lane = __lane_id();
    const uint32_t r16 = lane % 16, k4 = lane / 16;
    const uint32_t row = row0 + r16;
    for (uint32_t ks = 0; ks < 4; ++ks) {
        const uint32_t k = ks * 4 + k4;
        act[r16][k] = (row < n && k < kMlpIn) ? x[uint64_t(row) * kMlpIn + k] : 0.0f;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// One wave: the 16 input rows in act[0..16)[0..16) (w in the LDS layout) -> the
// outputs of rows row0 .. row0 + 15 that are below n, at y[(row - row0) kMlpOut + o]. Each layer reads all its operands
// first and then issues its MFMAs k-step by k-step across the four independent
// column tiles, so neither LDS latency nor the MFMA dependency is exposed per
// instruction; every column tile still sums its k-steps in order.
__device__ inline void mlpInferTile(const float* w, float (*act)[kMlpActRow], uint32_t n, uint32_t row0, float* y) {
    const uint32_t lane = __lane_id();
    const uint32_t r16 = lane % 16, k4 = lane / 16;
    mlpF4 acc[4];
    // layer 1: [16 x 16 (13 used)] x [16 x 64]
    {
        float a[4], b[4][4];
#pragma unroll
        for (uint32_t ks = 0; ks < 4; ++ks) {
            const uint32_t k = ks * 4 + k4;
            a[ks] = act[r16][k];
#pragma unroll
            for (uint32_t nt = 0; nt < 4; ++nt) b[nt][ks] = k < kMlpIn ? w[kW1 + (nt * 16 + r16) * kMlpIn + k] : 0.0f;
        }
#pragma unroll
        for (uint32_t nt = 0; nt < 4; ++nt) acc[nt] = mlpF4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (uint32_t ks = 0; ks < 4; ++ks)
#pragma unroll
            for (uint32_t nt = 0; nt < 4; ++nt)
                acc[nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[ks], b[nt][ks], acc[nt], 0, 0, 0);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // every lane has read the inputs
#pragma unroll
    for (uint32_t nt = 0; nt < 4; ++nt)
#pragma unroll
        for (uint32_t r = 0; r < 4; ++r) act[4 * k4 + r][nt * 16 + r16] = gelu(acc[nt][r] + w[kB1 + nt * 16 + r16]);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // layer 2: [16 x 64] x [64 x 64]
    {
        float a[16], b[4][16];
#pragma unroll
        for (uint32_t ks = 0; ks < 16; ++ks) {
            const uint32_t k = ks * 4 + k4;
            a[ks] = act[r16][k];
#pragma unroll
            for (uint32_t nt = 0; nt < 4; ++nt) b[nt][ks] = w[kMlpShW2 + (nt * 16 + r16) * kMlpW2Row + k];
        }
#pragma unroll
        for (uint32_t nt = 0; nt < 4; ++nt) acc[nt] = mlpF4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (uint32_t ks = 0; ks < 16; ++ks)
#pragma unroll
            for (uint32_t nt = 0; nt < 4; ++nt)
                acc[nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[ks], b[nt][ks], acc[nt], 0, 0, 0);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // every lane has read layer 1
#pragma unroll
    for (uint32_t nt = 0; nt < 4; ++nt)
#pragma unroll
        for (uint32_t r = 0; r < 4; ++r)
            act[4 * k4 + r][nt * 16 + r16] = gelu(acc[nt][r] + w[kMlpShB2 + nt * 16 + r16]);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // layer 3: [16 x 64] x [64 x 16 (2 used)]
    mlpF4 out = mlpF4{0.0f, 0.0f, 0.0f, 0.0f};
    {
        float a[16], b[16];
#pragma unroll
        for (uint32_t ks = 0; ks < 16; ++ks) {
            const uint32_t k = ks * 4 + k4;
            a[ks] = act[r16][k];
            b[ks] = r16 < kMlpOut ? w[kMlpShW3 + r16 * kMlpHid + k] : 0.0f;
        }
#pragma unroll
        for (uint32_t ks = 0; ks < 16; ++ks) out = __builtin_amdgcn_mfma_f32_16x16x4f32(a[ks], b[ks], out, 0, 0, 0);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // act is reused by the next tile
    if (r16 < kMlpOut)
        for (uint32_t r = 0; r < 4; ++r) {
            const uint32_t orow = row0 + 4 * k4 + r;
            if (orow < n) y[(4 * k4 + r) * kMlpOut + r16] = out[r] + w[kMlpShB3 + r16];
        }
}

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
// Batched inference (kMlpInfer): rows of x -> rows of y; the row count is *nPtr
// (device) when nPtr is given, else n.
void launchMlpInfer(const uint16_t* wb, const float* x, const uint32_t* nPtr, uint32_t n, float* y, uint32_t blocks,
                    hipStream_t s);

}  // namespace remo_hip

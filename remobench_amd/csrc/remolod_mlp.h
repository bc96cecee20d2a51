// Learned score: a 13 -> 64 -> 64 -> 2 MLP with exact (erf) GELU that predicts
// how much a node's surface variation and planarity will still change before the
// node is complete, from its partial accumulators: the prediction is a correction
// to the analytic values of the partial sums. BF16 weights, FP32 accumulation
// (inference rounds each layer's inputs to BF16 for the matrix cores; training
// keeps FP32 activations); trained with hand-written forward / backward / Adam
// kernels (mlp.hip) on (partial, final) pairs logged from the analytic pipeline.
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

// An open node takes the corrected values only where the model predicts a change
// of its surface variation above kLearnedChangeSv or of its planarity above
// kLearnedChangePl; otherwise (and for every closed node) the analytic values of
// its current sums stand, which are exact for a node whose points have all
// arrived.
constexpr float kLearnedChangeSv = 0.05f;
constexpr float kLearnedChangePl = 0.10f;

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

// ---- batched inference on the matrix cores (BF16) ----------------------------------
//
// One wave computes 16 rows with v_mfma_f32_16x16x32_bf16: BF16 operands (the
// weights as stored; each layer's inputs rounded to BF16), FP32 accumulation.
// Operand layout for lane l: A[row l % 16][k 8 (l / 16) .. 8 (l / 16) + 7],
// B[k 8 (l / 16) .. + 7][col l % 16], and the 16x16 result D[row 4 (l / 16) + r]
// [col l % 16] in element r. The layers are 13 (padded to 32) -> 64 -> 64 -> 2
// (padded to 16 columns): 4 + 8 + 2 MFMAs per 16 rows.
//
// Every lane keeps its own weight fragments in registers for the whole kernel
// (14 x 8 BF16 and 9 biases, loaded once from global memory), so no block stages
// the parameters; a layer's outputs go through a per-wave LDS tile (BF16 rows) to
// become the next layer's A operand.
typedef float mlpF4 __attribute__((ext_vector_type(4)));
typedef __bf16 mlpBf8 __attribute__((ext_vector_type(8)));
constexpr uint32_t kMlpWavesPerBlock = 4;
constexpr uint32_t kMlpInferThreads = 64 * kMlpWavesPerBlock;
constexpr uint32_t kMlpK1 = 32;                 // layer 1's K: 13 inputs, zero-padded
// per-wave activation tile: 16 rows of 64 BF16, padded to 72 (144 B: rows stay
// 16-byte aligned for the operand loads, and consecutive rows start 4 banks apart)
constexpr uint32_t kMlpActRow = kMlpHid + 8;

struct MlpFrag {
    mlpBf8 w1[4];      // layer 1: column tile nt, k = 8 (l / 16) + 0..7 (k >= 13: zero)
    mlpBf8 w2[4][2];   // layer 2: column tile nt, k-step ks: k = 32 ks + 8 (l / 16) + 0..7
    mlpBf8 w3[2];      // layer 3: k-step ks, column l % 16 (columns >= 2: zero)
    float b1[4], b2[4], b3;
};

__device__ __forceinline__ __bf16 mlpBf(uint16_t h) { return __builtin_bit_cast(__bf16, h); }
__device__ __forceinline__ __bf16 mlpToBf(float f) { return __builtin_bit_cast(__bf16, f32ToBf16(f)); }

// Eight consecutive BF16 parameters starting at p (16-byte aligned) as one operand.
__device__ __forceinline__ mlpBf8 mlpLoad8(const uint16_t* wb, uint32_t p) {
    const uint4 v = *reinterpret_cast<const uint4*>(wb + p);
    const uint32_t h[4] = {v.x, v.y, v.z, v.w};
    mlpBf8 r;
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q) {
        r[2 * q] = mlpBf(uint16_t(h[q] & 0xFFFFu));
        r[2 * q + 1] = mlpBf(uint16_t(h[q] >> 16));
    }
    return r;
}

// This lane's fragments of the BF16 parameters (the global layout of the C ABI:
// W[out][in] row-major, the buffer 16-byte aligned; W2 and W3 start at multiples
// of 8 parameters, W1's 13-wide rows are read element by element).
__device__ inline void mlpLoadFrag(const uint16_t* wb, MlpFrag& f) {
    const uint32_t lane = __lane_id(), c16 = lane % 16, g = lane / 16;
    static_assert(kW2 % 8 == 0 && kW3 % 8 == 0, "aligned weight blocks");
#pragma unroll
    for (uint32_t nt = 0; nt < 4; ++nt) {
        const uint32_t col = nt * 16 + c16;
#pragma unroll
        for (uint32_t i = 0; i < 8; ++i) {
            const uint32_t k = 8 * g + i;
            f.w1[nt][i] = k < kMlpIn ? mlpBf(wb[kW1 + col * kMlpIn + k]) : mlpBf(0);
        }
#pragma unroll
        for (uint32_t ks = 0; ks < 2; ++ks) f.w2[nt][ks] = mlpLoad8(wb, kW2 + col * kMlpHid + 32 * ks + 8 * g);
        f.b1[nt] = bf16ToF32(wb[kB1 + col]);
        f.b2[nt] = bf16ToF32(wb[kB2 + col]);
    }
#pragma unroll
    for (uint32_t ks = 0; ks < 2; ++ks) {
        if (c16 < kMlpOut) {
            f.w3[ks] = mlpLoad8(wb, kW3 + c16 * kMlpHid + 32 * ks + 8 * g);
        } else {
#pragma unroll
            for (uint32_t i = 0; i < 8; ++i) f.w3[ks][i] = mlpBf(0);
        }
    }
    f.b3 = c16 < kMlpOut ? bf16ToF32(wb[kB3 + c16]) : 0.0f;
}

// One wave: 16 input rows (FP32, kMlpIn used of 16 per row, in `in`) -> the
// outputs of rows row0 .. row0 + 15 that are below n, at y[(row - row0) kMlpOut + o].
// `act` is the wave's BF16 tile. All of a layer's operands are read before its
// MFMAs issue; every column tile sums its k-steps in order.
__device__ inline void mlpInferTile(const MlpFrag& f, const float (*in)[16], __bf16 (*act)[kMlpActRow], uint32_t n,
                                    uint32_t row0, float* y) {
    const uint32_t lane = __lane_id();
    const uint32_t r16 = lane % 16, g = lane / 16;
    mlpF4 acc[4];
    // layer 1: [16 x 32 (13 used)] x [32 x 64]
    {
        mlpBf8 a;
#pragma unroll
        for (uint32_t i = 0; i < 8; ++i) {
            const uint32_t k = 8 * g + i;
            a[i] = mlpToBf(k < 16 ? in[r16][k] : 0.0f);
        }
#pragma unroll
        for (uint32_t nt = 0; nt < 4; ++nt)
            acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, f.w1[nt], mlpF4{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
    }
#pragma unroll
    for (uint32_t nt = 0; nt < 4; ++nt)
#pragma unroll
        for (uint32_t r = 0; r < 4; ++r) act[4 * g + r][nt * 16 + r16] = mlpToBf(gelu(acc[nt][r] + f.b1[nt]));
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // layer 2: [16 x 64] x [64 x 64]
    {
        const mlpBf8 a0 = *reinterpret_cast<const mlpBf8*>(&act[r16][8 * g]);
        const mlpBf8 a1 = *reinterpret_cast<const mlpBf8*>(&act[r16][32 + 8 * g]);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // every lane has read layer 1
#pragma unroll
        for (uint32_t nt = 0; nt < 4; ++nt) {
            acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, f.w2[nt][0], mlpF4{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
            acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, f.w2[nt][1], acc[nt], 0, 0, 0);
        }
    }
#pragma unroll
    for (uint32_t nt = 0; nt < 4; ++nt)
#pragma unroll
        for (uint32_t r = 0; r < 4; ++r) act[4 * g + r][nt * 16 + r16] = mlpToBf(gelu(acc[nt][r] + f.b2[nt]));
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // layer 3: [16 x 64] x [64 x 16 (2 used)]
    mlpF4 out;
    {
        const mlpBf8 a0 = *reinterpret_cast<const mlpBf8*>(&act[r16][8 * g]);
        const mlpBf8 a1 = *reinterpret_cast<const mlpBf8*>(&act[r16][32 + 8 * g]);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // act is reused by the next tile
        out = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, f.w3[0], mlpF4{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
        out = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, f.w3[1], out, 0, 0, 0);
    }
    if (r16 < kMlpOut)
        for (uint32_t r = 0; r < 4; ++r) {
            const uint32_t orow = row0 + 4 * g + r;
            if (orow < n) y[(4 * g + r) * kMlpOut + r16] = out[r] + f.b3;
        }
}

// Rows [row0, row0 + 16) of x (n rows of kMlpIn) into in[0..16)[0..16), zero
// past the inputs and past n.
__device__ inline void mlpLoadRows(const float* x, uint32_t n, uint32_t row0, float (*in)[16]) {
    const uint32_t lane = __lane_id();
    const uint32_t r16 = lane % 16, k4 = lane / 16;
    const uint32_t row = row0 + r16;
    for (uint32_t ks = 0; ks < 4; ++ks) {
        const uint32_t k = ks * 4 + k4;
        in[r16][k] = (row < n && k < kMlpIn) ? x[uint64_t(row) * kMlpIn + k] : 0.0f;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// The per-wave LDS of the inference kernels.
struct alignas(16) MlpWaveShared {
    float in[16][16];
    __bf16 act[16][kMlpActRow];
};

// ---- features --------------------------------------------------------------------

// The 13 inputs of a node, from its rolled-up exact sums (units 2^-(20+e) of the
// node extent, centred on the node):
//   0      log(1 + n) / log(1 + 50,000)
//   1..6   covariance / trace (xx xy xz yy yz zz), node-extent units
//   7..9   2 * mean offset from the node centre (in [-1, 1])
//   10     level / 20
//   11     occupancy fill: inner nodes numVoxels / 128^3, leaves min(1, n / 50,000)
//   12     arrival: the fraction of the cloud's batches ingested so far
// (level, whether the node is inner, its voxel count)
__device__ inline void nodeFeaturesOf(uint32_t level, bool inner, uint32_t numVoxels, const RollupSums& S,
                                      float arrival, float f[kMlpIn]) {
    for (uint32_t k = 0; k < kMlpIn; ++k) f[k] = 0.0f;
    const double dn = double(S.n);
    f[0] = float(log1p(dn) / log1p(50000.0));
    f[10] = float(level) / 20.0f;
    f[11] = inner ? float(numVoxels) / 2097152.0f : float(fmin(1.0, dn / 50000.0));
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

__device__ inline void nodeFeatures(const DevTree& t, uint32_t node, const RollupSums& S, float arrival,
                                    float f[kMlpIn]) {
    nodeFeaturesOf(t.level[node], t.firstChild[node] >= 0, t.numVoxels[node], S, arrival, f);
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

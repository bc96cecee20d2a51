// Accumulate: fold each leaf's new points into its running sums.
//
// Follows kernel_accumulate / kernel_accum_verify
// (kernels/remolod/remolod_accum.cu:47-233): one workgroup per leaf whose
// watermark count is below numPoints folds slots [count, numPoints); an inner
// node holding sums is cleared; Morton watermark of the folded points.
//
// Besides the reference's FP64 geometry sums g[9] and FP32 colour sums c[6], every
// point adds its node-centred fixed-point coordinate q = floor(l * 2^20) - 2^19
// (l the node-local coordinate in [0,1], computed in the reference's FP64 order)
// into int64 sums q1[3], q2[6]. Integer addition is associative, so those sums
// are bit-exact and independent of thread order; they are what Analysis rolls up.
//
// The wave reduction is 64-wide (__shfl_down over 64 lanes), then one LDS slot
// per wave, where the reference used 32-lane shuffles and 8 warps.

#include "remolod_internal.h"

namespace remo_hip {

namespace {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kWaves = kBlock / 64;

__device__ __forceinline__ uint64_t mortonSpread(uint32_t v) {
    uint64_t x = uint64_t(v) & 0xFFFFFull;
    x = (x | (x << 32)) & 0x001F00000000FFFFull;
    x = (x | (x << 16)) & 0x001F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

template <typename T>
__device__ __forceinline__ T waveSum(T v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    return v;
}

__device__ __forceinline__ uint64_t waveMax(uint64_t v) {
    for (int off = 32; off > 0; off >>= 1) {
        const uint64_t o = __shfl_down(v, off);
        v = o > v ? o : v;
    }
    return v;
}

__global__ void kAccumBegin(DevTree t) {
    Globals* g = t.g;
    g->pointsFolded = 0;
    g->leavesFolded = 0;
    g->sumLeafCounts = 0;
    g->innerWithSums = 0;
}

__global__ __launch_bounds__(kBlock) void kAccumulate(DevTree t, uint32_t batchIndex) {
    __shared__ long long sq[9][kWaves];
    __shared__ double sg[9][kWaves];
    __shared__ float sc[6][kWaves];
    __shared__ unsigned long long sm[kWaves];

    Globals* g = t.g;
    const uint32_t numNodes = g->numNodes;
    const uint32_t lane = laneId(), wave = threadIdx.x >> 6;
    const double invSize = 1.0 / double(t.octreeSize);
    const double fGrid = double(1u << kMaxDepth);

    for (uint32_t node = xcdBlock(); node < numNodes; node += gridDim.x) {
        remo_node_accum* a = t.accums + node;
        if (t.firstChild[node] >= 0) {
            if (threadIdx.x == 0 && a->count != 0) *a = remo_node_accum{};
            continue;
        }
        const uint32_t from = a->count;
        const uint32_t to = t.numPoints[node];
        if (to <= from) continue;

        const uint32_t lvl = t.level[node];
        const double levelScale = double(1ull << lvl);
        const double NX = double(t.nx[node]), NY = double(t.ny[node]), NZ = double(t.nz[node]);
        const uint32_t* dir = t.pointDir + uint64_t(node) * kPointDir;

        long long q[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        double lg[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        float lc[6] = {0, 0, 0, 0, 0, 0};
        uint64_t lm = 0;
        for (uint32_t s = from + threadIdx.x; s < to; s += kBlock) {
            const uint32_t k = s / kPointsPerChunk;
            const uint32_t ref = k < kPointDir ? dir[k] : 0u;
            if (ref == 0) continue;
            const remo_point p = chunkPtr(t.arena, ref)[s % kPointsPerChunk];
            // remolod_accum.cu:115-132, octreeMin = 0
            const double lx = double(p.x) * invSize;
            const double ly = double(p.y) * invSize;
            const double lz = double(p.z) * invSize;
            const uint32_t maxCoord = (1u << kMaxDepth) - 1u;
            const uint32_t X = min(uint32_t(fGrid * lx), maxCoord);
            const uint32_t Y = min(uint32_t(fGrid * ly), maxCoord);
            const uint32_t Z = min(uint32_t(fGrid * lz), maxCoord);
            const uint64_t m = (mortonSpread(X) << 2) | (mortonSpread(Y) << 1) | mortonSpread(Z);
            lm = m > lm ? m : lm;
            double x = lx * levelScale - NX;
            double y = ly * levelScale - NY;
            double z = lz * levelScale - NZ;
            x = fmin(fmax(x, 0.0), 1.0);
            y = fmin(fmax(y, 0.0), 1.0);
            z = fmin(fmax(z, 0.0), 1.0);
            lg[0] += x; lg[1] += y; lg[2] += z;
            lg[3] += x * x; lg[4] += x * y; lg[5] += x * z;
            lg[6] += y * y; lg[7] += y * z; lg[8] += z * z;
            const long long qx = (long long)floor(x * 1048576.0) - 524288;
            const long long qy = (long long)floor(y * 1048576.0) - 524288;
            const long long qz = (long long)floor(z * 1048576.0) - 524288;
            q[0] += qx; q[1] += qy; q[2] += qz;
            q[3] += qx * qx; q[4] += qx * qy; q[5] += qx * qz;
            q[6] += qy * qy; q[7] += qy * qz; q[8] += qz * qz;
            const float cr = float((p.color >> 0) & 0xFFu) * (1.0f / 255.0f);
            const float cg = float((p.color >> 8) & 0xFFu) * (1.0f / 255.0f);
            const float cb = float((p.color >> 16) & 0xFFu) * (1.0f / 255.0f);
            lc[0] += cr; lc[1] += cg; lc[2] += cb;
            lc[3] += cr * cr; lc[4] += cg * cg; lc[5] += cb * cb;
        }
        for (int i = 0; i < 9; ++i) {
            const long long v = waveSum(q[i]);
            const double d = waveSum(lg[i]);
            if (lane == 0) {
                sq[i][wave] = v;
                sg[i][wave] = d;
            }
        }
        for (int i = 0; i < 6; ++i) {
            const float v = waveSum(lc[i]);
            if (lane == 0) sc[i][wave] = v;
        }
        {
            const uint64_t v = waveMax(lm);
            if (lane == 0) sm[wave] = v;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int i = 0; i < 9; ++i) {
                long long v = 0;
                double d = 0.0;
                for (uint32_t w = 0; w < kWaves; ++w) {
                    v += sq[i][w];
                    d += sg[i][w];
                }
                if (i < 3)
                    a->q1[i] += v;
                else
                    a->q2[i - 3] += v;
                a->g[i] += d;
            }
            for (int i = 0; i < 6; ++i) {
                float v = 0.0f;
                for (uint32_t w = 0; w < kWaves; ++w) v += sc[i][w];
                a->c[i] += v;
            }
            unsigned long long mm = 0;
            for (uint32_t w = 0; w < kWaves; ++w) mm = sm[w] > mm ? sm[w] : mm;
            a->count = to;
            a->last_touched_batch = batchIndex;
            if (mm > a->morton_max) a->morton_max = mm;
            atomicMax((unsigned long long*)&g->mortonWatermark, mm);
            atomicAdd((unsigned long long*)&g->numAccumulated, (unsigned long long)(to - from));
            atomicAdd(&g->pointsFolded, to - from);
            atomicAdd(&g->leavesFolded, 1u);
        }
        __syncthreads();
    }
}

// kernel_accum_verify (remolod_accum.cu:214-233).
__global__ __launch_bounds__(kBlock) void kAccumVerify(DevTree t) {
    Globals* g = t.g;
    const uint32_t n = g->numNodes;
    unsigned long long sum = 0;
    uint32_t inner = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (t.firstChild[i] < 0)
            sum += t.accums[i].count;
        else if (t.accums[i].count != 0)
            inner++;
    }
    sum = waveSum(sum);
    inner = waveSum(inner);
    if (laneId() == 0) {
        if (sum) atomicAdd((unsigned long long*)&g->sumLeafCounts, sum);
        if (inner) atomicAdd(&g->innerWithSums, inner);
    }
}

}  // namespace

void launchAccumulate(const DevTree& t, uint32_t batchIndex, const LaunchShape& ls, hipStream_t s) {
    hipLaunchKernelGGL(kAccumBegin, dim3(1), dim3(1), 0, s, t);
    hipLaunchKernelGGL(kAccumulate, dim3(ls.pointBlocks), dim3(kBlock), 0, s, t, batchIndex);
    hipLaunchKernelGGL(kAccumVerify, dim3(ls.nodeBlocks), dim3(kBlock), 0, s, t);
}

}  // namespace remo_hip

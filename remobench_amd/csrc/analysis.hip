// Analysis and Rasterize: the per-frame passes over a built tree.
//
// Analysis has no reference code (kernels/remolod/remodlod_analysis.cu is empty);
// it follows the spec in README.md:80 and wiki/02_KernelPipeline.md:205-224:
// test closure, roll leaf accumulators up the tree and finalise, score =
// geometric score x screen coverage, rebuild the work queue each frame.
//
//  * Roll-up is level-synchronous, bottom-up: one launch writes every leaf's
//    sums, then one launch per level adds each inner node's eight children in
//    child order (a launch per level beats a one-workgroup walk with barriers
//    and an arrival-counter chain on these tree sizes). The sums are __int128 in
//    units U = 2^-(20+D) of the root extent (D the deepest level), centred on
//    each node: a child's sums move to its parent's centre with the exact offset
//    +-2^(18+D-L) per axis, so the roll-up is exact and order-independent.
//  * Finalise: covariance in node-extent units from num = n*S2 - S1*S1^T (exact
//    int128 while it fits), cyclic Jacobi on cov/trace in FP32, redone in FP64
//    when the smallest eigen-gap is below 1e-4 of the trace.
//
// Rasterize follows kernel_render's selection and the shared rasteriser
// (kernels/remolod/remolod_render.cu:30-181, kernels/shared/remo_draw.cuh:118-167,
// kernels/shared/remo_framebuffer.cuh:214-248, kernels/shared/remo_math.cuh:16-79):
// per-node frustum + 8-corner screen extent -> large flags; each large inner
// node emits its visible, non-large children and a large visible leaf emits
// itself; one workgroup per draw item splats with a 64-bit atomicMin of
// (float_bits(depth) << 32) | rgba. EDL, wireframe and the GL resolve are
// presentation and not part of this path: the output is the u64 buffer.

#include <cstring>

#include "remolod_analysis.h"
#include "remolod_mlp.h"

namespace remo_hip {

namespace {

constexpr uint32_t kBlock = 256;
using i128 = __int128;
constexpr uint64_t kFbClearValue = 0x7f80000000332211ull;  // remo_framebuffer.cuh:11

__device__ __forceinline__ int bitsU128(unsigned __int128 u) {
    const uint64_t hi = uint64_t(u >> 64), lo = uint64_t(u);
    return hi ? 128 - __clzll((long long)hi) : (lo ? 64 - __clzll((long long)lo) : 0);
}
__device__ __forceinline__ int bitsI128(i128 x) {
    return bitsU128(x < 0 ? (unsigned __int128)(-x) : (unsigned __int128)x);
}

// Cyclic Jacobi on a symmetric 3x3 (a00 a01 a02 a11 a12 a22); eigenvalues ascending.
template <typename T>
__device__ __forceinline__ void jacobi3(const T a[6], T ev[3]) {
    T m[3][3] = {{a[0], a[1], a[2]}, {a[1], a[3], a[4]}, {a[2], a[4], a[5]}};
    for (int sweep = 0; sweep < 10; ++sweep) {
        const T off = m[0][1] * m[0][1] + m[0][2] * m[0][2] + m[1][2] * m[1][2];
        if (!(off > T(0))) break;
        for (int pq = 0; pq < 3; ++pq) {
            const int p = pq == 2 ? 1 : 0;
            const int q = pq == 0 ? 1 : 2;
            const T apq = m[p][q];
            if (apq == T(0)) continue;
            const T theta = (m[q][q] - m[p][p]) / (T(2) * apq);
            const T at = theta < T(0) ? -theta : theta;
            T t;
            if (at > T(1e15))
                t = T(0.5) / theta;
            else
                t = (theta < T(0) ? T(-1) : T(1)) / (at + sqrt(theta * theta + T(1)));
            const T c = T(1) / sqrt(t * t + T(1));
            const T sn = t * c;
            const int r = 3 - p - q;
            const T arp = m[r][p], arq = m[r][q];
            m[p][p] = m[p][p] - t * apq;
            m[q][q] = m[q][q] + t * apq;
            m[p][q] = m[q][p] = T(0);
            const T nrp = c * arp - sn * arq;
            const T nrq = sn * arp + c * arq;
            m[r][p] = m[p][r] = nrp;
            m[r][q] = m[q][r] = nrq;
        }
    }
    ev[0] = m[0][0];
    ev[1] = m[1][1];
    ev[2] = m[2][2];
    T tmp;
    if (ev[0] > ev[1]) { tmp = ev[0]; ev[0] = ev[1]; ev[1] = tmp; }
    if (ev[1] > ev[2]) { tmp = ev[1]; ev[1] = ev[2]; ev[2] = tmp; }
    if (ev[0] > ev[1]) { tmp = ev[0]; ev[0] = ev[1]; ev[1] = tmp; }
}

// ---- roll-up -------------------------------------------------------------------

// Left shift of a possibly negative __int128 (two's complement, no UB).
__device__ __forceinline__ i128 shl(i128 x, int k) { return i128((unsigned __int128)x << k); }

__device__ __forceinline__ void leafSums(const DevTree& t, uint32_t node, uint32_t D, RollupSums& S) {
    const remo_node_accum& a = t.accums[node];
    const int sh = int(D - t.level[node]);
    for (int i = 0; i < 3; ++i) S.s1[i] = shl(i128(a.q1[i]), sh);
    for (int i = 0; i < 6; ++i) S.s2[i] = shl(i128(a.q2[i]), 2 * sh);
    S.n = a.count;
}

// The child offsets are +-h with h = 2^k, so every product is a shift: child c
// contributes S1 += C1 + n o, S2 += C2 + o_a C1_b + C1_a o_b + n o_a o_b to its
// parent, o the offset of c's centre from the parent's.
__device__ __forceinline__ void childContribution(const RollupSums& C, uint32_t ci, int k, RollupSums& S) {
    const int pi[6] = {0, 0, 0, 1, 1, 2}, pj[6] = {0, 1, 2, 1, 2, 2};
    const bool pos[3] = {(ci & 4) != 0, (ci & 2) != 0, (ci & 1) != 0};
    const i128 n = i128(C.n);
    const i128 nh = shl(n, k);
    for (int i = 0; i < 3; ++i) S.s1[i] = C.s1[i] + (pos[i] ? nh : -nh);
    const i128 nn = shl(n, 2 * k);
    for (int q = 0; q < 6; ++q) {
        const int a_ = pi[q], b_ = pj[q];
        const i128 t1 = shl(C.s1[b_], k), t2 = shl(C.s1[a_], k);
        const i128 cross = (pos[a_] ? t1 : -t1) + (pos[b_] ? t2 : -t2);
        S.s2[q] = C.s2[q] + cross + (pos[a_] == pos[b_] ? nn : -nn);
    }
    S.n = C.n;
}

// Sum over the 8 lanes of a group (xor butterfly within width 8).
__device__ __forceinline__ i128 groupSum8(i128 v) {
    for (int m = 1; m < 8; m <<= 1) {
        const unsigned __int128 u = (unsigned __int128)v;
        const unsigned long long lo = __shfl_xor((unsigned long long)uint64_t(u), m, 8);
        const unsigned long long hi = __shfl_xor((unsigned long long)uint64_t(u >> 64), m, 8);
        v += i128(((unsigned __int128)hi << 64) | lo);
    }
    return v;
}

// ---- finalise ------------------------------------------------------------------

struct Bounds {
    float mn[3], mx[3];
};

// nodeBounds (remolod_render.cu:30-38); cubeMin = 0, 2^level is exact.
__device__ __forceinline__ Bounds nodeBounds(const DevTree& t, uint32_t node, float cubeSize) {
    Bounds b;
    const float size = cubeSize / float(1u << t.level[node]);
    b.mn[0] = 0.0f + float(t.nx[node]) * size;
    b.mn[1] = 0.0f + float(t.ny[node]) * size;
    b.mn[2] = 0.0f + float(t.nz[node]) * size;
    for (int i = 0; i < 3; ++i) b.mx[i] = b.mn[i] + size;
    return b;
}

// projectedExtent (remolod_render.cu:40-70).
__device__ __forceinline__ void projectedExtent(const Bounds& b, const remo_mat4& M, float width, float height,
                                                float& dx, float& dy) {
    dx = dy = 0.0f;
    float minX = 1e30f, minY = 1e30f, maxX = -1e30f, maxY = -1e30f;
    bool any = false;
    for (int corner = 0; corner < 8; ++corner) {
        const float x = (corner & 1) ? b.mx[0] : b.mn[0];
        const float y = (corner & 2) ? b.mx[1] : b.mn[1];
        const float z = (corner & 4) ? b.mx[2] : b.mn[2];
        const float cx = M.rows[0][0] * x + M.rows[0][1] * y + M.rows[0][2] * z + M.rows[0][3] * 1.0f;
        const float cy = M.rows[1][0] * x + M.rows[1][1] * y + M.rows[1][2] * z + M.rows[1][3] * 1.0f;
        const float cw = M.rows[3][0] * x + M.rows[3][1] * y + M.rows[3][2] * z + M.rows[3][3] * 1.0f;
        if (cw <= 0.0f) continue;
        const float invW = 1.0f / cw;
        const float sx = (cx * invW * 0.5f + 0.5f) * width;
        const float sy = (cy * invW * 0.5f + 0.5f) * height;
        minX = fminf(minX, sx);
        maxX = fmaxf(maxX, sx);
        minY = fminf(minY, sy);
        maxY = fmaxf(maxY, sy);
        any = true;
    }
    if (!any) return;
    dx = maxX - minX;
    dy = maxY - minY;
}

struct AnalysisArgs {
    RollupSums* rollup;
    uint8_t* closed;
    remo_node_analysis* results;
    uint32_t* queue;
    uint32_t* queueCount;
    const FrameParams* frame;    // uniforms and arrival of this frame (device)
    uint32_t margin;
};

// Closure, covariance, eigen-solve and score of one node from its rolled-up sums.
// With `pred` (the learned score's surface variation and planarity for this
// node), an open node takes the prediction.
__device__ void finalizeNode(const DevTree& t, const AnalysisArgs& aa, const remo_shared_uniforms& u,
                             uint32_t node, const RollupSums& S, bool closed, uint32_t D,
                             const float* pred = nullptr) {
    remo_node_analysis r;
    memset(&r, 0, sizeof(r));
    r.n = S.n;
    float dx, dy;
    projectedExtent(nodeBounds(t, node, t.octreeSize), u.transform_frozen, u.width, u.height, dx, dy);
    r.coverage = fminf(dx, u.width) * fminf(dy, u.height);
    r.state = closed ? 2u : 0u;
    bool scored = false;
    if (pred && !closed && S.n >= 3) {
        // learned score for open nodes; the analytic path is the fallback when the
        // prediction says the node is not planar (surface variation > 0.05)
        const float sv = pred[0], pl = pred[1];
        if (sv <= kLearnedFallbackSv) {
            r.surface_variation = fmaxf(sv, 0.0f);
            r.planarity = fminf(fmaxf(pl, 0.0f), 1.0f);
            r.linearity = 0.0f;
            r.flags |= 2u;
            scored = true;
        }
    }
    if (!scored && S.n >= 3) {
        const uint32_t e = 20 + D - t.level[node];
        const int pi[6] = {0, 0, 0, 1, 1, 2}, pj[6] = {0, 1, 2, 1, 2, 2};
        int maxS2 = 0, maxS1 = 0;
        for (int k = 0; k < 6; ++k) maxS2 = max(maxS2, bitsI128(S.s2[k]));
        for (int k = 0; k < 3; ++k) maxS1 = max(maxS1, bitsI128(S.s1[k]));
        const bool exact = bitsU128((unsigned __int128)S.n) + maxS2 <= 125 && 2 * maxS1 <= 125;
        const double scale = ldexp(1.0, -int(e));
        const double dn = double(S.n);
        double cov[6];
        for (int k = 0; k < 6; ++k) {
            double num;
            if (exact)
                num = i128ToDouble(i128(S.n) * S.s2[k] - S.s1[pi[k]] * S.s1[pj[k]]);
            else
                num = i128ToDouble(S.s2[k]) * dn - i128ToDouble(S.s1[pi[k]]) * i128ToDouble(S.s1[pj[k]]);
            cov[k] = num / (dn * dn) * scale * scale;
        }
        const double trace = cov[0] + cov[3] + cov[5];
        if (trace > 0.0) {
            double a64[6];
            float a32[6];
            for (int k = 0; k < 6; ++k) {
                a64[k] = cov[k] / trace;
                a32[k] = float(a64[k]);
            }
            float ev32[3];
            jacobi3<float>(a32, ev32);
            const float gap = fminf(ev32[1] - ev32[0], ev32[2] - ev32[1]);
            double ev[3];
            if (gap < 1e-4f) {
                double ev64[3];
                jacobi3<double>(a64, ev64);
                for (int i = 0; i < 3; ++i) ev[i] = ev64[i];
                r.flags |= 1u;
            } else {
                for (int i = 0; i < 3; ++i) ev[i] = double(ev32[i]);
            }
            for (int i = 0; i < 3; ++i) r.lambda[i] = float(ev[i] * trace);
            const double sum = ev[0] + ev[1] + ev[2];
            r.surface_variation = sum > 0.0 ? float(ev[0] / sum) : 0.0f;
            r.planarity = ev[2] > 0.0 ? float((ev[1] - ev[0]) / ev[2]) : 0.0f;
            r.linearity = ev[2] > 0.0 ? float((ev[2] - ev[1]) / ev[2]) : 0.0f;
        }
    }
    r.score = 3.0f * r.surface_variation * r.coverage;
    aa.results[node] = r;
    remo_node_accum* a = t.accums + node;
    a->state = r.state;
    a->score = r.score;
}

// Leaves: sums from the fixed-point accumulators, closure from the batch age,
// then their analysis.
// A leaf's rolled-up sums and closure (written for kFinalize, returned for the
// parent's roll-up).
__device__ __forceinline__ uint8_t leafRecord(const DevTree& t, const AnalysisArgs& aa, uint32_t node, uint32_t D,
                                              uint32_t current, RollupSums& S) {
    leafSums(t, node, D, S);
    S.e = 20 + D - t.level[node];
    const remo_node_accum& a = t.accums[node];
    const uint8_t cl = (a.count == 0 || current >= a.last_touched_batch + aa.margin) ? 1 : 0;
    aa.rollup[node] = S;
    aa.closed[node] = cl;
    return cl;
}

// Leaves: sums from the fixed-point accumulators, closure from the batch age.
// (Forming leaf children's sums inside the level launches instead was slower:
// 61 against 53 us per Analysis on the 100M-point tree.)
__global__ __launch_bounds__(kBlock) void kAnalysisLeaves(DevTree t, AnalysisArgs aa) {
    Globals* g = t.g;
    const uint32_t n = g->numNodes;
    const uint32_t D = g->maxLevel;
    const uint32_t current = g->batchletIndex;
    if (blockIdx.x == 0 && threadIdx.x == 0) *aa.queueCount = 0;  // kFinalize fills the queue afterwards
    for (uint32_t node = blockIdx.x * blockDim.x + threadIdx.x; node < n; node += gridDim.x * blockDim.x) {
        if (t.firstChild[node] >= 0) continue;
        RollupSums S;
        leafRecord(t, aa, node, D, current, S);
    }
}

// Inner nodes of one level, bottom-up, one launch per level: the eight children
// (one level deeper, finished by the previous launch) are added by 8 lanes, one
// child each, and summed across the group (exact int128 addition, so the order
// does not matter). Only the exact sums travel up the levels; the eigen-solve and
// score of every node run afterwards in one node-parallel pass (kFinalize).
__global__ __launch_bounds__(kBlock) void kAnalysisLevel(DevTree t, AnalysisArgs aa, uint32_t level) {
    Globals* g = t.g;
    const uint32_t n = g->numNodes;
    const uint32_t D = g->maxLevel;
    const uint32_t sub = threadIdx.x & 7u;
    const uint32_t groups = (gridDim.x * blockDim.x) >> 3;
    // every lane of a group sees the same node, so the group's lanes stay together
    for (uint32_t node = (blockIdx.x * blockDim.x + threadIdx.x) >> 3; node < n; node += groups) {
        if (t.level[node] != level) continue;
        const int32_t fc = t.firstChild[node];
        if (fc < 0) continue;
        const uint32_t child = uint32_t(fc) + sub;
        RollupSums C;
        childContribution(aa.rollup[child], sub, int(18 + D - level), C);
        RollupSums P;
        for (int i = 0; i < 3; ++i) P.s1[i] = groupSum8(C.s1[i]);
        for (int q = 0; q < 6; ++q) P.s2[q] = groupSum8(C.s2[q]);
        P.n = uint64_t(groupSum8(i128(C.n)));
        P.e = 20 + D - level;
        P.pad = 0;
        uint32_t cl = aa.closed[child];
        for (int m = 1; m < 8; m <<= 1) cl &= __shfl_xor(cl, m, 8);
        if (sub == 0) {
            aa.rollup[node] = P;
            aa.closed[node] = uint8_t(cl);
        }
    }
}

// Sum over the lanes 8 apart (one lane of each width-8 group), in every lane.
__device__ __forceinline__ i128 strideSum8(i128 v) {
    for (int m = 8; m < 64; m <<= 1) {
        const unsigned __int128 u = (unsigned __int128)v;
        const unsigned long long lo = __shfl_xor((unsigned long long)uint64_t(u), m);
        const unsigned long long hi = __shfl_xor((unsigned long long)uint64_t(u >> 64), m);
        v += i128(((unsigned __int128)hi << 64) | lo);
    }
    return v;
}

// Two levels in one launch: a wave per inner node of `level`; lane 8c + g adds
// grandchild g of child c, the 8 lanes of a group sum child c (written when the
// child is inner; a leaf child's sums are already final), and lanes 8 apart sum
// the children into the node. Halves the launches of the per-level pass (each
// ~5 us, mostly fixed cost); the integer sums are the same.
__global__ __launch_bounds__(kBlock) void kAnalysisLevel2(DevTree t, AnalysisArgs aa, uint32_t level) {
    Globals* g = t.g;
    const uint32_t n = g->numNodes;
    const uint32_t D = g->maxLevel;
    const uint32_t lane = laneId(), c = lane >> 3, gi = lane & 7u;
    const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t node = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; node < n; node += waves) {
        if (t.level[node] != level) continue;
        const int32_t fc = t.firstChild[node];
        if (fc < 0) continue;
        const uint32_t child = uint32_t(fc) + c;
        const int32_t gfc = t.firstChild[child];  // the same for the 8 lanes of a group
        RollupSums C;
        uint32_t cl;
        if (gfc >= 0) {
            const uint32_t gc = uint32_t(gfc) + gi;
            RollupSums G;
            childContribution(aa.rollup[gc], gi, int(18 + D - (level + 1)), G);
            for (int i = 0; i < 3; ++i) C.s1[i] = groupSum8(G.s1[i]);
            for (int q = 0; q < 6; ++q) C.s2[q] = groupSum8(G.s2[q]);
            C.n = uint64_t(groupSum8(i128(G.n)));
            C.e = 20 + D - (level + 1);
            C.pad = 0;
            cl = aa.closed[gc];
            for (int m = 1; m < 8; m <<= 1) cl &= __shfl_xor(cl, m, 8);
            if (gi == 0) {
                aa.rollup[child] = C;
                aa.closed[child] = uint8_t(cl);
            }
        } else {
            C = aa.rollup[child];
            cl = aa.closed[child];
        }
        RollupSums T;
        childContribution(C, c, int(18 + D - level), T);
        RollupSums P;
        for (int i = 0; i < 3; ++i) P.s1[i] = strideSum8(T.s1[i]);
        for (int q = 0; q < 6; ++q) P.s2[q] = strideSum8(T.s2[q]);
        P.n = uint64_t(strideSum8(i128(T.n)));
        P.e = 20 + D - level;
        P.pad = 0;
        for (int m = 8; m < 64; m <<= 1) cl &= __shfl_xor(cl, m);
        if (lane == 0) {
            aa.rollup[node] = P;
            aa.closed[node] = uint8_t(cl);
        }
    }
}

// Appends the wave's scored nodes to the work queue (one atomic per wave).
__device__ inline void queueScored(const AnalysisArgs& aa, uint32_t node, bool q) {
    const uint64_t m = __ballot(q);
    if (m == 0) return;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t at = 0;
    if (laneId() == uint32_t(leader)) at = atomicAdd(aa.queueCount, uint32_t(__popcll(m)));
    at = __shfl(at, leader);
    if (q) aa.queue[at + __popcll(m & lanesBelow())] = node;
}

// Every node's eigen-solve and score, and the work queue of scored nodes.
__global__ __launch_bounds__(kBlock) void kFinalize(DevTree t, AnalysisArgs aa) {
    Globals* g = t.g;
    const uint32_t n = g->numNodes;
    const uint32_t D = g->maxLevel;
    const remo_shared_uniforms& u = aa.frame->u;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const uint32_t node = base + threadIdx.x;
        bool q = false;
        if (node < n) {
            finalizeNode(t, aa, u, node, aa.rollup[node], aa.closed[node] != 0, D);
            q = aa.results[node].score > 0.0f;
        }
        queueScored(aa, node, q);
    }
}

// kFinalize with the learned score, 16 nodes per wave (the Analysis trees hold
// some 10^4 nodes, so a wave per 16 keeps more of the chip busy than a wave per
// 64 running four MLP tiles in a row): lanes 0..15 write their node's 13 inputs
// into LDS, the wave runs one MLP tile on the matrix cores (the same tile as
// kMlpInfer), and lanes 0..15 finalise their nodes, open nodes taking the
// prediction.
// `arrival` is the batch fraction of the inputs, or negative to read it from the
// frame parameters (a captured graph, whose arguments are fixed at capture).
__global__ __launch_bounds__(kBlock) void kFinalizeLearned(DevTree t, AnalysisArgs aa, const uint16_t* wb,
                                                           float arrival) {
    static_assert(kBlock == kMlpInferThreads, "one MLP wave per finalize wave");
    constexpr uint32_t kPerBlock = 16 * kMlpWavesPerBlock;
    __shared__ MlpInferShared sh;
    __shared__ float pred[kMlpWavesPerBlock][16 * kMlpOut];
    Globals* g = t.g;
    const uint32_t n = g->numNodes;
    if (blockIdx.x * kPerBlock >= n) return;  // the whole block: before the barrier
    const uint32_t D = g->maxLevel;
    const remo_shared_uniforms& u = aa.frame->u;
    if (arrival < 0.0f) arrival = aa.frame->arrival;
    const uint32_t wave = threadIdx.x / 64, lane = laneId();
    float(*act)[kMlpActRow] = sh.act[wave];
    mlpStageWeights(wb, sh.w);
    __syncthreads();
    for (uint32_t base = blockIdx.x * kPerBlock; base < n; base += gridDim.x * kPerBlock) {
        const uint32_t w0 = base + wave * 16;  // this wave's nodes
        const uint32_t node = w0 + lane;
        const bool mine = lane < 16 && node < n;
        if (w0 < n) {
            if (lane < 16) {
                float f[kMlpIn];
                if (node < n) nodeFeatures(t, node, aa.rollup[node], arrival, f);
                for (uint32_t k = 0; k < 16; ++k) act[lane][k] = (node < n && k < kMlpIn) ? f[k] : 0.0f;
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            mlpInferTile(sh.w, act, n, w0, pred[wave]);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // the predictions are in LDS
        }
        bool q = false;
        if (mine) {
            finalizeNode(t, aa, u, node, aa.rollup[node], aa.closed[node] != 0, D, &pred[wave][lane * kMlpOut]);
            q = aa.results[node].score > 0.0f;
        }
        queueScored(aa, node, q);
    }
}

// The learned score's inputs of every node (training pairs are logged from these).
__global__ __launch_bounds__(kBlock) void kFeatures(DevTree t, AnalysisArgs aa, float* out, uint32_t cap) {
    const uint32_t n = min(t.g->numNodes, cap);
    for (uint32_t node = blockIdx.x * blockDim.x + threadIdx.x; node < n; node += gridDim.x * blockDim.x) {
        float f[kMlpIn];
        nodeFeatures(t, node, aa.rollup[node], aa.frame->arrival, f);
        for (uint32_t k = 0; k < kMlpIn; ++k) out[uint64_t(node) * kMlpIn + k] = f[k];
    }
}

// Work queue: every node with a positive score, rebuilt each frame (one atomic per wave).
// ---- rasterize -----------------------------------------------------------------

struct Plane {
    float nx, ny, nz, d;
};

// Frustum::fromViewProj (remo_math.cuh:49-79).
__device__ __forceinline__ void frustumPlanes(const remo_mat4& m, Plane P[6]) {
    const float(*r)[4] = m.rows;
    const float raw[6][4] = {
        {r[3][0] - r[0][0], r[3][1] - r[0][1], r[3][2] - r[0][2], r[3][3] - r[0][3]},
        {r[3][0] + r[0][0], r[3][1] + r[0][1], r[3][2] + r[0][2], r[3][3] + r[0][3]},
        {r[3][0] + r[1][0], r[3][1] + r[1][1], r[3][2] + r[1][2], r[3][3] + r[1][3]},
        {r[3][0] - r[1][0], r[3][1] - r[1][1], r[3][2] - r[1][2], r[3][3] - r[1][3]},
        {r[3][0] - r[2][0], r[3][1] - r[2][1], r[3][2] - r[2][2], r[3][3] - r[2][3]},
        {r[3][0] + r[2][0], r[3][1] + r[2][1], r[3][2] + r[2][2], r[3][3] + r[2][3]},
    };
    for (int i = 0; i < 6; ++i) {
        const float a = raw[i][0], b = raw[i][1], c = raw[i][2], d = raw[i][3];
        const float len = sqrtf(a * a + b * b + c * c);
        const float inv = len > 0.0f ? 1.0f / len : 0.0f;
        P[i] = {a * inv, b * inv, c * inv, d * inv};
    }
}

__device__ __forceinline__ bool intersectsBox(const Plane P[6], const Bounds& b) {
    for (int i = 0; i < 6; ++i) {
        const Plane& p = P[i];
        const float px = p.nx > 0.0f ? b.mx[0] : b.mn[0];
        const float py = p.ny > 0.0f ? b.mx[1] : b.mn[1];
        const float pz = p.nz > 0.0f ? b.mx[2] : b.mn[2];
        if (p.nx * px + p.ny * py + p.nz * pz + p.d < 0.0f) return false;
    }
    return true;
}

struct RenderArgs {
    uint8_t* visible;
    uint8_t* large;
    DrawItem* items;
    uint32_t* counters;  // [0] items, [1] overflow, [2..3] samples, [4] units, [5] unit overflow
    uint32_t capacity;
    RasterUnit* units;
    uint32_t unitCapacity;
    uint32_t tile;       // samples per raster unit
    const FrameParams* frame;
};

// The frame's parameters into device memory (a kernel argument, so the copy is
// stream-ordered and needs no host buffer that outlives the call).
__global__ void kSetFrame(FrameParams* dst, FrameParams value) {
    if (threadIdx.x == 0) *dst = value;
}

__global__ void kFbClear(uint64_t* fb, uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        fb[i] = kFbClearValue;
}

__global__ __launch_bounds__(kBlock) void kRenderFlags(DevTree t, RenderArgs ra) {
    __shared__ Plane P[6];
    const remo_shared_uniforms& u = ra.frame->u;
    if (threadIdx.x == 0) frustumPlanes(u.transform_frozen, P);
    if (blockIdx.x == 0 && threadIdx.x < 8) ra.counters[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t n = min(t.g->numNodes, t.maxNodes);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const Bounds b = nodeBounds(t, i, t.octreeSize);
        const bool hasSamples = t.numPoints[i] > 0 || t.numVoxels[i] > 0;
        ra.visible[i] = (hasSamples && intersectsBox(P, b)) ? 1 : 0;
        float dx, dy;
        projectedExtent(b, u.transform_frozen, u.width, u.height, dx, dy);
        ra.large[i] = (dx > u.lod_pixel_budget || dy > u.lod_pixel_budget) ? 1 : 0;
    }
}

__device__ __forceinline__ void emit(const DevTree& t, const RenderArgs& ra, uint32_t node) {
    const uint32_t np = t.numPoints[node], nv = t.numVoxels[node];
    if (np == 0 && nv == 0) return;
    const uint32_t idx = atomicAdd(&ra.counters[0], 1u);
    if (idx >= ra.capacity) {
        ra.counters[1] = 1;
        return;
    }
    ra.items[idx] = DrawItem{node, np, nv, t.level[node]};
    atomicAdd(reinterpret_cast<unsigned long long*>(ra.counters + 2), (unsigned long long)(np + nv));
    // raster units of at most `tile` samples, so every workgroup gets similar work
    // (every slot below min(counter, capacity) is written, so the raster never reads
    // an unwritten unit; units past the capacity are dropped and flagged)
    const uint32_t up = (np + ra.tile - 1) / ra.tile, uv = (nv + ra.tile - 1) / ra.tile;
    const uint32_t base = atomicAdd(&ra.counters[4], up + uv);
    for (uint32_t k = 0; k < up + uv; ++k) {
        if (base + k >= ra.unitCapacity) {
            ra.counters[5] = 1;
            break;
        }
        const bool vox = k >= up;
        const uint32_t kk = vox ? k - up : k, total = vox ? nv : np;
        ra.units[base + k] = RasterUnit{node, vox ? 1u : 0u, kk * ra.tile, min(ra.tile, total - kk * ra.tile)};
    }
}

// Frontier selection (remolod_render.cu:134-179).
__global__ __launch_bounds__(kBlock) void kRenderEmit(DevTree t, RenderArgs ra) {
    const uint32_t n = min(t.g->numNodes, t.maxNodes);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (ra.large[i] == 0) continue;
        const int32_t fc = t.firstChild[i];
        if (fc >= 0) {
            for (uint32_t c = 0; c < 8; ++c) {
                const uint32_t child = uint32_t(fc) + c;
                if (child >= n) continue;
                if (ra.large[child] != 0) continue;
                if (ra.visible[child] == 0) continue;
                emit(t, ra, child);
            }
        } else if (ra.visible[i] != 0) {
            emit(t, ra, i);
        }
    }
}

__device__ __forceinline__ int32_t f2i_sat(float f) {
    if (!(f == f)) return 0;
    if (f >= 2147483647.0f) return 2147483647;
    if (f <= -2147483648.0f) return int32_t(0x80000000u);
    return int32_t(f);
}

__device__ __forceinline__ uint32_t hashColor(uint64_t key) {
    uint64_t h = (key + 0x9E3779B97F4A7C15ull) * 0xD1B54A32D192ED03ull;
    h ^= h >> 31;
    h *= 0x94D049BB133111EBull;
    h ^= h >> 29;
    const uint32_t r = 96u + ((uint32_t(h) & 0xFFu) * 159u) / 255u;
    const uint32_t gg = 96u + ((uint32_t(h >> 8) & 0xFFu) * 159u) / 255u;
    const uint32_t b = 96u + ((uint32_t(h >> 16) & 0xFFu) * 159u) / 255u;
    return r | (gg << 8) | (b << 16) | (255u << 24);
}

// remoDrawPoint + remoProject + remoFbPack (remo_framebuffer.cuh:33-61, remo_math.cuh:16-39).
__device__ __forceinline__ void drawPoint(uint64_t* fb, const remo_shared_uniforms& u, int32_t w, int32_t h,
                                          const remo_point& p, uint32_t color) {
    const auto& M = u.transform.rows;
    const float cx = M[0][0] * p.x + M[0][1] * p.y + M[0][2] * p.z + M[0][3] * 1.0f;
    const float cy = M[1][0] * p.x + M[1][1] * p.y + M[1][2] * p.z + M[1][3] * 1.0f;
    const float cw = M[3][0] * p.x + M[3][1] * p.y + M[3][2] * p.z + M[3][3] * 1.0f;
    if (cw <= 0.0f) return;
    const float invW = 1.0f / cw;
    const float px = (cx * invW * 0.5f + 0.5f) * u.width;
    const float py = (cy * invW * 0.5f + 0.5f) * u.height;
    const int64_t half = u.point_size / 2;
    const int64_t ix = f2i_sat(px), iy = f2i_sat(py);
    if (ix + half < 0 || iy + half < 0 || ix - half >= w || iy - half >= h) return;
    const uint64_t packed = (uint64_t(__float_as_uint(cw)) << 32) | color;
    for (int64_t oy = -half; oy <= half; ++oy) {
        const int64_t sy = iy + oy;
        if (sy < 0 || sy >= h) continue;
        for (int64_t ox = -half; ox <= half; ++ox) {
            const int64_t sx = ix + ox;
            if (sx < 0 || sx >= w) continue;
            unsigned long long* px64 = reinterpret_cast<unsigned long long*>(fb + sy * w + sx);
            // early depth test: the buffer only decreases, so a stale read is safe
            if (packed < *px64) atomicMin(px64, (unsigned long long)packed);
        }
    }
}

__global__ __launch_bounds__(kBlock) void kRenderRaster(DevTree t, RenderArgs ra, uint64_t* fb) {
    const remo_shared_uniforms& u = ra.frame->u;
    if (u.show_points == 0) return;
    const uint32_t numUnits = min(ra.counters[4], ra.unitCapacity);
    const int32_t w = int32_t(u.width), h = int32_t(u.height);
    const int32_t colorMode = u.color_mode;
    for (uint32_t ui = xcdBlock(); ui < numUnits; ui += gridDim.x) {
        const RasterUnit unit = ra.units[ui];
        uint32_t key = unit.node;
        if (colorMode == 2) key = uint32_t(t.level[unit.node] * 2654435761ull);
        const uint32_t* dir;
        uint32_t dirCap;
        if (unit.pass == 0) {
            dir = t.pointDir + uint64_t(unit.node) * kPointDir;
            dirCap = kPointDir;
        } else {
            const uint32_t gi = t.gridNo[unit.node];
            if (gi == kNone) continue;
            dir = t.voxelDir + uint64_t(gi) * kVoxelDir;
            dirCap = kVoxelDir;
        }
        const uint32_t end = unit.first + unit.count;
        for (uint32_t s = unit.first + threadIdx.x; s < end; s += blockDim.x) {
            const uint32_t k = s / kPointsPerChunk;
            if (k >= dirCap) break;
            const uint32_t ref = dir[k];
            if (ref == 0) continue;
            const remo_point p = chunkPtr(t.arena, ref)[s % kPointsPerChunk];
            uint32_t color = p.color;
            if (colorMode == 3) color = 0xFFFFFFFFu;
            else if (colorMode == 1 || colorMode == 2) color = hashColor(key);
            drawPoint(fb, u, w, h, p, color);
        }
    }
}

// The `flat` pipeline's Rasterize (kernels/flat/flat_render.cu:9-76): every
// point of a resident cloud, no LOD, the same projection and 64-bit splat. The
// reference hands out 65,536-point slices as draw items; a slice is the key of
// the per-item colour modes here too (level 0 for COLOR_LEVEL).
constexpr uint64_t kFlatSlice = 65536;

__global__ __launch_bounds__(kBlock) void kRenderFlat(const remo_point* pts, uint64_t n, remo_shared_uniforms u,
                                                      uint64_t* fb) {
    if (u.show_points == 0) return;
    const int32_t w = int32_t(u.width), h = int32_t(u.height);
    const int32_t colorMode = u.color_mode;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
        const remo_point p = pts[i];
        uint32_t color = p.color;
        if (colorMode == 3) color = 0xFFFFFFFFu;
        else if (colorMode == 1) color = hashColor(i / kFlatSlice);
        else if (colorMode == 2) color = hashColor(0);
        drawPoint(fb, u, w, h, p, color);
    }
}

// Eye-dome lighting on the u64 buffer (remoApplyEDL, remo_framebuffer.cuh:78-122):
// each covered pixel is darkened by exp(-24 * strength * mean over its covered
// 4-neighbours of max(0, log2 d - log2 d_n)), alpha 255; depth is kept, so the
// neighbours' depth reads do not depend on the order pixels are shaded in.
__global__ __launch_bounds__(kBlock) void kApplyEdl(uint64_t* fb, remo_shared_uniforms u) {
    const int32_t w = int32_t(u.width), h = int32_t(u.height);
    const uint64_t n = uint64_t(w) * uint64_t(h);
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
        const int32_t x = int32_t(i % uint64_t(w)), y = int32_t(i / uint64_t(w));
        const uint64_t pixel = fb[i];
        const float depth = __uint_as_float(uint32_t(pixel >> 32));
        if (!(depth < 3.0e38f)) continue;
        const float logDepth = log2f(depth);
        float response = 0.0f;
        int taps = 0;
        const int32_t off[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
        for (int k = 0; k < 4; ++k) {
            const int32_t nx = x + off[k][0], ny = y + off[k][1];
            if (nx < 0 || ny < 0 || nx >= w || ny >= h) continue;
            const float nd = __uint_as_float(uint32_t(fb[uint64_t(ny) * uint64_t(w) + uint64_t(nx)] >> 32));
            if (!(nd < 3.0e38f)) continue;
            response += fmaxf(0.0f, logDepth - log2f(nd));
            ++taps;
        }
        if (taps == 0) continue;
        response /= float(taps);
        const float shade = expf(-response * 24.0f * u.edl_strength);
        const uint32_t c = uint32_t(pixel);
        const uint32_t r = uint32_t(float(c & 0xFFu) * shade);
        const uint32_t g = uint32_t(float((c >> 8) & 0xFFu) * shade);
        const uint32_t b = uint32_t(float((c >> 16) & 0xFFu) * shade);
        fb[i] = (pixel & 0xFFFFFFFF00000000ull) | (r & 0xFFu) | ((g & 0xFFu) << 8) | ((b & 0xFFu) << 16) | (255u << 24);
    }
}

}  // namespace

void launchRenderFlat(const remo_point* pts, uint64_t n, const remo_shared_uniforms& u, uint64_t* fb, uint32_t blocks,
                      hipStream_t s) {
    const uint64_t pixels = uint64_t(u.width) * uint64_t(u.height);
    hipLaunchKernelGGL(kFbClear, dim3(1024), dim3(kBlock), 0, s, fb, pixels);
    if (n) hipLaunchKernelGGL(kRenderFlat, dim3(blocks), dim3(kBlock), 0, s, pts, n, u, fb);
}

void launchApplyEdl(uint64_t* fb, const remo_shared_uniforms& u, hipStream_t s) {
    hipLaunchKernelGGL(kApplyEdl, dim3(1024), dim3(kBlock), 0, s, fb, u);
}

int analysisAllocate(AnalysisState& a, uint64_t maxNodes, const AllocFn& alloc) {
    a.maxNodes = uint32_t(maxNodes);
    void* p = nullptr;
    int rc = REMO_OK;
    auto A = [&](void** dst, uint64_t bytes, const char* what) {
        if (rc == REMO_OK) rc = alloc(dst, bytes, what);
    };
    A(&p, sizeof(remo_node_analysis) * maxNodes, "analysis results");
    a.results = static_cast<remo_node_analysis*>(p);
    A(&p, sizeof(RollupSums) * maxNodes, "roll-up sums");
    a.rollup = static_cast<RollupSums*>(p);
    A(&p, maxNodes, "closure flags");
    a.closed = static_cast<uint8_t*>(p);
    A(&p, 4 * maxNodes, "work queue");
    a.queue = static_cast<uint32_t*>(p);
    A(&p, 16, "work queue count");
    a.queueCount = static_cast<uint32_t*>(p);
    A(&p, maxNodes, "visible flags");
    a.visible = static_cast<uint8_t*>(p);
    A(&p, maxNodes, "large flags");
    a.large = static_cast<uint8_t*>(p);
    A(&p, sizeof(DrawItem) * uint64_t(a.drawCapacity), "draw list");
    a.drawList = static_cast<DrawItem*>(p);
    A(&p, 32, "draw counters");
    a.drawCounters = static_cast<uint32_t*>(p);
    A(&p, sizeof(RasterUnit) * uint64_t(a.unitCapacity), "raster units");
    a.units = static_cast<RasterUnit*>(p);
    A(&p, sizeof(FrameParams), "frame parameters");
    a.frame = static_cast<FrameParams*>(p);
    if (rc == REMO_OK) rc = mlpAllocate(a, alloc);
    return rc;
}

void analysisReset(AnalysisState& a) {
    a.drawItems = 0;
    a.drawListOverflow = 0;
    a.drawSamples = 0;
}

namespace {

// The device copy of the frame's parameters; skipped when it already holds them
// (an earlier write on the same stream). Only the learned score reads the
// arrival fraction, which changes every step: without it, a frame of the same
// camera needs no write (one launch less per step).
void setFrame(AnalysisState& a, const remo_shared_uniforms& u, float arrival, bool needArrival, hipStream_t s) {
    FrameParams f{};
    f.u = u;
    f.arrival = (needArrival || !a.frameSet) ? arrival : a.lastFrame.arrival;
    if (a.frameSet && std::memcmp(&f, &a.lastFrame, sizeof(f)) == 0) return;
    hipLaunchKernelGGL(kSetFrame, dim3(1), dim3(64), 0, s, a.frame, f);
    a.lastFrame = f;
    a.frameSet = true;
}

// The Analysis kernels alone (no host work, so they can be captured).
// The learned score's arrival fraction goes in as an argument when arrival >= 0,
// else it is read from the frame parameters.
void launchAnalysis(AnalysisState& a, const DevTree& t, bool learned, uint32_t maxLevel, uint32_t numNodes,
                    float arrival, hipStream_t s) {
    const uint32_t nb = a.shape.nodeBlocks;
    AnalysisArgs aa{a.rollup, a.closed, a.results, a.queue, a.queueCount, a.frame, a.closureMargin};
    // one launch per level: measured faster than one workgroup walking all levels
    // with barriers (157 us against 101 us per Analysis on a 10,073-node tree)
    hipLaunchKernelGGL(kAnalysisLeaves, dim3(nb), dim3(kBlock), 0, s, t, aa);
    int level = int(maxLevel) - 1;  // the deepest level with inner nodes
    for (; level >= 1; level -= 2)
        hipLaunchKernelGGL(kAnalysisLevel2, dim3(nb), dim3(kBlock), 0, s, t, aa, uint32_t(level - 1));
    if (level == 0) hipLaunchKernelGGL(kAnalysisLevel, dim3(nb), dim3(kBlock), 0, s, t, aa, 0u);
    if (learned)
        hipLaunchKernelGGL(kFinalizeLearned, dim3(nb), dim3(kBlock), 0, s, t, aa, a.mlpWeights, arrival);
    else
        hipLaunchKernelGGL(kFinalize, dim3(nb), dim3(kBlock), 0, s, t, aa);
}

// The Rasterize kernels alone.
void launchRender(AnalysisState& a, const DevTree& t, uint64_t* fb, uint64_t pixels, hipStream_t s) {
    hipLaunchKernelGGL(kFbClear, dim3(1024), dim3(kBlock), 0, s, fb, pixels);
    RenderArgs ra{a.visible, a.large, a.drawList, a.drawCounters, a.drawCapacity,
                  a.units, a.unitCapacity, a.shape.rasterTile, a.frame};
    hipLaunchKernelGGL(kRenderFlags, dim3(a.shape.nodeBlocks), dim3(kBlock), 0, s, t, ra);
    hipLaunchKernelGGL(kRenderEmit, dim3(a.shape.nodeBlocks), dim3(kBlock), 0, s, t, ra);
    hipLaunchKernelGGL(kRenderRaster, dim3(a.shape.rasterBlocks), dim3(kBlock), 0, s, t, ra, fb);
}

}  // namespace

int runAnalysis(AnalysisState& a, const DevTree& t, const remo_shared_uniforms& u, int mode, uint32_t maxLevel,
                uint32_t numNodes, float arrival, hipStream_t s, std::string* err) {
    const bool learned = mode == REMO_ANALYSIS_LEARNED;
    if (learned && !a.mlpLoaded) {
        if (err) *err = "no trained score model loaded (remo_set_score_model)";
        return REMO_ERR_STATE;
    }
    a.arrival = arrival;
    // the arrival fraction, which changes every step, goes to the kernel as an
    // argument: the frame parameters are written only when the camera changes
    setFrame(a, u, arrival, false, s);
    launchAnalysis(a, t, learned, maxLevel, numNodes, arrival, s);
    return REMO_OK;
}

int runFeatures(AnalysisState& a, const DevTree& t, float* out, uint32_t cap, hipStream_t s) {
    // the inputs of the last Analysis: its arrival fraction on the device
    if (a.frameSet) setFrame(a, a.lastFrame.u, a.arrival, true, s);
    AnalysisArgs aa{a.rollup, a.closed, a.results, a.queue, a.queueCount, a.frame, a.closureMargin};
    hipLaunchKernelGGL(kFeatures, dim3(a.shape.nodeBlocks), dim3(kBlock), 0, s, t, aa, out, cap);
    return REMO_OK;
}

int runRender(AnalysisState& a, const DevTree& t, const remo_shared_uniforms& u, uint64_t* fb, hipStream_t s,
              std::string* err) {
    (void)err;
    setFrame(a, u, a.arrival, false, s);
    launchRender(a, t, fb, uint64_t(u.width) * uint64_t(u.height), s);
    return REMO_OK;
}

int runFrame(AnalysisState& a, const DevTree& t, const remo_shared_uniforms& u, int mode, uint32_t maxLevel,
             uint32_t numNodes, float arrival, uint64_t* fb, hipStream_t s, std::string* err) {
    const bool learned = mode == REMO_ANALYSIS_LEARNED;
    if (learned && !a.mlpLoaded) {
        if (err) *err = "no trained score model loaded (remo_set_score_model)";
        return REMO_ERR_STATE;
    }
    const uint64_t pixels = uint64_t(u.width) * uint64_t(u.height);
    const uint64_t key[6] = {maxLevel, uint64_t(mode),
                             reinterpret_cast<uint64_t>(fb), pixels,
                             (uint64_t(a.shape.nodeBlocks) << 32) | a.shape.rasterBlocks, a.shape.rasterTile};
    a.arrival = arrival;
    setFrame(a, u, arrival, learned, s);  // outside the graph: the graph's kernels read it
    if (!a.frameExec || std::memcmp(key, a.frameKey, sizeof(key)) != 0) {
        if (a.frameExec) (void)hipGraphExecDestroy(a.frameExec);
        a.frameExec = nullptr;
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            if (err) *err = "hipStreamBeginCapture failed";
            return REMO_ERR_DEVICE;
        }
        launchAnalysis(a, t, learned, maxLevel, numNodes, -1.0f, s);
        launchRender(a, t, fb, pixels, s);
        if (hipStreamEndCapture(s, &g) != hipSuccess || !g) {
            if (err) *err = "hipStreamEndCapture failed";
            return REMO_ERR_DEVICE;
        }
        const hipError_t e = hipGraphInstantiate(&a.frameExec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            a.frameExec = nullptr;
            if (err) *err = std::string("hipGraphInstantiate failed: ") + hipGetErrorString(e);
            return REMO_ERR_DEVICE;
        }
        std::memcpy(a.frameKey, key, sizeof(key));
    }
    if (hipGraphLaunch(a.frameExec, s) != hipSuccess) {
        if (err) *err = "hipGraphLaunch failed";
        return REMO_ERR_DEVICE;
    }
    return REMO_OK;
}

void analysisRelease(AnalysisState& a) {
    if (a.frameExec) (void)hipGraphExecDestroy(a.frameExec);
    a.frameExec = nullptr;
}

void analysisReadDiagnostics(AnalysisState& a, hipStream_t s) {
    uint32_t c[6] = {0, 0, 0, 0, 0, 0};
    if (hipMemcpyAsync(c, a.drawCounters, sizeof(c), hipMemcpyDeviceToHost, s) != hipSuccess) return;
    if (hipStreamSynchronize(s) != hipSuccess) return;
    a.drawItems = c[0];
    a.drawListOverflow = c[1] | c[5];
    uint64_t samples;
    std::memcpy(&samples, c + 2, 8);
    a.drawSamples = samples;
}

}  // namespace remo_hip

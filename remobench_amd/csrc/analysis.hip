// Analysis and Rasterize: the per-frame passes over a built tree.
//
// Analysis has no reference code (kernels/remolod/remodlod_analysis.cu is empty);
// it follows the spec in README.md:80 and wiki/02_KernelPipeline.md:205-224:
// test closure, roll leaf accumulators up the tree and finalise, score =
// geometric score x screen coverage, rebuild the work queue each frame.
//
//  * Roll-up: the sums are __int128 integers in units U = 2^-(20+D) of the root
//    extent (D the deepest level) in one frame for the whole tree, so a parent's
//    sums are the plain sum of its children's; a node's own (centred) frame is
//    taken, exactly, only where its shape or features are read. The first
//    Analysis after a reset or a deepening of the tree rolls up every node (one
//    launch for the leaves, one per two levels); later ones roll up only what
//    changed since (the leaves Accumulate folded, the nodes that split, their
//    ancestors) in one workgroup.
//  * Shape: covariance in node-extent units from num = n*S2 - S1*S1^T (exact
//    int128 while it fits), cyclic Jacobi on cov/trace in FP32 (eigenvalues of a
//    symmetric matrix move by at most the norm of a perturbation, so FP32 holds
//    them to ~1e-7 of the trace whatever the eigen-gap), solved once per change
//    of a node's sums and cached; the finalize reads the cache.
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

// ---- roll-up -------------------------------------------------------------------

// a * b modulo 2^128 for an unsigned 64-bit factor.
__device__ __forceinline__ i128 mulU64(i128 a, uint64_t b) {
    const unsigned __int128 u = (unsigned __int128)a;
    const uint64_t lo = uint64_t(u), hi = uint64_t(u >> 64);
    const unsigned __int128 plo = (unsigned __int128)lo * b;
    return i128(plo + ((unsigned __int128)(hi * b) << 64));
}

// The sums are kept in one frame for the whole tree: integers in units
// U = 2^-(20+D) of the root extent (D the deepest level), origin at the root's
// minimum corner, so a parent's sums are the plain sum of its children's. Every
// operation is on __int128 modulo 2^128: a record's node-centred form (below)
// fits in 128 bits, so it comes out exact even where a root-frame sum wrapped.
// A node's centre in these units is (2 i + 1) 2^(e-1) per axis, i its integer
// coordinate at its level and e = 20 + D - level (its extent is 2^e units).

// The accumulator fields and node coordinates a leaf's record is made from.
struct LeafIn {
    int64_t q1[3], q2[6];
    uint32_t count, last, level, x, y, z;
};

__device__ __forceinline__ LeafIn loadLeaf(const DevTree& t, uint32_t node) {
    const remo_node_accum& a = t.accums[node];
    LeafIn l;
    for (int i = 0; i < 3; ++i) l.q1[i] = a.q1[i];
    for (int i = 0; i < 6; ++i) l.q2[i] = a.q2[i];
    l.count = a.count;
    l.last = a.last_touched_batch;
    l.level = t.level[node];
    l.x = t.nx[node];
    l.y = t.ny[node];
    l.z = t.nz[node];
    return l;
}

// A leaf's record from its accumulator (leafRecordValue).
__device__ __forceinline__ void leafSumsOf(const LeafIn& a, uint32_t D, RollupSums& S) {
    const LeafFrame F = leafFrame(D, a.level, a.x, a.y, a.z);
    for (int f = 0; f < 3; ++f) S.s1[f] = leafRecordValue(F, a.q1, 0, a.count, f);
    for (int q = 0; q < 6; ++q) S.s2[q] = leafRecordValue(F, a.q1, a.q2[q], a.count, 3 + q);
    S.n = a.count;
    S.e = 20 + D - a.level;
    S.pad = 0;
}

__device__ __forceinline__ void leafSums(const DevTree& t, uint32_t node, uint32_t D, RollupSums& S) {
    leafSumsOf(loadLeaf(t, node), D, S);
}

// A record in the node's own frame (centred on the node, same units): what the
// shape, the learned-score inputs and the dump read. S1 - n c and
// S2_ab - c_a S1_b - S1_a c_b + n c_a c_b, c = odd * 2^(e-1).
__device__ __forceinline__ i128 centredS1(const RollupSums& R, const uint32_t odd[3], int i) {
    return R.s1[i] - shl(mulU32(i128(R.n), odd[i]), int(R.e) - 1);
}
// (S2_ab of the record, S1_a, S1_b, n, the odd factors of axes a and b, e - 1)
__device__ __forceinline__ i128 centredS2Of(i128 s2q, i128 s1a, i128 s1b, unsigned long long n, uint32_t oa,
                                            uint32_t ob, int s) {
    const i128 cross = mulU32(s1b, oa) + mulU32(s1a, ob);
    return s2q - shl(cross, s) + shl(mulU32(mulU32(i128(n), oa), ob), 2 * s);
}
__device__ __forceinline__ i128 centredS2(const RollupSums& R, const uint32_t odd[3], int q) {
    const int a = q < 3 ? 0 : (q < 5 ? 1 : 2);
    const int b = q < 3 ? q : (q < 5 ? q - 2 : 2);
    return centredS2Of(R.s2[q], R.s1[a], R.s1[b], R.n, odd[a], odd[b], int(R.e) - 1);
}

__device__ __forceinline__ i128 int128Zero() { return i128(0); }

// Value i of three or six, for an index known only at run time, as selects
// (an array indexed at run time would go to scratch memory).
template <typename T>
__device__ __forceinline__ T pick3(T v0, T v1, T v2, int i) {
    return i == 0 ? v0 : (i == 1 ? v1 : v2);
}
template <typename T>
__device__ __forceinline__ T pick6(T v0, T v1, T v2, T v3, T v4, T v5, int i) {
    return i < 3 ? pick3(v0, v1, v2, i) : pick3(v3, v4, v5, i - 3);
}

__device__ __forceinline__ RollupSums centredOdd(const RollupSums& R, const uint32_t odd[3]) {
    RollupSums S;
    for (int i = 0; i < 3; ++i) S.s1[i] = centredS1(R, odd, i);
    for (int q = 0; q < 6; ++q) S.s2[q] = centredS2(R, odd, q);
    S.n = R.n;
    S.e = R.e;
    S.pad = 0;
    return S;
}
__device__ __forceinline__ RollupSums centred(const DevTree& t, uint32_t node, const RollupSums& R) {
    const uint32_t odd[3] = {2u * t.nx[node] + 1u, 2u * t.ny[node] + 1u, 2u * t.nz[node] + 1u};
    return centredOdd(R, odd);
}

// Lane exchange within rows of 16 lanes by DPP (a VALU operation, no LDS
// round trip): quad_perm [1,0,3,2] pairs lanes i and i^1, [2,3,0,1] lanes i and
// i^2, row_half_mirror lanes i and 7-i of each group of 8.
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141;

template <int kCtrl>
__device__ __forceinline__ uint32_t dpp32(uint32_t v) {
    return uint32_t(__builtin_amdgcn_update_dpp(0, int(v), kCtrl, 0xF, 0xF, false));
}

template <int kCtrl>
__device__ __forceinline__ i128 dpp128(i128 v) {
    const unsigned __int128 u = (unsigned __int128)v;
    const uint32_t w0 = dpp32<kCtrl>(uint32_t(u)), w1 = dpp32<kCtrl>(uint32_t(u >> 32));
    const uint32_t w2 = dpp32<kCtrl>(uint32_t(u >> 64)), w3 = dpp32<kCtrl>(uint32_t(u >> 96));
    return i128((unsigned __int128)w0 | ((unsigned __int128)w1 << 32) | ((unsigned __int128)w2 << 64) |
                ((unsigned __int128)w3 << 96));
}

// Sum over the 8 lanes of a group (every lane of the group active), in every
// lane: pairs, then quads, then the two quads of the group.
__device__ __forceinline__ i128 groupSum8(i128 v) {
    v += dpp128<kDppXor1>(v);
    v += dpp128<kDppXor2>(v);
    v += dpp128<kDppHalfMirror>(v);
    return v;
}

template <int kCtrl>
__device__ __forceinline__ uint64_t dpp64(uint64_t v) {
    return uint64_t(dpp32<kCtrl>(uint32_t(v))) | (uint64_t(dpp32<kCtrl>(uint32_t(v >> 32))) << 32);
}
__device__ __forceinline__ uint64_t groupSum8U64(uint64_t v) {
    v += dpp64<kDppXor1>(v);
    v += dpp64<kDppXor2>(v);
    v += dpp64<kDppHalfMirror>(v);
    return v;
}

// ---- finalise ------------------------------------------------------------------

struct Bounds {
    float mn[3], mx[3];
};

// nodeBounds (remolod_render.cu:30-38); cubeMin = 0, 2^level is exact.
__device__ __forceinline__ Bounds nodeBoundsOf(uint32_t level, uint32_t x, uint32_t y, uint32_t z, float cubeSize) {
    Bounds b;
    const float size = cubeSize / float(1u << level);
    b.mn[0] = 0.0f + float(x) * size;
    b.mn[1] = 0.0f + float(y) * size;
    b.mn[2] = 0.0f + float(z) * size;
    for (int i = 0; i < 3; ++i) b.mx[i] = b.mn[i] + size;
    return b;
}
__device__ __forceinline__ Bounds nodeBounds(const DevTree& t, uint32_t node, float cubeSize) {
    return nodeBoundsOf(t.level[node], t.nx[node], t.ny[node], t.nz[node], cubeSize);
}

// projectedExtent (remolod_render.cu:40-70).
// One corner's screen position; false when it is behind the eye (w <= 0).
__device__ __forceinline__ bool cornerScreen(const Bounds& b, const remo_mat4& M, float width, float height,
                                             int corner, float& sx, float& sy) {
    const float x = (corner & 1) ? b.mx[0] : b.mn[0];
    const float y = (corner & 2) ? b.mx[1] : b.mn[1];
    const float z = (corner & 4) ? b.mx[2] : b.mn[2];
    const float cx = M.rows[0][0] * x + M.rows[0][1] * y + M.rows[0][2] * z + M.rows[0][3] * 1.0f;
    const float cy = M.rows[1][0] * x + M.rows[1][1] * y + M.rows[1][2] * z + M.rows[1][3] * 1.0f;
    const float cw = M.rows[3][0] * x + M.rows[3][1] * y + M.rows[3][2] * z + M.rows[3][3] * 1.0f;
    if (cw <= 0.0f) return false;
    const float invW = 1.0f / cw;
    sx = (cx * invW * 0.5f + 0.5f) * width;
    sy = (cy * invW * 0.5f + 0.5f) * height;
    return true;
}

__device__ __forceinline__ void projectedExtent(const Bounds& b, const remo_mat4& M, float width, float height,
                                                float& dx, float& dy) {
    dx = dy = 0.0f;
    float minX = 1e30f, minY = 1e30f, maxX = -1e30f, maxY = -1e30f;
    bool any = false;
    for (int corner = 0; corner < 8; ++corner) {
        float sx, sy;
        if (!cornerScreen(b, M, width, height, corner, sx, sy)) continue;
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
    uint32_t* queueCtl;          // the queue's two counters, its parity, the readers counter
    const FrameParams* frame;    // uniforms and arrival of this frame (device)
    uint32_t margin;
    uint32_t* closeAt;
    ShapeCache* shapes;
    uint4* markList;             // marked nodes: (node | level << 24, parent, first child), (x, y, z)
    // device slots of the incremental roll-up when its marked nodes exceed LDS
    RollupSums* slotSums;
    uint32_t* slotNode;
    uint32_t* slotFc;
    uint32_t* slotClose;
    uint32_t* slotChild;
    uint32_t* slotNk;
    uint32_t* slotOf;
    uint32_t ldsSlots;   // marked nodes the roll-up keeps in LDS at most (<= kLdsSlots)
};

// Cyclic Jacobi on a symmetric 3x3 (a00 a01 a02 a11 a12 a22) in FP32 with the
// hardware reciprocal, square root and reciprocal square root (about 1 ulp each;
// a rotation is then orthogonal to within a few ulps, so the eigenvalues stay
// within ~1e-7 of the trace): a rotation costs a few VALU operations instead of
// three correctly rounded divisions and two square roots. Sweeps stop when the
// off-diagonal mass is below 2^-48 of the diagonal's (the eigenvalues then sit
// within an FP32 rounding of their limit); eigenvalues ascending.
__device__ __forceinline__ void jacobi3Fast(const float a[6], float ev[3]) {
    float m[3][3] = {{a[0], a[1], a[2]}, {a[1], a[3], a[4]}, {a[2], a[4], a[5]}};
    for (int sweep = 0; sweep < 10; ++sweep) {
        const float off = m[0][1] * m[0][1] + m[0][2] * m[0][2] + m[1][2] * m[1][2];
        const float dg = m[0][0] * m[0][0] + m[1][1] * m[1][1] + m[2][2] * m[2][2];
        if (!(off > 3.5527137e-15f * dg)) break;
        for (int pq = 0; pq < 3; ++pq) {
            const int p = pq == 2 ? 1 : 0;
            const int q = pq == 0 ? 1 : 2;
            const float apq = m[p][q];
            if (apq == 0.0f) continue;
            const float theta = (m[q][q] - m[p][p]) * 0.5f * __builtin_amdgcn_rcpf(apq);
            const float at = fabsf(theta);
            const float t = at > 1e15f ? 0.5f * __builtin_amdgcn_rcpf(theta)
                                       : copysignf(__builtin_amdgcn_rcpf(at + __builtin_amdgcn_sqrtf(at * at + 1.0f)),
                                                   theta);
            const float c = __builtin_amdgcn_rsqf(t * t + 1.0f);
            const float sn = t * c;
            const int r = 3 - p - q;
            const float arp = m[r][p], arq = m[r][q];
            m[p][p] = m[p][p] - t * apq;
            m[q][q] = m[q][q] + t * apq;
            m[p][q] = m[q][p] = 0.0f;
            const float nrp = c * arp - sn * arq;
            const float nrq = sn * arp + c * arq;
            m[r][p] = m[p][r] = nrp;
            m[r][q] = m[q][r] = nrq;
        }
    }
    ev[0] = m[0][0];
    ev[1] = m[1][1];
    ev[2] = m[2][2];
    float tmp;
    if (ev[0] > ev[1]) { tmp = ev[0]; ev[0] = ev[1]; ev[1] = tmp; }
    if (ev[1] > ev[2]) { tmp = ev[1]; ev[1] = ev[2]; ev[2] = tmp; }
    if (ev[0] > ev[1]) { tmp = ev[0]; ev[0] = ev[1]; ev[1] = tmp; }
}

// The analytic shape of a node from its centred sums: covariance numerators
// num = n*S2 - S1*S1^T (exact int128 while they fit, FP64 otherwise), the
// matrix num / trace(num) in FP32, Jacobi as above; eigenvalues ascending, in
// node-extent units (times trace(num) / n^2 / 4^e), surface variation
// l0/(l0+l1+l2), planarity (l1-l0)/l2, linearity (l2-l1)/l2.
// Whether the numerators are exact in int128: n * S2 and S1 * S1 fit (then
// |S1| < 2^63, a signed 64-bit value).
__device__ __forceinline__ bool shapeExact(unsigned long long n, int maxS2, int maxS1) {
    return bitsU128((unsigned __int128)n) + maxS2 <= 125 && 2 * maxS1 <= 125;
}

// Covariance numerator S2_ab n - S1_a S1_b of centred sums.
__device__ __forceinline__ double shapeNumOf(i128 s1a, i128 s1b, i128 s2k, unsigned long long n, bool exact) {
    if (exact) return i128ToDouble(mulU64(s2k, n) - i128(int64_t(s1a)) * i128(int64_t(s1b)));
    return i128ToDouble(s2k) * double(n) - i128ToDouble(s1a) * i128ToDouble(s1b);
}
// Numerator k (xx xy xz yy yz zz).
__device__ __forceinline__ double shapeNum(const i128 s1[3], i128 s2k, unsigned long long n, int k, bool exact) {
    const int a = k < 3 ? 0 : (k < 5 ? 1 : 2);
    const int b = k < 3 ? k : (k < 5 ? k - 2 : 2);
    return shapeNumOf(s1[a], s1[b], s2k, n, exact);
}

// The shape from the six numerators (sh zeroed, sh.n set, n >= 3).
__device__ __forceinline__ void shapeFinish(const double num[6], unsigned long long n, uint32_t e, ShapeCache& sh) {
    const double dn = double(n);
    const double tn = num[0] + num[3] + num[5];
    if (!(tn > 0.0)) return;
    const double inv = 1.0 / tn;
    float a32[6];
    for (int k = 0; k < 6; ++k) a32[k] = float(num[k] * inv);
    float ev[3];
    jacobi3Fast(a32, ev);
    const double scale = ldexp(1.0, -int(e));
    const double trace = tn / (dn * dn) * scale * scale;
    for (int i = 0; i < 3; ++i) sh.lambda[i] = float(double(ev[i]) * trace);
    const float sum = ev[0] + ev[1] + ev[2];
    sh.sv = sum > 0.0f ? ev[0] * __builtin_amdgcn_rcpf(sum) : 0.0f;
    const float r2 = ev[2] > 0.0f ? __builtin_amdgcn_rcpf(ev[2]) : 0.0f;
    sh.planarity = (ev[1] - ev[0]) * r2;
    sh.linearity = (ev[2] - ev[1]) * r2;
}

__device__ void shapeOf(const RollupSums& S, ShapeCache& sh) {
    memset(&sh, 0, sizeof(sh));
    sh.n = S.n;
    if (S.n < 3) return;
    int maxS2 = 0, maxS1 = 0;
    for (int k = 0; k < 6; ++k) maxS2 = max(maxS2, bitsI128(S.s2[k]));
    for (int k = 0; k < 3; ++k) maxS1 = max(maxS1, bitsI128(S.s1[k]));
    const bool exact = shapeExact(S.n, maxS2, maxS1);
    double num[6];
    for (int k = 0; k < 6; ++k) num[k] = shapeNum(S.s1, S.s2[k], S.n, k, exact);
    shapeFinish(num, S.n, S.e, sh);
}

// Recomputes the cached shape of every node whose record changed since its
// shape was solved (after the full roll-up, which rewrites every record).
__global__ __launch_bounds__(kBlock) void kShapes(DevTree t, AnalysisArgs aa) {
    const uint32_t n = t.g->numNodes;
    for (uint32_t node = blockIdx.x * blockDim.x + threadIdx.x; node < n; node += gridDim.x * blockDim.x) {
        if (!t.shapeStale[node]) continue;
        ShapeCache sh;
        shapeOf(centred(t, node, aa.rollup[node]), sh);
        aa.shapes[node] = sh;
        t.shapeStale[node] = 0;
    }
}

// Closure, shape, coverage and score of one node. With `pred` (the learned
// score's predicted change of this node's surface variation and planarity), an
// open node takes its analytic values plus that change where the change is large
// (above 0.05 in surface variation or 0.1 in planarity); otherwise, and for
// closed nodes, the analytic values of its current sums stand.
// (dx, dy: the node's projected extent.)
__device__ void finalizeExtent(const DevTree& t, const AnalysisArgs& aa, const remo_shared_uniforms& u,
                               uint32_t node, const ShapeCache& sh, bool closed, float dx, float dy,
                               const float* pred = nullptr) {
    remo_node_analysis r;
    memset(&r, 0, sizeof(r));
    r.n = sh.n;
    r.coverage = fminf(dx, u.width) * fminf(dy, u.height);
    r.state = closed ? 2u : 0u;
    for (int i = 0; i < 3; ++i) r.lambda[i] = sh.lambda[i];
    r.surface_variation = sh.sv;
    r.planarity = sh.planarity;
    r.linearity = sh.linearity;
    if (pred && !closed && sh.n >= 3) {
        // the predicted change to the analytic values of the current sums
        const float dsv = pred[0], dpl = pred[1];
        if (fabsf(dsv) > kLearnedChangeSv || fabsf(dpl) > kLearnedChangePl) {
            r.surface_variation = fminf(fmaxf(sh.sv + dsv, 0.0f), 1.0f / 3.0f);
            r.planarity = fminf(fmaxf(sh.planarity + dpl, 0.0f), 1.0f);
            r.flags |= 2u;
        }
    }
    r.score = 3.0f * r.surface_variation * r.coverage;
    aa.results[node] = r;
    aa.closed[node] = closed ? 1 : 0;
    remo_node_accum* a = t.accums + node;
    a->state = r.state;
    a->score = r.score;
}

__device__ __forceinline__ void finalizeNode(const DevTree& t, const AnalysisArgs& aa, const remo_shared_uniforms& u,
                                             uint32_t node, const ShapeCache& sh, bool closed,
                                             const float* pred = nullptr) {
    float dx, dy;
    projectedExtent(nodeBounds(t, node, t.octreeSize), u.transform_frozen, u.width, u.height, dx, dy);
    finalizeExtent(t, aa, u, node, sh, closed, dx, dy, pred);
}

// The batch index from which a leaf counts as closed: untouched for `margin`
// batches, or empty (closed from the start).
__device__ __forceinline__ uint32_t leafCloseAt(const remo_node_accum& a, uint32_t margin) {
    return a.count == 0 ? 0u : a.last_touched_batch + margin;
}

// A leaf's record and closing batch from its accumulator.
__device__ __forceinline__ void leafRecord(const DevTree& t, const AnalysisArgs& aa, uint32_t node, uint32_t D) {
    RollupSums S;
    leafSums(t, node, D, S);
    aa.rollup[node] = S;
    aa.closeAt[node] = leafCloseAt(t.accums[node], aa.margin);
    markRecordChanged(t, node);
}

// The full roll-up (the first Analysis, or one after the tree got deeper, which
// rescales every sum): every leaf, then the inner nodes level by level, then
// every node's shape (kShapes); it empties the dirty list.
// Leaves: sums from the fixed-point accumulators, closing batch from the batch
// age. (Forming leaf children's sums inside the level launches instead was
// slower: 61 against 53 us per Analysis on the 100M-point tree.)
__global__ __launch_bounds__(kBlock) void kAnalysisLeaves(DevTree t, AnalysisArgs aa) {
    Globals* g = t.g;
    const uint32_t n = g->numNodes;
    const uint32_t D = g->maxLevel;
    if (blockIdx.x == 0 && threadIdx.x == 0) *t.dirtyCount = 0;
    for (uint32_t node = blockIdx.x * blockDim.x + threadIdx.x; node < n; node += gridDim.x * blockDim.x) {
        t.dirtyListed[node] = 0;
        if (t.firstChild[node] >= 0) continue;
        leafRecord(t, aa, node, D);
    }
}

// Max over the 8 lanes of a group (every lane of the group active).
__device__ __forceinline__ uint32_t groupMax8(uint32_t v) {
    v = max(v, dpp32<kDppXor1>(v));
    v = max(v, dpp32<kDppXor2>(v));
    return max(v, dpp32<kDppHalfMirror>(v));
}

// Inner nodes of one level, bottom-up, one launch per level: the eight children
// (one level deeper, finished by the previous launch) are added by 8 lanes, one
// child each, and summed across the group (exact int128 addition, so the order
// does not matter). Only the exact sums travel up the levels; the eigen-solve and
// score of every node run afterwards (kShapes, then the Analysis launch).
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
        const RollupSums C = aa.rollup[child];
        RollupSums P;
        for (int i = 0; i < 3; ++i) P.s1[i] = groupSum8(C.s1[i]);
        for (int q = 0; q < 6; ++q) P.s2[q] = groupSum8(C.s2[q]);
        P.n = uint64_t(groupSum8(i128(C.n)));
        P.e = 20 + D - level;
        P.pad = 0;
        const uint32_t ca = groupMax8(aa.closeAt[child]);
        if (sub == 0) {
            aa.rollup[node] = P;
            aa.closeAt[node] = ca;
            markRecordChanged(t, node);
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
        uint32_t ca;
        if (gfc >= 0) {
            const uint32_t gc = uint32_t(gfc) + gi;
            const RollupSums G = aa.rollup[gc];
            for (int i = 0; i < 3; ++i) C.s1[i] = groupSum8(G.s1[i]);
            for (int q = 0; q < 6; ++q) C.s2[q] = groupSum8(G.s2[q]);
            C.n = uint64_t(groupSum8(i128(G.n)));
            C.e = 20 + D - (level + 1);
            C.pad = 0;
            ca = groupMax8(aa.closeAt[gc]);
            if (gi == 0) {
                aa.rollup[child] = C;
                aa.closeAt[child] = ca;
                markRecordChanged(t, child);
            }
        } else {
            C = aa.rollup[child];
            ca = aa.closeAt[child];
        }
        RollupSums P;
        for (int i = 0; i < 3; ++i) P.s1[i] = strideSum8(C.s1[i]);
        for (int q = 0; q < 6; ++q) P.s2[q] = strideSum8(C.s2[q]);
        P.n = uint64_t(strideSum8(i128(C.n)));
        P.e = 20 + D - level;
        P.pad = 0;
        for (int m = 8; m < 64; m <<= 1) ca = max(ca, uint32_t(__shfl_xor(ca, m)));
        if (lane == 0) {
            aa.rollup[node] = P;
            aa.closeAt[node] = ca;
            markRecordChanged(t, node);
        }
    }
}

// Analysis in one launch. Workgroup 0 rolls up what changed since the last
// Analysis; the other workgroups (the sweep) finalise every other node at the
// same time.
//
// The records of leaves are written where their sums change: Accumulate's finish
// writes a folded leaf's record, a split writes its new children's (empty) ones.
// So only inner nodes are rolled up here: the nodes on the dirty list that are
// inner (they split) and every ancestor of a node on the list are marked (an LDS
// bitmap, so each is taken once; ~53 per 1M-point step on the 10k-node tree).
// Each mark is a slot. For every slot at once, eight lanes take its children: a
// child that is not marked adds its record to the slot's partial sum (summed over
// the eight lanes by DPP). A slot's final sums are its partial plus the partials
// of every marked node below it, so each marked node adds its partial to each
// marked ancestor (non-returning LDS atomics on 42-bit limbs, exact modulo 2^128:
// the result equals a full roll-up's). More marked nodes than the LDS holds take
// slots in device memory, placed by level, with a pass per level. Last each
// slot's record is written and its shape, extent and score computed, 4 lanes per
// node (with the learned score, a wave's 16 nodes are one MLP tile).
//
// A sweep workgroup takes ranges of nodes. It marks the inner nodes of its ranges
// the same way (every node on the dirty list climbs to the root), and finalises
// every node of its ranges that is not marked, first solving the shape of a leaf
// whose record changed since its shape was solved.
//
// The dirty list is read by every workgroup, so the last to have read it (a
// counter) empties it. The work queue of scored nodes: the sweep appends to the
// queue's front, on one of two counters (the one the parity word names; the last
// reader clears the other and flips the parity for the next launch), and
// workgroup 0 fills its back segment from the end, counted in LDS.
constexpr uint32_t kRollThreads = 512;
constexpr uint32_t kLevels = kMaxDepth + 1;
constexpr uint32_t kLdsSlots = 384;
static_assert(AnalysisState{}.rollLdsSlots == kLdsSlots, "default LDS slot limit");
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;
static_assert(kLdsSlots <= kRollThreads, "a slot per thread when the LDS slots push to their parents");
constexpr uint32_t kSweepBlocks = 128;
// the work-queue control words: cursors 0 and 1 (the queue's front), the parity,
// the readers counter, and the count of the queue's back segment (the rolled-up
// nodes, written from the end of the queue by workgroup 0) for parity 0 and 1
constexpr uint32_t kQueueParity = 2, kQueueReaders = 3, kQueueBack = 4;

__device__ __forceinline__ void addSums(RollupSums& P, const RollupSums& T) {
    for (int i = 0; i < 3; ++i) P.s1[i] += T.s1[i];
    for (int q = 0; q < 6; ++q) P.s2[q] += T.s2[q];
    P.n += T.n;
}

__device__ __forceinline__ ShapeCache solveShape(const DevTree& t, const AnalysisArgs& aa, uint32_t node,
                                                 const RollupSums& R) {
    ShapeCache sh;
    shapeOf(centred(t, node, R), sh);
    aa.shapes[node] = sh;
    return sh;
}

// The node's shape: the cached one, or solved now when its record changed since.
__device__ __forceinline__ ShapeCache currentShape(const DevTree& t, const AnalysisArgs& aa, uint32_t node) {
    if (!t.shapeStale[node]) return aa.shapes[node];
    const ShapeCache sh = solveShape(t, aa, node, aa.rollup[node]);
    t.shapeStale[node] = 0;
    return sh;
}

// Appends the wave's scored nodes to the work queue (one atomic per wave; the
// queue holds maxNodes entries).
__device__ inline void queueScored(const DevTree& t, const AnalysisArgs& aa, uint32_t* cursor, uint32_t node,
                                   bool q) {
    const uint64_t m = __ballot(q);
    if (m == 0) return;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t at = 0;
    if (laneId() == uint32_t(leader)) at = atomicAdd(cursor, uint32_t(__popcll(m)));
    at = __shfl(at, leader);
    const uint32_t slot = at + __popcll(m & lanesBelow());
    if (q && slot < t.maxNodes) aa.queue[slot] = node;
}

// Empties the dirty list (every workgroup has read it) and hands the work queue's
// other counter, cleared, to the next launch.
__device__ __forceinline__ void resetDirty(const DevTree& t, const AnalysisArgs& aa, uint32_t nd, uint32_t parity) {
    for (uint32_t e = threadIdx.x; e < nd; e += blockDim.x) t.dirtyListed[t.dirtyList[e]] = 0;
    if (threadIdx.x == 0) {
        // (the control words share a line with the live counter: atomics, not stores)
        atomicExch(t.dirtyCount, 0u);
        atomicExch(aa.queueCtl + (parity ^ 1u), 0u);
        atomicExch(aa.queueCtl + kQueueParity, parity ^ 1u);
        atomicExch(aa.queueCtl + kQueueReaders, 0u);
    }
}

// One slot of the device-memory path: its partial plus its marked children's
// final sums (their slots listed first in the slot's eight child links).
__device__ __forceinline__ void rollSlot(uint32_t sl, RollupSums* sSums, uint32_t* sClose, const uint32_t* sChild,
                                         const uint32_t* sNk) {
    RollupSums P = sSums[sl];
    uint32_t ca = sClose[sl];
    const uint32_t nk = sNk[sl];
    for (uint32_t j = 0; j < nk; ++j) {
        const uint32_t cs = sChild[sl * 8 + j];
        addSums(P, sSums[cs]);
        ca = max(ca, sClose[cs]);
    }
    sSums[sl] = P;
    sClose[sl] = ca;
}

// The LDS path keeps a slot's sums as 64-bit accumulators of limbs: each of the
// nine __int128 sums as bits 0..41, 42..83 and 84..127, then n. A marked node
// adds its final sums to its parent's with non-returning LDS atomics; a limb
// accumulator takes up to 2^19 additions of 44-bit limbs without overflow, and
// the sum recombined modulo 2^128 is the exact __int128 sum.
constexpr uint32_t kLimbs = 28;
// a slot's row of limbs padded to an odd number of 8-byte words, so that slots
// 4 apart do not share LDS banks (the ancestor pass adds from up to 384 lanes)
constexpr uint32_t kLimbStride = 29;
constexpr uint64_t kLimbMask = (1ull << 42) - 1ull;

__device__ __forceinline__ void limbsOf(i128 v, uint64_t* out) {
    const unsigned __int128 u = (unsigned __int128)v;
    out[0] = uint64_t(u) & kLimbMask;
    out[1] = uint64_t(u >> 42) & kLimbMask;
    out[2] = uint64_t(u >> 84);
}

__device__ __forceinline__ i128 fromLimbs(const uint64_t* l) {
    return i128((unsigned __int128)l[0] + ((unsigned __int128)l[1] << 42) + ((unsigned __int128)l[2] << 84));
}

// Min / max of floats over the 4 lanes of a quad (every lane active).
__device__ __forceinline__ float quadF(float v, bool mx) {
    const float a = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), kDppXor1, 0xF, 0xF, false));
    v = mx ? fmaxf(v, a) : fminf(v, a);
    const float b = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), kDppXor2, 0xF, 0xF, false));
    return mx ? fmaxf(v, b) : fminf(v, b);
}
__device__ __forceinline__ uint32_t quadMax(uint32_t v) {
    v = max(v, dpp32<kDppXor1>(v));
    return max(v, dpp32<kDppXor2>(v));
}

// shapeOf of a node's root-frame record by the 4 lanes of a quad (each holding
// the record; o* the node's odd factors 2 x + 1): lane j forms numerators j and
// j + 4 (when < 6), the quad shares them, and every lane finishes the solve with
// the same operations as shapeOf, so the result is the same.
__device__ __forceinline__ ShapeCache shapeOfQuad(const RollupSums& R, uint32_t ox, uint32_t oy, uint32_t oz,
                                                  uint32_t sub) {
    ShapeCache sh;
    memset(&sh, 0, sizeof(sh));
    sh.n = R.n;
    if (R.n < 3) return sh;  // uniform over the quad
    const int s = int(R.e) - 1;
    const i128 c0 = R.s1[0] - shl(mulU32(i128(R.n), ox), s);
    const i128 c1 = R.s1[1] - shl(mulU32(i128(R.n), oy), s);
    const i128 c2 = R.s1[2] - shl(mulU32(i128(R.n), oz), s);
    const int maxS1 = max(bitsI128(c0), max(bitsI128(c1), bitsI128(c2)));
    i128 s2[2];
    int bits = 0;
    for (int h = 0; h < 2; ++h) {
        const int q = int(sub) + 4 * h < 6 ? int(sub) + 4 * h : 0;
        const int a = q < 3 ? 0 : (q < 5 ? 1 : 2);
        const int b = q < 3 ? q : (q < 5 ? q - 2 : 2);
        s2[h] = centredS2Of(pick6(R.s2[0], R.s2[1], R.s2[2], R.s2[3], R.s2[4], R.s2[5], q),
                            pick3(R.s1[0], R.s1[1], R.s1[2], a), pick3(R.s1[0], R.s1[1], R.s1[2], b), R.n,
                            pick3(ox, oy, oz, a), pick3(ox, oy, oz, b), s);
        if (int(sub) + 4 * h < 6) bits = max(bits, bitsI128(s2[h]));
    }
    const bool exact = shapeExact(R.n, int(quadMax(uint32_t(bits))), maxS1);
    double mine[2];
    for (int h = 0; h < 2; ++h) {
        const int q = int(sub) + 4 * h < 6 ? int(sub) + 4 * h : 0;
        const int a = q < 3 ? 0 : (q < 5 ? 1 : 2);
        const int b = q < 3 ? q : (q < 5 ? q - 2 : 2);
        mine[h] = shapeNumOf(pick3(c0, c1, c2, a), pick3(c0, c1, c2, b), s2[h], R.n, exact);
    }
    const int base = int(laneId() & ~3u);
    double num[6];
    for (int k = 0; k < 4; ++k) num[k] = __shfl(mine[0], base + k);
    for (int k = 4; k < 6; ++k) num[k] = __shfl(mine[1], base + k - 4);
    shapeFinish(num, R.n, R.e, sh);
    return sh;
}

// The projected extent of a node by the 4 lanes of a quad, two corners each (the
// minima and maxima of projectedExtent over the same values).
__device__ __forceinline__ void extentQuad(const Bounds& bd, const remo_shared_uniforms& u, uint32_t sub, float& dx,
                                           float& dy) {
    float minX = 1e30f, minY = 1e30f, maxX = -1e30f, maxY = -1e30f;
    bool in = false;
    for (int h = 0; h < 2; ++h) {
        float sx, sy;
        if (!cornerScreen(bd, u.transform_frozen, u.width, u.height, int(sub) + 4 * h, sx, sy)) continue;
        minX = fminf(minX, sx);
        maxX = fmaxf(maxX, sx);
        minY = fminf(minY, sy);
        maxY = fmaxf(maxY, sy);
        in = true;
    }
    minX = quadF(minX, false);
    maxX = quadF(maxX, true);
    minY = quadF(minY, false);
    maxY = quadF(maxY, true);
    const bool any = (uint32_t(__ballot(in) >> (laneId() & ~3u)) & 0xFu) != 0u;
    dx = any ? maxX - minX : 0.0f;
    dy = any ? maxY - minY : 0.0f;
}

// The Analysis launch's LDS, one allocation shared by its roles (LDS declared in
// separate functions would be allocated side by side).
struct alignas(16) AnalysisLds {
    union {
        uint32_t markBits[kMarkBits / 32];  // workgroup 0: the marked (rolled-up) inner nodes
        struct {  // then, with the learned score, each wave's MLP tile (the sweep's too)
            MlpWaveShared tile[kRollThreads / 64];
            float pred[kRollThreads / 64][16 * kMlpOut];
        } mlp;
    };
    uint64_t limbs[kLdsSlots][kLimbStride];
    uint32_t node[kLdsSlots], fc[kLdsSlots], close[kLdsSlots], up[kLdsSlots];
    uint4 xyz[kLdsSlots];  // (x, y, z, voxel count)
    uint32_t count[kLevels], start[kLevels], cursor[kLevels];
    uint32_t sweepMarks[64];
    remo_shared_uniforms u;
    uint32_t marked, lastReader, backCount, current;
};
static_assert(sizeof(AnalysisLds::mlp) <= sizeof(AnalysisLds::markBits), "the MLP tiles reuse the bitmap");

// The slot phases over slot arrays in LDS (kLds: up to kLdsSlots marked nodes,
// limb accumulators) or in device memory (__int128 sums, a pass per level):
// placement by level, partials, the sums of marked descendants, then the
// records, shapes and finalize of the marked nodes, 4 lanes per node (with the
// learned score, a wave's 16 nodes are one MLP tile).
template <bool kLearned, bool kLds>
__device__ __forceinline__ void rollSlots(const DevTree& t, const AnalysisArgs& aa, uint32_t D, uint32_t M,
                                          AnalysisLds& L, RollupSums* sSums, uint32_t* sNode, uint32_t* sFc,
                                          uint32_t* sClose, uint32_t* sChild, uint32_t* sNk, uint32_t* queueCursor,
                                          const uint16_t* wb, float arrival) {
    const uint32_t tid = threadIdx.x;
    uint64_t(*sLimbs)[kLimbStride] = L.limbs;
    uint32_t* sUp = L.up;
    const uint32_t* markBits = L.markBits;
    if constexpr (!kLds) {
        // slots by level (the LDS slots are the marks themselves, in marking order)
        if (tid == 0) {
            uint32_t acc = 0;
            for (uint32_t Lv = 0; Lv < kLevels; ++Lv) {
                L.start[Lv] = acc;
                L.cursor[Lv] = acc;
                acc += L.count[Lv];
            }
        }
        __syncthreads();
        for (uint32_t k = tid; k < M; k += kRollThreads) {
            const uint4 v = aa.markList[2 * k];  // (node | level << 24, parent, first child)
            const uint32_t p = v.x & 0xFFFFFFu, Lv = v.x >> 24;
            const uint32_t pos = atomicAdd(&L.cursor[Lv], 1u);
            aa.slotOf[p] = pos;
            sNode[pos] = p;
            sFc[pos] = v.z | (Lv << 24);
        }
        __syncthreads();
    }
    // partials: 8 lanes per slot, one child each
    const uint32_t sub = tid & 7u;
    for (uint32_t grp = tid >> 3; grp < M; grp += kRollThreads / 8) {
        const uint32_t fcw = sFc[grp];
        const uint32_t Lv = fcw >> 24, child = (fcw & 0xFFFFFFu) + sub;
        const bool mk = (markBits[child >> 5] >> (child & 31u)) & 1u;
        const RollupSums R = aa.rollup[child];
        const uint32_t caR = aa.closeAt[child];
        uint32_t slotv = 0, up = kNoSlot;
        if constexpr (kLds) {
            if (sub == 0) {
                const uint32_t par = sUp[grp];
                if (par != kNone) up = aa.slotOf[par];  // every ancestor of a marked node is marked
            }
        } else {
            slotv = aa.slotOf[child];
        }
        RollupSums T;
        for (int i = 0; i < 3; ++i) T.s1[i] = mk ? i128(0) : R.s1[i];
        for (int q = 0; q < 6; ++q) T.s2[q] = mk ? i128(0) : R.s2[q];
        T.n = mk ? 0ull : R.n;
        uint32_t ca = mk ? 0u : caR;
        uint32_t gbits = 0;
        if constexpr (!kLds) {
            // the marked children's slots, packed to the front of the slot's links
            gbits = uint32_t(__ballot(mk) >> (laneId() & ~7u)) & 0xFFu;
            if (mk) sChild[grp * 8 + __popc(gbits & ((1u << sub) - 1u))] = slotv;
        }
        RollupSums P;
        for (int i = 0; i < 3; ++i) P.s1[i] = groupSum8(T.s1[i]);
        for (int q = 0; q < 6; ++q) P.s2[q] = groupSum8(T.s2[q]);
        P.n = groupSum8U64(T.n);
        P.e = 20 + D - Lv;
        P.pad = 0;
        ca = groupMax8(ca);
        if constexpr (kLds) {
            // the partial as limbs: lane j value j (selects, no branch per value),
            // lane 0 also value 8, lane 1 the count
            const i128 mine = pick3(pick3(P.s1[0], P.s1[1], P.s1[2], int(sub)),
                                    pick3(P.s2[0], P.s2[1], P.s2[2], int(sub) - 3),
                                    pick3(P.s2[3], P.s2[4], int128Zero(), int(sub) - 6), int(sub) / 3);
            limbsOf(mine, &sLimbs[grp][3 * sub]);
            if (sub == 0) limbsOf(P.s2[5], &sLimbs[grp][24]);
            if (sub == 1) sLimbs[grp][27] = P.n;
            if (sub == 0) {
                sClose[grp] = ca;
                sUp[grp] = up;
            }
        } else if (sub == 0) {
            sSums[grp] = P;
            sClose[grp] = ca;
            sNk[grp] = __popc(gbits);
        }
    }
    __syncthreads();
    if constexpr (kLds) {
        // a slot's final sums are its partial plus the partials of every marked
        // node below it: each adds its partial to each marked ancestor at once
        // (M <= kRollThreads: a slot per thread; the partials are read first)
        uint64_t own[kLimbs];
        uint32_t caOwn = 0, up = kNoSlot;
        if (tid < M) {
#pragma unroll
            for (uint32_t j = 0; j < kLimbs; ++j) own[j] = sLimbs[tid][j];
            caOwn = sClose[tid];
            up = sUp[tid];
        }
        __syncthreads();
        for (uint32_t a = up; a != kNoSlot; a = sUp[a]) {
#pragma unroll
            for (uint32_t j = 0; j < kLimbs; ++j) atomicAdd((unsigned long long*)&sLimbs[a][j], (unsigned long long)own[j]);
            atomicMax(&sClose[a], caOwn);
        }
    } else {
        for (int Lv = int(D) - 1; Lv >= 0; --Lv) {
            const uint32_t cnt = L.count[Lv];
            if (cnt == 0) continue;  // uniform: the barrier below is skipped by every thread
            for (uint32_t sl = L.start[Lv] + tid; sl < L.start[Lv] + cnt; sl += kRollThreads)
                rollSlot(sl, sSums, sClose, sChild, sNk);
            __syncthreads();
        }
    }
    __syncthreads();  // (the bitmap is free from here: the MLP tiles take its place)
    // the slots' records, shapes and results, 4 lanes per slot: the first 64
    // slots of a pass on waves 0..3 (a wave per SIMD), the next 64 on waves 4..7
    const remo_shared_uniforms& u = L.u;
    const uint32_t wave = tid >> 6, qs = tid & 3u, qi = (tid >> 2) & 15u;
    const uint32_t inPass = (wave >= 4 ? 64u : 0u) + (wave & 3u) * 16u + qi;
    for (uint32_t base = 0; base < M; base += kRollThreads / 4) {
        const uint32_t k = base + inPass;  // the same for the 4 lanes of a quad
        const bool live = k < M;
        uint32_t node = 0, Lv = 0, ca = 0;
        uint4 xyz = make_uint4(0, 0, 0, 0);
        RollupSums S;
        ShapeCache sh;
        float dx = 0.0f, dy = 0.0f;
        if (live) {
            node = sNode[k];
            Lv = sFc[k] >> 24;
            if constexpr (kLds) {
                xyz = L.xyz[k];
                const uint64_t* l = sLimbs[k];
#pragma unroll
                for (int i = 0; i < 3; ++i) S.s1[i] = fromLimbs(l + 3 * i);
#pragma unroll
                for (int i = 0; i < 6; ++i) S.s2[i] = fromLimbs(l + 9 + 3 * i);
                S.n = l[27];
                S.e = 20 + D - Lv;
                S.pad = 0;
            } else {
                xyz = make_uint4(t.nx[node], t.ny[node], t.nz[node], t.numVoxels[node]);
                S = sSums[k];
            }
            ca = sClose[k];
            // the record, a value per lane
            RollupSums* r = aa.rollup + node;
#pragma unroll
            for (int v = 0; v < 9; ++v) {
                if (qs != uint32_t(v & 3)) continue;
                if (v < 3)
                    r->s1[v] = S.s1[v];
                else
                    r->s2[v - 3] = S.s2[v - 3];
            }
            if (qs == 1) {
                r->n = S.n;
                r->e = S.e;
                r->pad = 0;
            }
            sh = shapeOfQuad(S, 2u * xyz.x + 1u, 2u * xyz.y + 1u, 2u * xyz.z + 1u, qs);
            extentQuad(nodeBoundsOf(Lv, xyz.x, xyz.y, xyz.z, t.octreeSize), u, qs, dx, dy);
        }
        const float* pred = nullptr;
        if constexpr (kLearned) {
            // the wave's 16 quads are one MLP tile: quad i's node is row i (the
            // weight fragments are loaded here, in flight beside the inputs)
            auto& tile = L.mlp.tile[wave];
            MlpFrag frag;
            if (__ballot(live) != 0) mlpLoadFrag(wb, frag);
            if (qs == 0) {
                float f[kMlpIn];
                if (live) {
                    const uint32_t odd[3] = {2u * xyz.x + 1u, 2u * xyz.y + 1u, 2u * xyz.z + 1u};
                    nodeFeaturesOf(Lv, true, xyz.w, centredOdd(S, odd), arrival, f);
                }
                for (uint32_t j = 0; j < 16; ++j) tile.in[qi][j] = (live && j < kMlpIn) ? f[j] : 0.0f;
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            if (__ballot(live) != 0) mlpInferTile(frag, tile.in, tile.act, 16, 0, L.mlp.pred[wave]);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // the predictions are in LDS
            pred = &L.mlp.pred[wave][qi * kMlpOut];
        }
        bool q = false;
        if (live && qs == 0) {
            aa.closeAt[node] = ca;
            aa.shapes[node] = sh;
            t.shapeStale[node] = 0;
            finalizeExtent(t, aa, u, node, sh, L.current >= ca, dx, dy, pred);
            q = aa.results[node].score > 0.0f;
        }
        // the queue's back segment, counted in LDS
        const uint64_t m = __ballot(q);
        if (m != 0) {
            const int leader = __ffsll((unsigned long long)m) - 1;
            uint32_t at = 0;
            if (laneId() == uint32_t(leader)) at = atomicAdd(&L.backCount, uint32_t(__popcll(m)));
            at = __shfl(at, leader);
            const uint32_t slot = at + __popcll(m & lanesBelow());
            if (q && slot < t.maxNodes) aa.queue[t.maxNodes - 1u - slot] = node;
        }
    }
}

template <bool kLearned>
__device__ __forceinline__ void rollupBlock(const DevTree& t, const AnalysisArgs& aa, uint32_t nd, uint32_t parity,
                            uint32_t readers, AnalysisLds& L, const uint16_t* wb, float arrival) {
    const uint32_t tid = threadIdx.x;
    const uint32_t D = t.g->maxLevel;
    for (uint32_t k = tid; k < kMarkBits / 128; k += kRollThreads)
        reinterpret_cast<uint4*>(L.markBits)[k] = make_uint4(0, 0, 0, 0);
    if (tid < kLevels) L.count[tid] = 0;
    if (tid == 0) {
        L.marked = 0;
        L.backCount = 0;
    }
    __syncthreads();
    // the frame's uniforms into LDS, by the top threads (which rarely have a node
    // to climb from) beside the marking below: the finalize at the end of this
    // workgroup's path would otherwise wait for them there
    constexpr uint32_t kUWords = sizeof(remo_shared_uniforms) / 4;
    static_assert(sizeof(remo_shared_uniforms) % 4 == 0 && kUWords < kRollThreads, "uniforms copy as words");
    if (tid >= kRollThreads - kUWords) {
        const uint32_t w = tid - (kRollThreads - kUWords);
        reinterpret_cast<uint32_t*>(&L.u)[w] = reinterpret_cast<const uint32_t*>(&aa.frame->u)[w];
        if (w == 0) L.current = t.g->batchletIndex;
    }
    // mark: a node on the list that is inner (it split) is rolled up itself, a
    // leaf from its parent; every ancestor with it. A mark carries the node's
    // level, parent, first child, coordinates and voxel count (loaded beside the
    // climb's loads).
    for (uint32_t e = tid; e < nd; e += kRollThreads) {
        const uint32_t node = t.dirtyList[e];
        const bool inner = t.firstChild[node] >= 0;
        uint32_t lp = t.level[node];
        uint32_t p = node;
        if (!inner) {
            p = t.parent[node];
            --lp;
        }
        while (p != kNone) {
            const uint32_t next = t.parent[p];
            const uint32_t fc = uint32_t(t.firstChild[p]);
            const uint4 xyz = make_uint4(t.nx[p], t.ny[p], t.nz[p], t.numVoxels[p]);
            const uint32_t bit = 1u << (p & 31u);
            if (atomicOr(&L.markBits[p >> 5], bit) & bit) break;  // marked, with its ancestors
            const uint32_t m = atomicAdd(&L.marked, 1u);
            aa.markList[2 * m] = make_uint4(p | (lp << 24), next, fc, 0u);
            aa.markList[2 * m + 1] = xyz;
            aa.slotOf[p] = m;  // the LDS slot (the device path places by level)
            if (m < kLdsSlots) {
                L.node[m] = p;
                L.fc[m] = fc | (lp << 24);
                L.up[m] = next;  // the parent's node; its slot in the partial pass
                L.xyz[m] = xyz;
            }
            atomicAdd(&L.count[lp], 1u);
            p = next;
            --lp;
        }
    }
    __syncthreads();
    // the dirty list is read: count this workgroup (from a wave with no loads in
    // the partial pass below when fewer than 57 nodes are marked)
    uint32_t readBefore = 0;
    if (tid == kRollThreads - 64) readBefore = atomicAdd(aa.queueCtl + kQueueReaders, 1u);
    const uint32_t M = L.marked;
    uint32_t* queueCursor = aa.queueCtl + parity;
    if (M <= min(aa.ldsSlots, kLdsSlots))  // uniform
        rollSlots<kLearned, true>(t, aa, D, M, L, nullptr, L.node, L.fc, L.close, nullptr, nullptr, queueCursor, wb,
                                  arrival);
    else
        rollSlots<kLearned, false>(t, aa, D, M, L, aa.slotSums, aa.slotNode, aa.slotFc, aa.slotClose, aa.slotChild,
                                   aa.slotNk, queueCursor, wb, arrival);
    if (tid == kRollThreads - 64) L.lastReader = readBefore == readers - 1u;
    __syncthreads();
    if (tid == 0) aa.queueCtl[kQueueBack + parity] = min(L.backCount, t.maxNodes);
    if (L.lastReader) resetDirty(t, aa, nd, parity);
}

// A sweep workgroup: its ranges are node ranges b, b + kSweepBlocks, ... of 512
// nodes (a node per thread) or, with the learned score, of 128 (16 per wave, one
// MLP tile: the trees hold some 10^4 nodes, so a tile per wave keeps more of the
// chip busy than four tiles in a row).
template <bool kLearned>
__device__ __forceinline__ void sweepBlock(const DevTree& t, const AnalysisArgs& aa, uint32_t nd, uint32_t parity,
                                           uint32_t readers,
                           uint32_t n, AnalysisLds& L, const uint16_t* wb, float arrival) {
    constexpr uint32_t kRange = kLearned ? 128u : kRollThreads;
    constexpr uint32_t kWords = kRange / 32;
    // ranges a workgroup marks in LDS (pools up to kMarkBits nodes; larger ones
    // take the full roll-up every time, so nothing on them is marked)
    constexpr uint32_t kRanges = (kMarkBits + kRange * kSweepBlocks - 1) / (kRange * kSweepBlocks);
    static_assert(kRanges * kWords <= sizeof(AnalysisLds::sweepMarks) / 4, "the sweep's marks fit");
    uint32_t* marks = L.sweepMarks;
    const uint32_t tid = threadIdx.x;
    const uint32_t b = blockIdx.x - 1u;
    const uint32_t nRanges = (n + kRange - 1) / kRange;
    for (uint32_t k = tid; k < kRanges * kWords; k += kRollThreads) marks[k] = 0;
    __syncthreads();
    // the inner nodes of this workgroup's ranges that workgroup 0 rolls up
    for (uint32_t e = tid; e < nd; e += kRollThreads) {
        const uint32_t node = t.dirtyList[e];
        uint32_t p = t.firstChild[node] >= 0 ? node : t.parent[node];
        while (p != kNone) {
            const uint32_t r = p / kRange, k = r / kSweepBlocks;
            if (r % kSweepBlocks == b && k < kRanges) {
                const uint32_t i = p % kRange;
                atomicOr(&marks[k * kWords + i / 32], 1u << (i & 31u));
            }
            p = t.parent[p];
        }
    }
    __syncthreads();
    const uint32_t current = t.g->batchletIndex;
    const remo_shared_uniforms& u = aa.frame->u;
    uint32_t* queueCursor = aa.queueCtl + parity;
    const uint32_t wave = tid >> 6, lane = laneId();
    for (uint32_t k = 0, r = b; r < nRanges; ++k, r += kSweepBlocks) {
        const uint32_t i = kLearned ? wave * 16 + lane : tid;  // within the range
        const uint32_t node = r * kRange + i;
        const bool mine = (!kLearned || lane < 16) && node < n;
        const bool mk = mine && k < kRanges && ((marks[k * kWords + i / 32] >> (i & 31u)) & 1u);
        const bool fin = mine && !mk;
        ShapeCache sh;
        if (fin) sh = currentShape(t, aa, node);
        const float* pred = nullptr;
        if constexpr (kLearned) {
            auto& tile = L.mlp.tile[wave];
            const uint32_t w0 = r * kRange + wave * 16;  // this wave's nodes
            if (w0 < n) {  // uniform over the wave
                MlpFrag frag;  // (loaded beside the inputs)
                mlpLoadFrag(wb, frag);
                if (lane < 16) {
                    float f[kMlpIn];
                    if (fin) nodeFeatures(t, node, centred(t, node, aa.rollup[node]), arrival, f);
                    for (uint32_t j = 0; j < 16; ++j) tile.in[lane][j] = (fin && j < kMlpIn) ? f[j] : 0.0f;
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                mlpInferTile(frag, tile.in, tile.act, n, w0, L.mlp.pred[wave]);
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // the predictions are in LDS
            }
            pred = &L.mlp.pred[wave][(lane & 15u) * kMlpOut];
        }
        bool q = false;
        if (fin) {
            finalizeNode(t, aa, u, node, sh, current >= aa.closeAt[node], pred);
            q = aa.results[node].score > 0.0f;
        }
        queueScored(t, aa, queueCursor, node, q);
    }
    // counted as a reader only now (it read the list before its climbs), so that
    // the reset falls to a sweep workgroup rather than to workgroup 0, which
    // counts itself after its marking and finishes last
    if (tid == 0) L.lastReader = atomicAdd(aa.queueCtl + kQueueReaders, 1u) == readers - 1u;
    __syncthreads();
    if (L.lastReader) resetDirty(t, aa, nd, parity);
}

// Analysis in one launch: workgroup 0 the roll-up, 1 + kSweepBlocks workgroups
// in all. With the learned score (kLearned) `wb` holds its BF16 parameters and
// `arrival` the batch fraction of its inputs, or a negative value to read it from
// the frame parameters (a captured graph, whose arguments are fixed at capture).
template <bool kLearned>
__global__ __launch_bounds__(kRollThreads) void kAnalysisStep(DevTree t, AnalysisArgs aa, const uint16_t* wb,
                                                              float arrival) {
    __shared__ AnalysisLds L;
    const uint32_t n = t.g->numNodes;
    // (pools beyond the bitmap always take the full roll-up, which empties the list)
    const uint32_t nd = t.maxNodes <= kMarkBits ? min(*t.dirtyCount, t.maxNodes) : 0u;
    const uint32_t parity = aa.queueCtl[kQueueParity];
    if (kLearned && arrival < 0.0f) arrival = aa.frame->arrival;
    // the workgroups that read the dirty list: the roll-up and each sweep
    // workgroup with a range
    constexpr uint32_t kRange = kLearned ? 128u : kRollThreads;
    const uint32_t readers = 1u + min(kSweepBlocks, (n + kRange - 1) / kRange);
    if (blockIdx.x == 0)
        rollupBlock<kLearned>(t, aa, nd, parity, readers, L, wb, arrival);
    else if (blockIdx.x < readers)
        sweepBlock<kLearned>(t, aa, nd, parity, readers, n, L, wb, arrival);
}

// Every node's record in its own frame (the roll-up dump). A record never
// written by a roll-up (e = 0) is copied as it is.
__global__ __launch_bounds__(kBlock) void kCentredRollups(DevTree t, const RollupSums* in, RollupSums* out,
                                                         uint32_t n) {
    for (uint32_t node = blockIdx.x * blockDim.x + threadIdx.x; node < n; node += gridDim.x * blockDim.x) {
        const RollupSums R = in[node];
        out[node] = R.e ? centred(t, node, R) : R;
    }
}

// The learned score's inputs of every node (training pairs are logged from these).
__global__ __launch_bounds__(kBlock) void kFeatures(DevTree t, AnalysisArgs aa, float* out, uint32_t cap) {
    const uint32_t n = min(t.g->numNodes, cap);
    for (uint32_t node = blockIdx.x * blockDim.x + threadIdx.x; node < n; node += gridDim.x * blockDim.x) {
        float f[kMlpIn];
        nodeFeatures(t, node, centred(t, node, aa.rollup[node]), aa.frame->arrival, f);
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
    A(&p, 32, "work queue counters");
    a.queueCtl = static_cast<uint32_t*>(p);
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
    A(&p, 4 * maxNodes, "closing batches");
    a.closeAt = static_cast<uint32_t*>(p);
    A(&p, 4 * maxNodes, "stale shape flags");
    a.shapeStale = static_cast<uint32_t*>(p);
    A(&p, sizeof(ShapeCache) * maxNodes, "shape cache");
    a.shapes = static_cast<ShapeCache*>(p);
    A(&p, 32 * maxNodes, "marked nodes");
    a.markList = static_cast<uint4*>(p);
    const uint64_t slots = maxNodes < kMarkBits ? maxNodes : kMarkBits;
    A(&p, sizeof(RollupSums) * slots, "roll-up slot sums");
    a.slotSums = static_cast<RollupSums*>(p);
    A(&p, 4 * slots, "roll-up slot nodes");
    a.slotNode = static_cast<uint32_t*>(p);
    A(&p, 4 * slots, "roll-up slot children");
    a.slotFc = static_cast<uint32_t*>(p);
    A(&p, 4 * slots, "roll-up slot closing batches");
    a.slotClose = static_cast<uint32_t*>(p);
    A(&p, 32 * slots, "roll-up slot child links");
    a.slotChild = static_cast<uint32_t*>(p);
    A(&p, 4 * slots, "roll-up slot marked children");
    a.slotNk = static_cast<uint32_t*>(p);
    A(&p, 4 * maxNodes, "roll-up slot of each node");
    a.slotOf = static_cast<uint32_t*>(p);
    // (the first Analysis rolls up every node, which marks every record changed)
    if (rc == REMO_OK && (hipMemset(a.queueCtl, 0, 32) != hipSuccess ||
                          hipMemset(a.shapeStale, 0, 4 * maxNodes) != hipSuccess))
        rc = REMO_ERR_DEVICE;
    if (rc == REMO_OK) rc = mlpAllocate(a, alloc);
    return rc;
}

void analysisReset(AnalysisState& a) {
    a.drawItems = 0;
    a.drawListOverflow = 0;
    a.drawSamples = 0;
    a.needFull = true;
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

AnalysisArgs argsOf(const AnalysisState& a) {
    return AnalysisArgs{a.rollup, a.closed, a.results, a.queue, a.queueCtl, a.frame, a.closureMargin,
                        a.closeAt, a.shapes, a.markList,
                        a.slotSums, a.slotNode, a.slotFc, a.slotClose, a.slotChild, a.slotNk, a.slotOf,
                        a.rollLdsSlots};
}

// Whether this Analysis must roll up every node: the first one after a reset,
// one after the tree got deeper (the sums' units are 2^-(20 + depth) of the root
// extent, so every sum rescales), or a node pool beyond the marking bitmap.
bool takeFull(AnalysisState& a, const DevTree& t, uint32_t maxLevel) {
    const bool full = a.needFull || maxLevel != a.lastD || t.maxNodes > kMarkBits;
    a.needFull = false;
    a.lastD = maxLevel;
    return full;
}

// The full roll-up: every leaf, then the inner nodes one launch per two levels
// (faster than one workgroup walking all levels with barriers: 157 against 101 us
// per Analysis on a 10,073-node tree, one launch per level).
void launchFullRollup(AnalysisState& a, const DevTree& t, uint32_t maxLevel, hipStream_t s) {
    const uint32_t nb = a.shape.nodeBlocks;
    const AnalysisArgs aa = argsOf(a);
    hipLaunchKernelGGL(kAnalysisLeaves, dim3(nb), dim3(kBlock), 0, s, t, aa);
    int level = int(maxLevel) - 1;  // the deepest level with inner nodes
    for (; level >= 1; level -= 2)
        hipLaunchKernelGGL(kAnalysisLevel2, dim3(nb), dim3(kBlock), 0, s, t, aa, uint32_t(level - 1));
    if (level == 0) hipLaunchKernelGGL(kAnalysisLevel, dim3(nb), dim3(kBlock), 0, s, t, aa, 0u);
    hipLaunchKernelGGL(kShapes, dim3(nb), dim3(kBlock), 0, s, t, aa);
}

// The per-frame Analysis kernels (no host work, so they can be captured): the
// incremental roll-up of what changed since the last Analysis with the finalize
// of every node in the same launch. The learned score's arrival fraction goes in
// as an argument when arrival >= 0, else it is read from the frame parameters.
void launchAnalysis(AnalysisState& a, const DevTree& t, bool learned, float arrival, hipStream_t s) {
    const AnalysisArgs aa = argsOf(a);
    if (learned)
        hipLaunchKernelGGL(kAnalysisStep<true>, dim3(1 + kSweepBlocks), dim3(kRollThreads), 0, s, t, aa,
                           a.mlpWeights, arrival);
    else
        hipLaunchKernelGGL(kAnalysisStep<false>, dim3(1 + kSweepBlocks), dim3(kRollThreads), 0, s, t, aa,
                           static_cast<const uint16_t*>(nullptr), 0.0f);
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

void launchCentredRollups(const AnalysisState& a, const DevTree& t, RollupSums* out, uint32_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(kCentredRollups, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, t, a.rollup, out, n);
}

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
    (void)numNodes;
    if (takeFull(a, t, maxLevel)) launchFullRollup(a, t, maxLevel, s);
    launchAnalysis(a, t, learned, arrival, s);
    return REMO_OK;
}

int runFeatures(AnalysisState& a, const DevTree& t, float* out, uint32_t cap, hipStream_t s) {
    // the inputs of the last Analysis: its arrival fraction on the device
    if (a.frameSet) setFrame(a, a.lastFrame.u, a.arrival, true, s);
    const AnalysisArgs aa = argsOf(a);
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
    (void)numNodes;
    // the graph holds the incremental roll-up and the finalize pass; a full
    // roll-up, when one is due, runs eagerly before it
    if (takeFull(a, t, maxLevel)) launchFullRollup(a, t, maxLevel, s);
    const uint64_t key[6] = {a.rollLdsSlots, uint64_t(mode),
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
        launchAnalysis(a, t, learned, -1.0f, s);
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

void analysisQueueDiagnostics(const AnalysisState& a, uint32_t* host, hipStream_t s) {
    (void)hipMemcpyAsync(host, a.drawCounters, 6 * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
}

void analysisTakeDiagnostics(AnalysisState& a, const uint32_t* host) {
    a.drawItems = host[0];
    a.drawListOverflow = host[1] | host[5];
    uint64_t samples;
    std::memcpy(&samples, host + 2, 8);
    a.drawSamples = samples;
}

void analysisReadDiagnostics(AnalysisState& a, hipStream_t s) {
    uint32_t c[6] = {0, 0, 0, 0, 0, 0};
    if (hipMemcpyAsync(c, a.drawCounters, sizeof(c), hipMemcpyDeviceToHost, s) != hipSuccess) return;
    if (hipStreamSynchronize(s) != hipSuccess) return;
    analysisTakeDiagnostics(a, c);
}

}  // namespace remo_hip

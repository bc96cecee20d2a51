// Accumulate: fold each leaf's new points into its running sums.
//
// Follows kernel_accumulate / kernel_accum_verify
// (kernels/remolod/remolod_accum.cu:47-233): a leaf whose watermark count is
// below numPoints folds slots [count, numPoints); an inner node holding sums is
// cleared; the Morton watermark of the folded points advances; Σ leaf counts and
// inner-with-sums are verified.
//
// Dirty leaves carry 1k-50k new points each, so one workgroup per leaf (the
// reference's shape) leaves most of the GPU idle behind the largest leaves.
// Here a plan pass cuts every dirty range into tiles of 2,048 points and lists
// the leaves it planned, one persistent launch folds the tiles into partial
// sums, and a finish pass adds each listed leaf's partials in tile order (a
// workgroup per leaf), which keeps the FP64 and FP32 sums deterministic for a
// given point order. The finish also writes the leaf's roll-up record for
// Analysis (its sums in the tree's root frame), so that Analysis rolls up only
// inner nodes. (Finishing a leaf in the workgroup that folds its last tile needs a
// device-scope release and acquire per tile, which write back and invalidate the
// per-XCD L2s here: 61 against 40 us per step.) The plan gathers the verify sums
// (sum of leaf counts, inner nodes holding sums) as it visits every node.
//
// Besides the reference's FP64 geometry sums g[9] and FP32 colour sums c[6], every
// point adds its node-centred fixed-point coordinate q = floor(l * 2^20) - 2^19
// (l the node-local coordinate in [0,1], computed in the reference's FP64 order)
// into int64 sums q1[3], q2[6]. Integer addition is associative, so those sums
// are bit-exact and independent of thread order; they are what Analysis rolls up.

#include "remolod_analysis.h"

namespace remo_hip {

namespace {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kWaves = kBlock / 64;
constexpr uint32_t kSlotsPerThread = AccumWork::kTile / kBlock;
static_assert(AccumWork::kTile % kBlock == 0, "a tile is a whole number of slots per thread");

__device__ __forceinline__ uint64_t mortonSpread(uint32_t v) {
    uint64_t x = uint64_t(v) & 0xFFFFFull;
    x = (x | (x << 32)) & 0x001F00000000FFFFull;
    x = (x | (x << 16)) & 0x001F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

// Morton order of (X, Y, Z) with X the most significant bit of each level,
// without interleaving: the dimension whose coordinates differ in the highest
// bit decides (Chan's most-significant-differing-bit comparison).
__device__ __forceinline__ bool msbLess(uint32_t a, uint32_t b) { return a < b && a < (a ^ b); }
__device__ __forceinline__ bool mortonLess(uint32_t aX, uint32_t aY, uint32_t aZ, uint32_t bX, uint32_t bY,
                                           uint32_t bZ) {
    const uint32_t dx = aX ^ bX, dy = aY ^ bY, dz = aZ ^ bZ;
    uint32_t a = aX, b = bX, d = dx;
    if (msbLess(d, dy)) {
        a = aY;
        b = bY;
        d = dy;
    }
    if (msbLess(d, dz)) {
        a = aZ;
        b = bZ;
    }
    return a < b;
}

template <typename T>
__device__ __forceinline__ T waveSum(T v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    return v;
}

// Wave reductions by DPP (VALU lane exchanges, no LDS round trip), the total in
// lane 63: pairs and quads, the mirrored halves of each row of 8 and of 16 lanes
// (every lane of a row then holds the row's total), row 0's total into row 1 and
// row 2's into row 3 (row_bcast:15), rows 0-1's into rows 2-3 (row_bcast:31). A
// 64-bit value moves as two 32-bit halves. The order of the additions is fixed,
// so FP sums are deterministic.
constexpr int kQuadXor1 = 0xB1, kQuadXor2 = 0x4E, kRowHalfMirror = 0x141, kRowMirror = 0x140;
constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143;

template <int kCtrl, int kRowMask>
__device__ __forceinline__ uint32_t dppU32(uint32_t v) {
    return uint32_t(__builtin_amdgcn_update_dpp(0, int(v), kCtrl, kRowMask, 0xF, false));
}
template <int kCtrl, int kRowMask, typename T>
__device__ __forceinline__ T dppMove(T v) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit values");
    if constexpr (sizeof(T) == 4) {
        const uint32_t r = dppU32<kCtrl, kRowMask>(__builtin_bit_cast(uint32_t, v));
        return __builtin_bit_cast(T, r);
    } else {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        const uint64_t r = uint64_t(dppU32<kCtrl, kRowMask>(uint32_t(u))) |
                           (uint64_t(dppU32<kCtrl, kRowMask>(uint32_t(u >> 32))) << 32);
        return __builtin_bit_cast(T, r);
    }
}
// (rows outside the mask receive 0, the identity of both reductions below)
template <typename T>
__device__ __forceinline__ T waveTotalDpp(T v) {
    v += dppMove<kQuadXor1, 0xF>(v);
    v += dppMove<kQuadXor2, 0xF>(v);
    v += dppMove<kRowHalfMirror, 0xF>(v);
    v += dppMove<kRowMirror, 0xF>(v);
    v += dppMove<kRowBcast15, 0xA>(v);
    v += dppMove<kRowBcast31, 0xC>(v);
    return v;
}
__device__ __forceinline__ uint64_t waveMaxDpp(uint64_t v) {
    uint64_t o;
    o = dppMove<kQuadXor1, 0xF>(v); v = o > v ? o : v;
    o = dppMove<kQuadXor2, 0xF>(v); v = o > v ? o : v;
    o = dppMove<kRowHalfMirror, 0xF>(v); v = o > v ? o : v;
    o = dppMove<kRowMirror, 0xF>(v); v = o > v ? o : v;
    o = dppMove<kRowBcast15, 0xA>(v); v = o > v ? o : v;
    o = dppMove<kRowBcast31, 0xC>(v); v = o > v ? o : v;
    return v;
}

__device__ __forceinline__ uint64_t waveMax(uint64_t v) {
    for (int off = 32; off > 0; off >>= 1) {
        const uint64_t o = __shfl_down(v, off);
        v = o > v ? o : v;
    }
    return v;
}

// Plan: per node, clear an inner node's sums; cut a dirty leaf's new range into
// tiles. A leaf whose tiles do not fit waits for the next launch (the watermark
// makes the pass incremental, plans/03_AccumulatorHook.md:562-573).
__global__ __launch_bounds__(kBlock) void kAccumPlan(DevTree t, AccumWork w, uint32_t parity) {
    Globals* g = t.g;
    const uint32_t n = g->numNodes;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        g->pointsFolded = 0;
        g->leavesFolded = 0;
    }
    // the verify sums (kernel_accum_verify, remolod_accum.cu:214-233) over the
    // state this launch leaves: a leaf's count after its fold, inner nodes' sums
    // after they are cleared
    unsigned long long sumCounts = 0;
    uint32_t innerSums = 0;
    const uint32_t lane = laneId();
    for (uint32_t base0 = blockIdx.x * blockDim.x; base0 < n; base0 += gridDim.x * blockDim.x) {
        const uint32_t node = base0 + threadIdx.x;
        uint32_t from = 0, to = 0, nt = 0;
        if (node < n) {
            remo_node_accum* a = t.accums + node;
            if (t.firstChild[node] >= 0) {
                if (a->count != 0) *a = remo_node_accum{};
                innerSums += a->count != 0 ? 1u : 0u;
            } else {
                from = a->count;
                to = t.numPoints[node];
                if (to > from) nt = (to - from + AccumWork::kTile - 1) / AccumWork::kTile;
            }
        }
        // the wave's tiles and planned leaves reserved with one atomic each (a
        // returning atomic per leaf on these two counters queued at the memory side)
        uint32_t incl = nt;
        for (uint32_t off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(incl, off);
            if (lane >= off) incl += y;
        }
        const uint32_t total = __shfl(incl, 63);
        const uint64_t planning = __ballot(nt != 0);
        uint32_t tileAt = 0, leafAt = 0;
        if (lane == 0 && total != 0) {
            tileAt = atomicAdd(w.numTiles + parity, total);
            leafAt = atomicAdd(w.numPlanned + parity, uint32_t(__popcll(planning)));
        }
        tileAt = __shfl(tileAt, 0);
        leafAt = __shfl(leafAt, 0);
        if (node < n && t.firstChild[node] < 0) {
            uint32_t after = from;
            if (nt != 0) {
                const uint32_t base = tileAt + incl - nt;
                if (uint64_t(base) + nt > w.tileCap) {
                    // no room: mark the reserved slots below the capacity empty,
                    // and the leaf's list entry with no tiles
                    for (uint32_t k = 0; base + k < w.tileCap && k < nt; ++k)
                        w.tiles[base + k] = make_uint4(kNone, 0, 0, 0);
                    w.planned[leafAt + __popcll(planning & lanesBelow())] = make_uint4(kNone, 0, 0, 0);
                } else {
                    for (uint32_t k = 0; k < nt; ++k) {
                        const uint32_t b = from + k * AccumWork::kTile;
                        w.tiles[base + k] = make_uint4(node, b, min(to, b + AccumWork::kTile), k);
                    }
                    w.planned[leafAt + __popcll(planning & lanesBelow())] = make_uint4(node, base, nt, to);
                    after = to;
                }
            }
            sumCounts += after;
        }
    }
    sumCounts = waveSum(sumCounts);
    innerSums = waveSum(innerSums);
    if (laneId() == 0) {
        if (sumCounts) atomicAdd((unsigned long long*)&g->sumLeafCountsAcc, sumCounts);
        if (innerSums) atomicAdd(&g->innerWithSumsAcc, innerSums);
    }
}

// Persistent fold of all tiles into partial sums. Tile i goes to hardware block
// i: a launch holds fewer tiles (~760 on the 100M-point scene) than blocks, and
// hardware blocks go round-robin to the 8 XCDs, so every XCD gets an equal share
// (an XCD-contiguous renumbering would hand all of them to the first three).
// The leaf's partials [base, base + nt) added in tile order into its sums (one
// field per thread, from LDS), its watermark count advanced, the folded counts
// into the globals, the leaf onto Analysis' dirty list: kernel_accumulate's
// per-leaf tail (remolod_accum.cu:190-212); then its roll-up record. The
// partials are loaded in one round by the workgroup (a field's loads one after
// another, tile by tile, were a chain of dependent memory round trips).
constexpr uint32_t kMaxLeafTiles = (kMaxPointsPerNode + AccumWork::kTile - 1) / AccumWork::kTile + 1;
static_assert(sizeof(AccumPartial) % 8 == 0, "partials load as 8-byte words");

// The roles of the workgroup's waves: wave 0 the geometry sums (threads 0..8),
// wave 1 the colour sums (threads 64..69) and the Morton maximum (thread 70),
// wave 2 the leaf's roll-up record (threads 128..137), wave 3 the counters and the
// dirty-list entry (thread 192). Each wave then waits only on its own memory
// operations: the record's stores do not wait behind the sums' stores or the
// watermark's atomic, and the dirty list's chain of atomics runs from the start.
// Waves 0..2 stage the partials.
constexpr uint32_t kColourThread = 64;
constexpr uint32_t kMortonThread = 70;
constexpr uint32_t kRecordThread = 128;
constexpr uint32_t kCountThread = 192;
constexpr uint32_t kStageThreads = 192;
static_assert(kCountThread >= kStageThreads && kCountThread < kBlock, "the count thread stages nothing");

// The folded totals a leaf's record is made from (LDS, per workgroup).
struct LeafTotals {
    int64_t q[9];
    uint32_t count;
};

// `plan`: (leaf, first tile, tiles, count after the fold), as the plan listed it.
__device__ __forceinline__ void finishLeaf(const DevTree& t, const AccumWork& w, uint4 plan, uint32_t batchIndex,
                                           AccumPartial* sp, LeafTotals& tot) {
    const uint32_t node = plan.x, base = plan.y, nt = plan.z;
    const uint32_t nLoad = min(nt, kMaxLeafTiles);
    constexpr uint32_t kVecs = sizeof(AccumPartial) / 8;
    const uint32_t f = threadIdx.x;
    remo_node_accum* a = t.accums + node;
    const AccumPartial* gp = w.partials + base;  // tiles past the LDS copy (leaves over 50,000 points)
    // what the record needs besides the totals, in flight beside the partials
    const uint32_t rf = f - kRecordThread;
    uint32_t lvl = 0, X = 0, Y = 0, Z = 0, D = 0;
    if (rf < 10) {
        lvl = t.level[node];
        X = t.nx[node];
        Y = t.ny[node];
        Z = t.nz[node];
        D = t.g->maxLevel;
    }
    if (f == kCountThread) {
        const uint32_t from = a->count, to = plan.w;
        a->count = to;
        tot.count = to;
        a->last_touched_batch = batchIndex;
        Globals* g = t.g;
        atomicAdd((unsigned long long*)&g->numAccumulated, (unsigned long long)(to - from));
        atomicAdd(&g->pointsFolded, to - from);
        atomicAdd(&g->leavesFolded, 1u);
        listDirty(t, node);  // its sums changed: Analysis rolls its ancestors up again
    }
    const uint2* src = reinterpret_cast<const uint2*>(gp);
    uint2* dst = reinterpret_cast<uint2*>(sp);
    if (f < kStageThreads)
        for (uint32_t v = f; v < nLoad * kVecs; v += kStageThreads) dst[v] = src[v];
    __syncthreads();
    // sums in tile order: the LDS copy, then any tiles past it
    if (f < 9) {
        long long v = 0;
        double d = 0.0;
        for (uint32_t k = 0; k < nLoad; ++k) {
            v += sp[k].q[f];
            d += sp[k].g[f];
        }
        for (uint32_t k = nLoad; k < nt; ++k) {
            v += gp[k].q[f];
            d += gp[k].g[f];
        }
        int64_t total;
        if (f < 3)
            a->q1[f] = total = a->q1[f] + v;
        else
            a->q2[f - 3] = total = a->q2[f - 3] + v;
        tot.q[f] = total;
        a->g[f] += d;
    } else if (f - kColourThread < 6) {
        const uint32_t c = f - kColourThread;
        float v = 0.0f;
        for (uint32_t k = 0; k < nLoad; ++k) v += sp[k].c[c];
        for (uint32_t k = nLoad; k < nt; ++k) v += gp[k].c[c];
        a->c[c] += v;
    } else if (f == kMortonThread) {
        unsigned long long mm = 0;
        for (uint32_t k = 0; k < nLoad; ++k) mm = sp[k].morton > mm ? sp[k].morton : mm;
        for (uint32_t k = nLoad; k < nt; ++k) mm = gp[k].morton > mm ? gp[k].morton : mm;
        if (mm > a->morton_max) a->morton_max = mm;
        if (mm) atomicMax((unsigned long long*)&t.g->mortonWatermark, mm);
    }
    __syncthreads();  // sp is reused for the next leaf; the totals are in LDS
    if (rf < 10) {
        // the leaf's roll-up record from its new totals, a value per thread
        RollupSums* r = t.rollup + node;
        const LeafFrame F = leafFrame(D, lvl, X, Y, Z);
        const int64_t q1[3] = {tot.q[0], tot.q[1], tot.q[2]};
        if (rf < 3) {
            r->s1[rf] = leafRecordValue(F, q1, 0, tot.count, int(rf));
        } else if (rf < 9) {
            r->s2[rf - 3] = leafRecordValue(F, q1, tot.q[rf], tot.count, int(rf));
        } else {
            r->n = tot.count;
            r->e = 20 + D - lvl;
            r->pad = 0;
            t.closeAt[node] = batchIndex + t.closureMargin;  // count > 0: last touched now
            markRecordChanged(t, node);
        }
    }
    __syncthreads();  // tot is reused for the next leaf
}

__global__ __launch_bounds__(kBlock) void kAccumTiles(DevTree t, AccumWork w, uint32_t parity) {
    __shared__ long long sq[9][kWaves];
    __shared__ double sg[9][kWaves];
    __shared__ float sc[6][kWaves];
    __shared__ unsigned long long sm[kWaves];
    const uint32_t lane = laneId(), wave = threadIdx.x >> 6;
    const double invSize = 1.0 / double(t.octreeSize);
    const double fGrid = double(1u << kMaxDepth);
    const uint32_t numTiles = min(w.numTiles[parity], uint32_t(w.tileCap));
    // the next launch plans into the other counter
    if (blockIdx.x == 0 && threadIdx.x == 0) w.numTiles[parity ^ 1u] = 0;

    for (uint32_t ti = blockIdx.x; ti < numTiles; ti += gridDim.x) {
        const uint4 tile = w.tiles[ti];
        const uint32_t node = tile.x;
        if (node == kNone) continue;  // planned past the capacity
        const uint32_t lvl = t.level[node];
        const double levelScale = double(1ull << lvl);
        const double NX = double(t.nx[node]), NY = double(t.ny[node]), NZ = double(t.nz[node]);
        const uint32_t* dir = t.pointDir + uint64_t(node) * kPointDir;

        long long q[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        double lg[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        float lc[6] = {0, 0, 0, 0, 0, 0};
        uint32_t bX = 0, bY = 0, bZ = 0;  // the lane's greatest point in Morton order so far
        // the lane's loads first, then the sums in slot order (the order of one
        // slot at a time, so the FP sums are the same)
        uint32_t ref[kSlotsPerThread];
        remo_point pts[kSlotsPerThread];
        for (uint32_t j = 0; j < kSlotsPerThread; ++j) {
            const uint32_t s = tile.y + threadIdx.x + j * kBlock;
            const uint32_t k = s / kPointsPerChunk;
            ref[j] = (s < tile.z && k < kPointDir) ? dir[k] : 0u;
        }
        for (uint32_t j = 0; j < kSlotsPerThread; ++j) {
            const uint32_t s = tile.y + threadIdx.x + j * kBlock;
            if (ref[j] != 0) pts[j] = chunkPtr(t.arena, ref[j])[s % kPointsPerChunk];
        }
        for (uint32_t j = 0; j < kSlotsPerThread; ++j) {
            if (ref[j] == 0) continue;
            const remo_point p = pts[j];
            // remolod_accum.cu:115-132, octreeMin = 0
            const double lx = double(p.x) * invSize;
            const double ly = double(p.y) * invSize;
            const double lz = double(p.z) * invSize;
            const uint32_t maxCoord = (1u << kMaxDepth) - 1u;
            const uint32_t X = min(uint32_t(fGrid * lx), maxCoord);
            const uint32_t Y = min(uint32_t(fGrid * ly), maxCoord);
            const uint32_t Z = min(uint32_t(fGrid * lz), maxCoord);
            if (mortonLess(bX, bY, bZ, X, Y, Z)) {
                bX = X;
                bY = Y;
                bZ = Z;
            }
            double x = lx * levelScale - NX;
            double y = ly * levelScale - NY;
            double z = lz * levelScale - NZ;
            x = fmin(fmax(x, 0.0), 1.0);
            y = fmin(fmax(y, 0.0), 1.0);
            z = fmin(fmax(z, 0.0), 1.0);
            lg[0] += x; lg[1] += y; lg[2] += z;
            lg[3] += x * x; lg[4] += x * y; lg[5] += x * z;
            lg[6] += y * y; lg[7] += y * z; lg[8] += z * z;
            // x in [0, 1]: floor(x * 2^20) is an exact int32, and so are the
            // centred values; their products need 64 bits
            const int32_t qx = int32_t(floor(x * 1048576.0)) - 524288;
            const int32_t qy = int32_t(floor(y * 1048576.0)) - 524288;
            const int32_t qz = int32_t(floor(z * 1048576.0)) - 524288;
            q[0] += qx; q[1] += qy; q[2] += qz;
            q[3] += (long long)qx * qx; q[4] += (long long)qx * qy; q[5] += (long long)qx * qz;
            q[6] += (long long)qy * qy; q[7] += (long long)qy * qz; q[8] += (long long)qz * qz;
            const float cr = float((p.color >> 0) & 0xFFu) * (1.0f / 255.0f);
            const float cg = float((p.color >> 8) & 0xFFu) * (1.0f / 255.0f);
            const float cb = float((p.color >> 16) & 0xFFu) * (1.0f / 255.0f);
            lc[0] += cr; lc[1] += cg; lc[2] += cb;
            lc[3] += cr * cr; lc[4] += cg * cg; lc[5] += cb * cb;
        }
        for (int i = 0; i < 9; ++i) {
            const long long v = waveTotalDpp(q[i]);
            const double d = waveTotalDpp(lg[i]);
            if (lane == 63) {
                sq[i][wave] = v;
                sg[i][wave] = d;
            }
        }
        for (int i = 0; i < 6; ++i) {
            const float v = waveTotalDpp(lc[i]);
            if (lane == 63) sc[i][wave] = v;
        }
        {
            const uint64_t lm = (mortonSpread(bX) << 2) | (mortonSpread(bY) << 1) | mortonSpread(bZ);
            const uint64_t v = waveMaxDpp(lm);
            if (lane == 63) sm[wave] = v;
        }
        __syncthreads();
        if (threadIdx.x < 25) {
            AccumPartial* out = w.partials + ti;
            const uint32_t f = threadIdx.x;
            if (f < 9) {
                long long v = 0;
                double d = 0.0;
                for (uint32_t wv = 0; wv < kWaves; ++wv) {
                    v += sq[f][wv];
                    d += sg[f][wv];
                }
                out->q[f] = v;
                out->g[f] = d;
            } else if (f < 15) {
                float v = 0.0f;
                for (uint32_t wv = 0; wv < kWaves; ++wv) v += sc[f - 9][wv];
                out->c[f - 9] = v;
            } else if (f == 15) {
                unsigned long long mm = 0;
                for (uint32_t wv = 0; wv < kWaves; ++wv) mm = sm[wv] > mm ? sm[wv] : mm;
                out->morton = mm;
            }
        }
        __syncthreads();
    }
}

// The folded leaves (the plan listed them), a workgroup each: finishLeaf. Block 0
// publishes the plan's verify sums.
__global__ __launch_bounds__(kBlock) void kAccumFinish(DevTree t, AccumWork w, uint32_t batchIndex,
                                                       uint32_t parity) {
    __shared__ AccumPartial sp[kMaxLeafTiles];
    __shared__ LeafTotals tot;
    const uint32_t n = w.numPlanned[parity];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        w.numPlanned[parity ^ 1u] = 0;  // the next launch lists into the other counter
        Globals* g = t.g;
        g->sumLeafCounts = g->sumLeafCountsAcc;
        g->innerWithSums = g->innerWithSumsAcc;
        g->sumLeafCountsAcc = 0;
        g->innerWithSumsAcc = 0;
    }
    for (uint32_t k = blockIdx.x; k < n; k += gridDim.x) {
        const uint4 plan = w.planned[k];
        if (plan.x != kNone) finishLeaf(t, w, plan, batchIndex, sp, tot);  // (uniform: a leaf per workgroup)
    }
}

// The FP32-atomics comparator: every stored point of every leaf adds its
// node-local coordinate (FP32 arithmetic, clamped to [0,1]) into per-node FP32
// sums with one atomicAdd per sum, the naive alternative to the int64 fixed point.
// out[node * 10 + k]: k = 0 count, 1..3 sum x y z, 4..9 sum xx xy xz yy yz zz.
__global__ __launch_bounds__(kBlock) void kAccumFp32Atomics(DevTree t, float* out, uint32_t cap) {
    const uint32_t n = min(t.g->numNodes, cap);
    for (uint32_t node = xcdBlock(); node < n; node += gridDim.x) {
        if (t.firstChild[node] >= 0) continue;
        const uint32_t np = t.numPoints[node];
        const float scale = float(1u << t.level[node]) / t.octreeSize;
        const float NX = float(t.nx[node]), NY = float(t.ny[node]), NZ = float(t.nz[node]);
        const uint32_t* dir = t.pointDir + uint64_t(node) * kPointDir;
        float* o = out + uint64_t(node) * 10;
        for (uint32_t s = threadIdx.x; s < np; s += kBlock) {
            const uint32_t k = s / kPointsPerChunk;
            const uint32_t ref = k < kPointDir ? dir[k] : 0u;
            if (ref == 0) continue;
            const remo_point p = chunkPtr(t.arena, ref)[s % kPointsPerChunk];
            const float x = fminf(fmaxf(p.x * scale - NX, 0.0f), 1.0f);
            const float y = fminf(fmaxf(p.y * scale - NY, 0.0f), 1.0f);
            const float z = fminf(fmaxf(p.z * scale - NZ, 0.0f), 1.0f);
            atomicAdd(o + 0, 1.0f);
            atomicAdd(o + 1, x);
            atomicAdd(o + 2, y);
            atomicAdd(o + 3, z);
            atomicAdd(o + 4, x * x);
            atomicAdd(o + 5, x * y);
            atomicAdd(o + 6, x * z);
            atomicAdd(o + 7, y * y);
            atomicAdd(o + 8, y * z);
            atomicAdd(o + 9, z * z);
        }
    }
}

__global__ void kZeroF32(float* p, uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = 0.0f;
}

}  // namespace

void launchAccumFp32Atomics(const DevTree& t, float* out, uint32_t cap, const LaunchShape& ls, hipStream_t s) {
    hipLaunchKernelGGL(kZeroF32, dim3(256), dim3(kBlock), 0, s, out, uint64_t(cap) * 10);
    hipLaunchKernelGGL(kAccumFp32Atomics, dim3(ls.pointBlocks), dim3(kBlock), 0, s, t, out, cap);
}

void launchAccumulate(const DevTree& t, const AccumWork& w, uint32_t batchIndex, uint32_t parity,
                      const LaunchShape& ls, hipStream_t s) {
    hipLaunchKernelGGL(kAccumPlan, dim3(ls.nodeBlocks), dim3(kBlock), 0, s, t, w, parity);
    hipLaunchKernelGGL(kAccumTiles, dim3(ls.pointBlocks), dim3(kBlock), 0, s, t, w, parity);
    hipLaunchKernelGGL(kAccumFinish, dim3(ls.nodeBlocks), dim3(kBlock), 0, s, t, w, batchIndex, parity);
}

}  // namespace remo_hip

// Refinement, colour filtering: every gridded node's voxel list rewritten in
// cell order with colours averaged over all the points of the node's subtree
// that fall in each cell.
//
// The reference specifies Refinement without code (README.md:81,
// wiki/02_KernelPipeline.md:228-242); its constraint for colour filtering
// (CLAUDE.md:181-184) is that a dense RGBA grid per node (~16 MB) is too big, so
// the sums must be sparse and keyed by the voxels. Here the key of a cell is its
// rank among the set bits of the node's 128^3 occupancy grid (x fastest, then y,
// then z): a per-grid prefix of set-bit counts over 8-word blocks turns a cell
// into its rank with one 32-byte load, and the sums are a dense array of one
// entry per occupied cell, three u64 each (r + g<<32, b + a<<32, n).
//
// A gridded node's cells are exactly its subtree's points' cells (every point
// sets its cell at every gridded ancestor; a node's grid is zeroed only when it
// splits, when all its region's points pass through it again), so its sums are
//   * the points of its leaf children (or its own points, for a root that is
//     still a leaf with a grid), each added at its cell's rank, and
//   * its inner children's sums: child cell (x', y', z') of octant (ox, oy, oz)
//     is parent cell (64 ox + x'/2, 64 oy + y'/2, 64 oz + z'/2), because both are
//     bit fields of the same 28-bit quantised coordinate.
// Levels run bottom-up, one launch of each phase per level, so a pass reads every
// point once and every child sum once. Sums persist per grid between calls; a
// node is recomputed only when a leaf below it gained points or its voxel count
// changed (dirty flags walk up the parent chain).
//
// Emission rewrites the voxel chunk list: the cell of rank k goes to slot k, at
// the cell centre kVoxelSample computes, with colour (sum + n/2) / n per channel.

#include <algorithm>
#include <chrono>
#include <cstring>

#include "remolod_refine.h"

namespace remo_hip {

namespace {

constexpr uint32_t kBlock = 256;
constexpr unsigned long long kCellPointLimit = 16843009ull;  // 255 n < 2^32: no carry between halves

struct RefineArgs {
    uint32_t* refPoints;
    uint32_t* refVoxels;
    uint32_t* dirty;
    uint32_t* status;
    uint32_t* cells;
    uint32_t* list;
    uint32_t* kids;
    uint32_t* prefix;
    unsigned long long* accOff;
    uint32_t* accCap;
    unsigned long long* acc;
    unsigned long long accEntries;
    RefineCounters* ctr;
};

RefineArgs argsOf(const RefineState& r) {
    return RefineArgs{r.refPoints, r.refVoxels, r.dirty, r.status, r.cells, r.list, r.kids, r.prefix,
                      r.accOff, r.accCap, r.acc, r.accEntries, r.ctr};
}

__device__ __forceinline__ const uint32_t* gridWords(const DevTree& t, uint32_t g) {
    return reinterpret_cast<const uint32_t*>(t.arena + t.gridOffset[g]);
}

// The 8 words of rank block `blk` (32 bytes; grids are 16-byte aligned).
__device__ __forceinline__ void loadBlock(const uint32_t* words, uint32_t blk, uint32_t w[8]) {
    const uint4* p = reinterpret_cast<const uint4*>(words) + 2 * blk;
    const uint4 a = p[0], b = p[1];
    w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w;
    w[4] = b.x; w[5] = b.y; w[6] = b.z; w[7] = b.w;
}

// Rank of cell vi among the grid's set bits; false when its bit is clear.
__device__ __forceinline__ bool rankOf(const uint32_t* words, const uint32_t* prefix, uint32_t vi,
                                       uint32_t& rank) {
    const uint32_t wi = vi >> 5, blk = wi >> 3, j = wi & 7u;
    uint32_t w[8];
    loadBlock(words, blk, w);
    uint32_t r = prefix[blk], word = 0;
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) {
        if (k < j) r += __popc(w[k]);
        if (k == j) word = w[k];
    }
    const uint32_t bit = 1u << (vi & 31u);
    rank = r + __popc(word & (bit - 1u));
    return (word & bit) != 0;
}

__device__ __forceinline__ unsigned long long packRG(uint32_t c) {
    return (unsigned long long)(c & 0xFFu) | ((unsigned long long)((c >> 8) & 0xFFu) << 32);
}
__device__ __forceinline__ unsigned long long packBA(uint32_t c) {
    return (unsigned long long)((c >> 16) & 0xFFu) | ((unsigned long long)(c >> 24) << 32);
}

__device__ __forceinline__ void addSums(unsigned long long* e, unsigned long long a, unsigned long long b,
                                        unsigned long long n) {
    atomicAdd(e, a);
    atomicAdd(e + 1, b);
    atomicAdd(e + 2, n);
}

// Exclusive prefix sum over the 256 threads of a workgroup; `total` is the sum.
__device__ __forceinline__ uint32_t blockScan(uint32_t v, uint32_t* lds4, uint32_t& total) {
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) lds4[wave] = x;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (uint32_t w = 0; w < kBlock / 64; ++w) {
        const uint32_t s = lds4[w];
        if (w < wave) base += s;
        tot += s;
    }
    __syncthreads();  // lds4 is rewritten by the next scan
    total = tot;
    return base + x - v;
}

// ---- per call: flags, dirty marking, per-level work ---------------------------

__global__ __launch_bounds__(kBlock) void kRefineClear(RefineArgs r, uint32_t numNodes) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < numNodes; i += gridDim.x * blockDim.x) {
        r.dirty[i] = 0;
        r.status[i] = 0;
    }
    if (blockIdx.x == 0) {
        RefineCounters* c = r.ctr;
        const unsigned long long used = c->arenaUsed;
        __syncthreads();  // every thread has read arenaUsed before it is rewritten
        uint32_t* w = reinterpret_cast<uint32_t*>(c);
        for (uint32_t k = threadIdx.x; k < sizeof(RefineCounters) / 4; k += blockDim.x) w[k] = 0;
        __syncthreads();
        if (threadIdx.x == 0) c->arenaUsed = used;
    }
}

// A gridded node whose voxel count differs from what its last refinement wrote
// (or whose sums are invalid), and the parent of every leaf whose point count
// changed, are dirty with all their ancestors.
__global__ __launch_bounds__(kBlock) void kRefineMark(DevTree t, RefineArgs r, uint32_t numNodes) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < numNodes; i += gridDim.x * blockDim.x) {
        const bool gridded = t.gridNo[i] != kNone;
        const bool leaf = t.firstChild[i] < 0;
        const bool selfDirty = gridded && r.refVoxels[i] != t.numVoxels[i];
        const bool srcDirty = leaf && t.numPoints[i] != r.refPoints[i];
        uint32_t x = kNone;
        if (selfDirty || (srcDirty && gridded)) x = i;
        else if (srcDirty) x = t.parent[i];
        while (x != kNone) {
            if (atomicExch(&r.dirty[x], 1u) != 0u) break;  // the rest of the chain is marked
            x = t.parent[x];
        }
    }
}

__global__ __launch_bounds__(kBlock) void kRefineCount(DevTree t, RefineArgs r, uint32_t numNodes) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < numNodes; i += gridDim.x * blockDim.x) {
        if (!r.dirty[i] || t.gridNo[i] == kNone) continue;
        const uint32_t L = min(t.level[i], uint32_t(kMaxDepth));
        const int32_t fc = t.firstChild[i];
        unsigned long long pts = 0;
        uint32_t kids = 0;
        if (fc < 0) {
            pts = t.numPoints[i];
        } else {
            for (uint32_t o = 0; o < 8; ++o) {
                const uint32_t c = uint32_t(fc) + o;
                if (t.firstChild[c] < 0) pts += t.numPoints[c];
                else ++kids;
            }
        }
        atomicAdd(&r.ctr->levelCount[L], 1u);
        atomicAdd(&r.ctr->levelPoints[L], pts);
        atomicAdd(&r.ctr->levelVoxels[L], (unsigned long long)t.numVoxels[i]);
        if (kids) atomicAdd(&r.ctr->levelKids[L], kids);
    }
}

__global__ __launch_bounds__(kBlock) void kRefineList(DevTree t, RefineArgs r, uint32_t numNodes,
                                                      RefineLevelOffsets lo) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < numNodes; i += gridDim.x * blockDim.x) {
        if (!r.dirty[i] || t.gridNo[i] == kNone) continue;
        const uint32_t L = min(t.level[i], uint32_t(kMaxDepth));
        if (L < lo.lo || L > lo.hi) continue;
        r.list[lo.off[L] + atomicAdd(&r.ctr->levelFill[L], 1u)] = i;
    }
}

// ---- per level ----------------------------------------------------------------

// One workgroup per node: the rank prefix of its grid, the soundness checks, its
// accumulator entries (reserved or re-used) zeroed.
__global__ __launch_bounds__(kBlock) void kRefinePrep(DevTree t, RefineArgs r, uint32_t first, uint32_t count,
                                                      uint32_t level, uint32_t kidFirst) {
    __shared__ uint32_t lds4[kBlock / 64];
    __shared__ uint32_t okShared, dirBad;
    __shared__ unsigned long long offShared;
    for (uint32_t k = blockIdx.x; k < count; k += gridDim.x) {
        const uint32_t node = r.list[first + k];
        const uint32_t g = t.gridNo[node];
        const uint32_t* words = gridWords(t, g);
        uint32_t* prefix = r.prefix + uint64_t(g) * kRefineBlocks;
        if (threadIdx.x == 0) dirBad = 0;
        uint32_t run = 0;
        for (uint32_t round = 0; round < kRefineBlocks / kBlock; ++round) {
            const uint32_t blk = round * kBlock + threadIdx.x;
            uint32_t w[8];
            loadBlock(words, blk, w);
            uint32_t c = 0;
#pragma unroll
            for (uint32_t q = 0; q < 8; ++q) c += __popc(w[q]);
            uint32_t tot;
            const uint32_t ex = blockScan(c, lds4, tot);
            prefix[blk] = run + ex;
            run += tot;
        }
        // every voxel chunk the rewritten list needs must exist
        const uint32_t nChunks = (run + kPointsPerChunk - 1) / kPointsPerChunk;
        const uint32_t* dir = t.voxelDir + uint64_t(g) * kVoxelDir;
        for (uint32_t q = threadIdx.x; q < nChunks; q += blockDim.x)
            if (q >= kVoxelDir || dir[q] == 0u) dirBad = 1;
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t nv = t.numVoxels[node];
            bool ok = t.numVoxelsStored[node] == nv && nv >= run && !dirBad &&
                      uint64_t(t.voxelChunks[g]) * kPointsPerChunk >= run;
            const int32_t fc = t.firstChild[node];
            if (fc >= 0) {
                for (uint32_t o = 0; o < 8; ++o) {
                    const uint32_t c = uint32_t(fc) + o;
                    if (t.firstChild[c] < 0) continue;
                    const uint32_t cg = t.gridNo[c];
                    const uint32_t rv = r.refVoxels[c];
                    if (cg == kNone || rv == kNone || rv != t.numVoxels[c] || r.accCap[cg] == 0) ok = false;
                }
            }
            if (ok && r.accCap[g] < run) {
                const unsigned long long need = run + run / 8 + 64;
                const unsigned long long base = atomicAdd(&r.ctr->arenaUsed, need);
                if (base + need > r.accEntries) {
                    ok = false;
                    r.ctr->arenaFull = 1;
                } else {
                    r.accOff[g] = base;
                    r.accCap[g] = uint32_t(need);
                }
            }
            if (ok && fc >= 0) {
                uint32_t kids = 0;
                for (uint32_t o = 0; o < 8; ++o) {
                    const uint32_t c = uint32_t(fc) + o;
                    if (t.firstChild[c] < 0) continue;
                    r.kids[kidFirst + atomicAdd(&r.ctr->levelKidFill[level], 1u)] = c;
                    ++kids;
                }
                if (kids) atomicAdd(&r.ctr->childSums, (unsigned long long)kids);
            }
            r.refVoxels[node] = kNone;  // the sums are invalid until this call's emission
            r.status[node] = ok ? 1u : 0u;
            r.cells[node] = run;
            okShared = ok ? 1u : 0u;
            offShared = r.accOff[g];
        }
        __syncthreads();
        if (okShared) {
            unsigned long long* e = r.acc + 3ull * offShared;
            for (uint32_t q = threadIdx.x; q < 3u * run; q += blockDim.x) e[q] = 0ull;
        }
        __syncthreads();  // okShared / offShared are rewritten for the next node
    }
}

// Units (node, octant): the points of a leaf child (octant 0 of a gridded leaf:
// its own points) into the node's cell sums. A workgroup folds its leaf in rounds
// of up to 2,032 points through an LDS table keyed by cell rank (a round always
// fits), then adds each distinct cell's sums with one set of global atomics. In
// the table a round's channel sums (at most 255 * 2,032 < 2^21) pack three to a
// u64 (r | g << 21 | b << 42) and alpha with the count (a | n << 32).
constexpr uint32_t kHashSlots = 2032;  // four workgroups' tables fit one CU's LDS
constexpr uint32_t kRoundPoints = kHashSlots;
constexpr unsigned long long kField21 = (1ull << 21) - 1ull;
struct PointTable {
    uint32_t key[kHashSlots];
    unsigned long long rgb[kHashSlots];
    unsigned long long an[kHashSlots];
};

__global__ __launch_bounds__(kBlock) void kRefinePoints(DevTree t, RefineArgs r, uint32_t first, uint32_t count,
                                                        uint32_t level) {
    __shared__ PointTable h;
    __shared__ uint32_t unitOk;
    const uint32_t units = count * 8u;
    const uint32_t sh = (kMaxDepth + 1) - level;
    const float size = t.octreeSize;
    for (uint32_t s = threadIdx.x; s < kHashSlots; s += blockDim.x) {
        h.key[s] = kNone;
        h.rgb[s] = 0;
        h.an[s] = 0;
    }
    for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
        const uint32_t node = r.list[first + u / 8u], o = u % 8u;
        const int32_t fc = t.firstChild[node];
        uint32_t src = node;
        if (threadIdx.x == 0) {
            // read once: another workgroup may clear the status meanwhile
            bool ok = r.status[node] != 0u;
            if (fc < 0) ok = ok && o == 0;
            else ok = ok && t.firstChild[uint32_t(fc) + o] < 0;
            unitOk = ok ? 1u : 0u;
        }
        __syncthreads();
        const bool ok = unitOk != 0u;
        __syncthreads();  // unitOk is rewritten by the next unit
        if (!ok) continue;
        if (fc >= 0) src = uint32_t(fc) + o;
        const uint32_t n = min(t.numPoints[src], kPointDir * kPointsPerChunk);
        const uint32_t g = t.gridNo[node];
        const uint32_t* words = gridWords(t, g);
        const uint32_t* prefix = r.prefix + uint64_t(g) * kRefineBlocks;
        unsigned long long* acc = r.acc + 3ull * r.accOff[g];
        const uint32_t* dir = t.pointDir + uint64_t(src) * kPointDir;
        uint32_t missing = 0;
        for (uint32_t base = 0; base < n; base += kRoundPoints) {
            const uint32_t end = min(n, base + kRoundPoints);
            for (uint32_t i = base + threadIdx.x; i < end; i += blockDim.x) {
                const uint32_t ref = dir[i / kPointsPerChunk];
                if (ref == 0) {
                    ++missing;
                    continue;
                }
                const remo_point p = chunkPtr(t.arena, ref)[i % kPointsPerChunk];
                const uint32_t vi = ((quant28(p.x, size) >> sh) & (kGridSize - 1)) +
                                    ((quant28(p.y, size) >> sh) & (kGridSize - 1)) * kGridSize +
                                    ((quant28(p.z, size) >> sh) & (kGridSize - 1)) * kGridSize * kGridSize;
                uint32_t rank;
                if (!rankOf(words, prefix, vi, rank)) {
                    ++missing;
                    continue;
                }
                // open addressing; a round holds at most kHashSlots distinct ranks
                uint32_t slot = (rank * 2654435761u) % kHashSlots;
                for (;;) {
                    const uint32_t prev = atomicCAS(&h.key[slot], kNone, rank);
                    if (prev == kNone || prev == rank) break;
                    slot = slot + 1u == kHashSlots ? 0u : slot + 1u;
                }
                const uint32_t c = p.color;
                atomicAdd(&h.rgb[slot], (unsigned long long)(c & 0xFFu) |
                                            ((unsigned long long)((c >> 8) & 0xFFu) << 21) |
                                            ((unsigned long long)((c >> 16) & 0xFFu) << 42));
                atomicAdd(&h.an[slot], (unsigned long long)(c >> 24) | (1ull << 32));
            }
            __syncthreads();
            for (uint32_t q = threadIdx.x; q < kHashSlots; q += blockDim.x) {
                const uint32_t key = h.key[q];
                if (key == kNone) continue;
                const unsigned long long rgb = h.rgb[q], an = h.an[q];
                addSums(acc + 3ull * key, (rgb & kField21) | (((rgb >> 21) & kField21) << 32),
                        (rgb >> 42) | ((an & 0xFFFFFFFFull) << 32), an >> 32);
                h.key[q] = kNone;
                h.rgb[q] = 0;
                h.an[q] = 0;
            }
            __syncthreads();
        }
        if (missing) {
            atomicAdd(&r.ctr->missingCells, missing);
            r.status[node] = 0;
        }
        if (threadIdx.x == 0) atomicAdd(&r.ctr->pointsFolded, (unsigned long long)n);
    }
}

// Tasks of an inner child: one per parent cell of its octant, (pz, py, px) in
// 64^3, x fastest. A task reads the child's four words holding cells
// (2px + {0,1}, 2py + {0,1}, 2pz + {0,1}) and adds their sums. Every parent cell
// lies in exactly one child's octant, so its sums are written once, by one task,
// with plain stores.
constexpr uint32_t kChildTasks = 64u * 64u * 64u;

__device__ __forceinline__ void wordAndBase(const uint32_t* words, const uint32_t* prefix, uint32_t wi,
                                            uint32_t& word, uint32_t& base) {
    uint32_t w[8];
    const uint32_t blk = wi >> 3, j = wi & 7u;
    loadBlock(words, blk, w);
    uint32_t b = prefix[blk], x = 0;
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) {
        if (k < j) b += __popc(w[k]);
        if (k == j) x = w[k];
    }
    word = x;
    base = b;
}

__global__ __launch_bounds__(kBlock) void kRefineChildren(DevTree t, RefineArgs r, uint32_t level, uint32_t kidFirst) {
    const uint64_t total = uint64_t(r.ctr->levelKidFill[level]) * kChildTasks;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t task = uint32_t(i % kChildTasks);
        const uint32_t child = r.kids[kidFirst + uint32_t(i / kChildTasks)];
        const uint32_t node = t.parent[child];
        if (!r.status[node]) continue;
        const uint32_t o = child - uint32_t(t.firstChild[node]);
        const uint32_t px = task & 63u, py = (task >> 6) & 63u, pz = task >> 12;
        const uint32_t cg = t.gridNo[child];
        const uint32_t* cwords = gridWords(t, cg);
        // child x cells 2px, 2px + 1: word px / 16, bits 2 (px % 16) and the next
        const uint32_t sh = 2u * (px & 15u);
        uint32_t cw[4];
        bool any = false;
#pragma unroll
        for (uint32_t d = 0; d < 4; ++d) {
            const uint32_t wi = (px >> 4) + 4u * (2u * py + (d & 1u)) + 512u * (2u * pz + (d >> 1));
            cw[d] = cwords[wi];
            any |= ((cw[d] >> sh) & 3u) != 0u;
        }
        if (!any) continue;
        const uint32_t* cprefix = r.prefix + uint64_t(cg) * kRefineBlocks;
        const unsigned long long* cacc = r.acc + 3ull * r.accOff[cg];
        unsigned long long s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
        for (uint32_t d = 0; d < 4; ++d) {
            const uint32_t m = (cw[d] >> sh) & 3u;
            if (!m) continue;
            const uint32_t wi = (px >> 4) + 4u * (2u * py + (d & 1u)) + 512u * (2u * pz + (d >> 1));
            uint32_t word, rank;
            wordAndBase(cwords, cprefix, wi, word, rank);
            rank += __popc(word & ((1u << sh) - 1u));
            if (m & 1u) {
                const unsigned long long* e = cacc + 3ull * rank;
                s0 += e[0];
                s1 += e[1];
                s2 += e[2];
                ++rank;
            }
            if (m & 2u) {
                const unsigned long long* e = cacc + 3ull * rank;
                s0 += e[0];
                s1 += e[1];
                s2 += e[2];
            }
        }
        const uint32_t g = t.gridNo[node];
        const uint32_t ox = (o >> 2) & 1u, oy = (o >> 1) & 1u, oz = o & 1u;
        const uint32_t pvi = (64u * ox + px) + kGridSize * (64u * oy + py) + kGridSize * kGridSize * (64u * oz + pz);
        uint32_t prank;
        if (!rankOf(gridWords(t, g), r.prefix + uint64_t(g) * kRefineBlocks, pvi, prank)) {
            atomicAdd(&r.ctr->missingCells, 1u);
            r.status[node] = 0;
            continue;
        }
        unsigned long long* e = r.acc + 3ull * (r.accOff[g] + prank);
        e[0] = s0;
        e[1] = s1;
        e[2] = s2;
    }
}

// floor((s + h) / n) for s <= 255 n, h < n: the float quotient is within one of it.
__device__ __forceinline__ uint32_t channelMean(unsigned long long s, unsigned long long n, unsigned long long h) {
    const unsigned long long num = s + h;
    uint32_t q = uint32_t(float(num) / float(n));
    if (q > 0u && (unsigned long long)q * n > num) --q;
    if ((unsigned long long)(q + 1u) * n <= num) ++q;
    return q;
}

// Units of 64 grid words, one per wave (1,024 per node). The wave scans the
// words' set-bit counts, then writes its voxels 64 at a time, one per lane in
// rank order (so the sums it reads and the chunk slots it writes are contiguous):
// lane k finds its word by a binary search over the scan and its bit by a
// binary search over popcounts. Positions are the cell centres kVoxelSample
// computes.
constexpr uint32_t kEmitWords = 64;
constexpr uint32_t kEmitUnits = kGridWords / kEmitWords;

__global__ __launch_bounds__(kBlock) void kRefineEmit(DevTree t, RefineArgs r, uint32_t first, uint32_t count,
                                                      uint32_t level) {
    const float size = t.octreeSize;
    const float nodeSize = size / float(1u << level);
    const uint32_t lane = __lane_id();
    constexpr uint32_t kWaves = kBlock / 64;
    const uint64_t units = uint64_t(count) * kEmitUnits;
    for (uint64_t u = uint64_t(blockIdx.x) * kWaves + (threadIdx.x >> 6); u < units;
         u += uint64_t(gridDim.x) * kWaves) {
        const uint32_t node = r.list[first + uint32_t(u / kEmitUnits)];
        const uint32_t seg = uint32_t(u % kEmitUnits);
        const uint32_t ok = __shfl(r.status[node], 0);  // one value for the wave
        if (!ok) {
            if (seg == 0 && lane == 0) atomicAdd(&r.ctr->skipped, 1u);
            continue;
        }
        const uint32_t g = t.gridNo[node];
        const uint32_t word = gridWords(t, g)[seg * kEmitWords + lane];
        uint32_t incl = __popc(word);
#pragma unroll
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(incl, d, 64);
            if (lane >= d) incl += y;
        }
        const uint32_t total = __shfl(incl, 63);
        const uint32_t rank0 = r.prefix[uint64_t(g) * kRefineBlocks + seg * (kEmitWords / kRefineBlockWords)];
        const unsigned long long* acc = r.acc + 3ull * r.accOff[g];
        const uint32_t* dir = t.voxelDir + uint64_t(g) * kVoxelDir;
        const float minx = float(t.nx[node]) * nodeSize;
        const float miny = float(t.ny[node]) * nodeSize;
        const float minz = float(t.nz[node]) * nodeSize;
        uint32_t empty = 0, overflow = 0;
        for (uint32_t j = 0; j < total; j += 64) {
            const uint32_t k = j + lane;
            // the first lane whose inclusive count exceeds k holds voxel k
            uint32_t wl = 0;
#pragma unroll
            for (uint32_t step = 32; step; step >>= 1) {
                const uint32_t v = __shfl(incl, wl + step - 1);
                if (v <= k) wl += step;
            }
            const uint32_t w = __shfl(word, wl);
            const uint32_t before = __shfl(incl, wl) - __popc(w);
            if (k >= total) continue;
            // the (k - before)-th set bit of w
            uint32_t left = k - before, bw = w, bit = 0;
#pragma unroll
            for (uint32_t half = 16; half; half >>= 1) {
                const uint32_t c = __popc(bw & ((1u << half) - 1u));
                if (left >= c) {
                    left -= c;
                    bw >>= half;
                    bit += half;
                }
            }
            const uint32_t vi = (seg * kEmitWords + wl) * 32u + bit;
            const uint32_t cx = vi & (kGridSize - 1), cy = (vi >> 7) & (kGridSize - 1), cz = vi >> 14;
            const uint32_t rank = rank0 + k;
            const unsigned long long* e = acc + 3ull * rank;
            const unsigned long long rg = e[0], ba = e[1], n = e[2];
            uint32_t color = 0;
            if (n == 0ull) {
                ++empty;
            } else {
                if (n > kCellPointLimit) overflow = 1;
                const unsigned long long h = n >> 1;
                color = (channelMean(rg & 0xFFFFFFFFull, n, h) & 0xFFu) |
                        ((channelMean(rg >> 32, n, h) & 0xFFu) << 8) |
                        ((channelMean(ba & 0xFFFFFFFFull, n, h) & 0xFFu) << 16) |
                        ((channelMean(ba >> 32, n, h) & 0xFFu) << 24);
            }
            remo_point v;
            v.x = minx + (nodeSize * (float(cx) + 0.5f)) / 128.0f;
            v.y = miny + (nodeSize * (float(cy) + 0.5f)) / 128.0f;
            v.z = minz + (nodeSize * (float(cz) + 0.5f)) / 128.0f;
            v.color = color;
            chunkPtr(t.arena, dir[rank / kPointsPerChunk])[rank % kPointsPerChunk] = v;
        }
        if (empty) atomicAdd(&r.ctr->emptyCells, empty);
        if (overflow) r.ctr->cellOverflow = 1;
        if (seg == 0 && lane == 0) {
            const uint32_t cells = r.cells[node];
            t.numVoxels[node] = cells;
            t.numVoxelsStored[node] = cells;
            r.refVoxels[node] = cells;
            const int32_t fc = t.firstChild[node];
            if (fc < 0) {
                r.refPoints[node] = t.numPoints[node];
            } else {
                for (uint32_t o = 0; o < 8; ++o) {
                    const uint32_t c = uint32_t(fc) + o;
                    if (t.firstChild[c] < 0) r.refPoints[c] = t.numPoints[c];
                }
            }
            atomicAdd(&r.ctr->refined, 1u);
            atomicAdd(&r.ctr->voxelsWritten, (unsigned long long)cells);
        }
    }
}

__global__ void kFill32(uint32_t* p, uint64_t n, uint32_t v) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

// Cost model (microseconds), fitted to per-launch times of a full pass over a
// 100M-point tree (6 levels, 1,259 gridded nodes, 53M voxels).
constexpr double kLevelUs = 150.0;      // a level's launches and their fixed costs
constexpr double kNodeUs = 0.13;        // rank prefix of one 256 KB grid
constexpr double kPointUs = 0.00009;    // one leaf point: load, rank, LDS fold, its share of the flush
constexpr double kChildUs = 2.7;        // one inner child: its 64^3 parent-cell tasks
constexpr double kVoxelUs = 0.000035;   // one voxel written

double nodesCostUs(const RefineCounters& c, uint32_t L) {
    return kNodeUs * c.levelCount[L] + kPointUs * double(c.levelPoints[L]) + kChildUs * c.levelKids[L] +
           kVoxelUs * double(c.levelVoxels[L]);
}

uint32_t blocksFor(uint64_t units, uint32_t cap) {
    return uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(units, cap)));
}

}  // namespace

int refineAllocate(RefineState& r, const DevTree& t, uint64_t entries, const RefineAllocFn& alloc, hipStream_t s) {
    const uint64_t N = t.maxNodes, G = t.maxGrids;
    int rc = 0;
    auto A = [&](auto** p, uint64_t bytes, const char* what) {
        if (rc != 0) return;
        void* q = nullptr;
        rc = alloc(&q, std::max<uint64_t>(bytes, 16), what);
        *p = static_cast<std::remove_pointer_t<decltype(p)>>(q);
    };
    A(&r.refPoints, 4 * N, "refinement point marks");
    A(&r.refVoxels, 4 * N, "refinement voxel marks");
    A(&r.dirty, 4 * N, "refinement dirty flags");
    A(&r.status, 4 * N, "refinement status");
    A(&r.cells, 4 * N, "refinement cell counts");
    A(&r.list, 4 * N, "refinement node list");
    A(&r.kids, 4 * N, "refinement child list");
    A(&r.prefix, 4 * G * kRefineBlocks, "refinement rank prefixes");
    A(&r.accOff, 8 * G, "refinement sum offsets");
    A(&r.accCap, 4 * G, "refinement sum capacities");
    A(&r.ctr, sizeof(RefineCounters), "refinement counters");
    if (rc != 0) return rc;
    if (hipHostMalloc(reinterpret_cast<void**>(&r.host), sizeof(RefineCounters), hipHostMallocDefault) != hipSuccess)
        return -2;
    std::memset(r.host, 0, sizeof(RefineCounters));
    if (hipEventCreate(&r.t0) != hipSuccess || hipEventCreate(&r.t1) != hipSuccess) return -3;
    if (hipMalloc(&r.acc, std::max<uint64_t>(entries, 1) * 24) != hipSuccess) {
        r.acc = nullptr;
        return -2;
    }
    r.accEntries = entries;
    r.allocated = true;
    refineReset(r, t, s);
    return 0;
}

void refineReset(RefineState& r, const DevTree& t, hipStream_t s) {
    if (!r.allocated) return;
    const uint64_t N = t.maxNodes, G = t.maxGrids;
    hipLaunchKernelGGL(kFill32, dim3(256), dim3(kBlock), 0, s, r.refVoxels, N, kNone);
    hipLaunchKernelGGL(kFill32, dim3(256), dim3(kBlock), 0, s, r.refPoints, N, 0u);
    hipLaunchKernelGGL(kFill32, dim3(256), dim3(kBlock), 0, s, r.accCap, G, 0u);
    hipLaunchKernelGGL(kFill32, dim3(1), dim3(kBlock), 0, s, reinterpret_cast<uint32_t*>(r.ctr),
                       uint64_t(sizeof(RefineCounters) / 4), 0u);
    std::memset(r.host, 0, sizeof(RefineCounters));
}

int refineGrow(RefineState& r, const DevTree& t, uint64_t entries, hipStream_t s, std::string* err) {
    if (hipStreamSynchronize(s) != hipSuccess) return -3;
    (void)hipFree(r.acc);
    r.acc = nullptr;
    r.accEntries = 0;
    if (hipMalloc(&r.acc, std::max<uint64_t>(entries, 1) * 24) != hipSuccess) {
        r.acc = nullptr;
        if (err) *err = "hipMalloc failed for the refinement sums (" + std::to_string((entries * 24) >> 20) + " MiB)";
        return -2;
    }
    r.accEntries = entries;
    refineReset(r, t, s);
    return 0;
}

int runRefine(RefineState& r, const DevTree& t, uint32_t numNodes, float budgetMs, uint32_t maxLevels,
              const LaunchShape& ls, hipStream_t s, RefineStatsOut* out, std::string* err) {
    std::memset(out, 0, sizeof(*out));
    const auto wall0 = std::chrono::steady_clock::now();
    const RefineArgs ra = argsOf(r);
    const uint32_t nb = ls.nodeBlocks;
    hipLaunchKernelGGL(kRefineClear, dim3(nb), dim3(kBlock), 0, s, ra, numNodes);
    hipLaunchKernelGGL(kRefineMark, dim3(nb), dim3(kBlock), 0, s, t, ra, numNodes);
    hipLaunchKernelGGL(kRefineCount, dim3(nb), dim3(kBlock), 0, s, t, ra, numNodes);
    if (hipMemcpyAsync(r.host, r.ctr, sizeof(RefineCounters), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        if (err) *err = "refinement: reading the dirty counts";
        return -3;
    }
    const RefineCounters c = *r.host;
    int deepest = -1, shallowest = -1;
    for (int L = int(kRefineLevels) - 1; L >= 0; --L) {
        if (!c.levelCount[L]) continue;
        if (deepest < 0) deepest = L;
        shallowest = L;
        ++out->levelsPending;
        out->nodesDirty += c.levelCount[L];
    }
    if (deepest < 0 || budgetMs == 0.0f) {
        out->complete = deepest < 0 ? 1u : 0u;
        return 0;
    }
    // entries the dirty nodes may take (prep reserves cells + cells/8 + 64): when
    // they would not fit beside what is handed out, the caller replaces the store
    // first (reclaiming what grown nodes left behind) and calls again
    {
        uint64_t need = 0;
        for (uint32_t L = 0; L < kRefineLevels; ++L)
            need += c.levelVoxels[L] + c.levelVoxels[L] / 8 + 64ull * c.levelCount[L];
        if (c.arenaUsed > 0 && c.arenaUsed + need > r.accEntries) return kRefineReplaceStore;
    }
    // Levels from the deepest while the estimate fits. The last level taken may be
    // cut to the nodes that fit at its average node cost (nodes of one level do
    // not depend on each other; the rest stay dirty); at least one node runs. The
    // kernels get the budget less the call's own cost (the marking kernels, two
    // readbacks, the launches: a decaying peak of what recent calls measured),
    // with 20 % kept back for the model's error, so the call's wall time holds it.
    const double kernelBudget =
        budgetMs > 0.0f ? std::max(0.0, (double(budgetMs) - r.overheadMs) * 0.80) : double(budgetMs);
    double est = 0.0;
    int lo = deepest;
    uint32_t taken = 0, lastCount = 0;
    for (int L = deepest; L >= 0; --L) {
        if (!c.levelCount[L]) continue;
        if (maxLevels && taken >= maxLevels) break;
        // the fitted model, corrected by what recent calls measured against it
        const double fixed = kLevelUs * 1e-3 * r.modelScale, nodes = nodesCostUs(c, uint32_t(L)) * 1e-3 * r.modelScale;
        uint32_t k = c.levelCount[L];
        if (budgetMs > 0.0f && est + fixed + nodes > kernelBudget) {
            const double room = kernelBudget - est - fixed;
            const double per = nodes / double(c.levelCount[L]);
            k = room > 0.0 ? uint32_t(std::min<double>(double(c.levelCount[L]), room / per)) : 0u;
            if (k == 0 && taken > 0) break;
            k = std::max(k, 1u);
        }
        est += fixed + nodes * double(k) / double(c.levelCount[L]);
        lo = L;
        lastCount = k;
        ++taken;
        if (k < c.levelCount[L]) break;
    }
    RefineLevelOffsets offs{};
    uint32_t kidOff[kRefineLevels] = {};
    uint32_t run = 0, kidRun = 0;
    for (int L = deepest; L >= lo; --L) {
        offs.off[L] = run;
        run += c.levelCount[L];
        kidOff[L] = kidRun;
        kidRun += c.levelKids[L];
    }
    offs.lo = uint32_t(lo);
    offs.hi = uint32_t(deepest);
    hipLaunchKernelGGL(kRefineList, dim3(nb), dim3(kBlock), 0, s, t, ra, numNodes, offs);
    (void)hipEventRecord(r.t0, s);
    for (int L = deepest; L >= lo; --L) {
        const uint32_t n = L == lo ? lastCount : c.levelCount[L];
        if (!n) continue;
        const uint32_t first = offs.off[L];
        hipLaunchKernelGGL(kRefinePrep, dim3(blocksFor(n, 4096)), dim3(kBlock), 0, s, t, ra, first, n, uint32_t(L),
                           kidOff[L]);
        hipLaunchKernelGGL(kRefinePoints, dim3(blocksFor(uint64_t(n) * 8, 8192)), dim3(kBlock), 0, s, t, ra, first,
                           n, uint32_t(L));
        if (c.levelKids[L])
            hipLaunchKernelGGL(kRefineChildren, dim3(blocksFor(uint64_t(c.levelKids[L]) * kChildTasks / kBlock, 32768)),
                               dim3(kBlock), 0, s, t, ra, uint32_t(L), kidOff[L]);
        hipLaunchKernelGGL(kRefineEmit, dim3(blocksFor(uint64_t(n) * kEmitUnits / (kBlock / 64), 16384)),
                           dim3(kBlock), 0, s, t, ra, first, n, uint32_t(L));
    }
    (void)hipEventRecord(r.t1, s);
    if (hipGetLastError() != hipSuccess) {
        if (err) *err = "refinement kernels failed to launch";
        return -3;
    }
    if (hipMemcpyAsync(r.host, r.ctr, sizeof(RefineCounters), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        if (err) *err = "refinement: reading the counters";
        return -3;
    }
    const RefineCounters& d = *r.host;
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, r.t0, r.t1);
    out->levelsProcessed = taken;
    out->nodesRefined = d.refined;
    out->nodesSkipped = d.skipped;
    out->complete = (lo == shallowest && lastCount == c.levelCount[lo] && d.skipped == 0) ? 1u : 0u;
    out->pointsFolded = d.pointsFolded;
    out->childSums = d.childSums;
    out->voxelsWritten = d.voxelsWritten;
    out->missingCells = d.missingCells;
    out->emptyCells = d.emptyCells;
    out->arenaFull = d.arenaFull;
    out->cellOverflow = d.cellOverflow;
    out->estimateMs = float(est);
    // a running correction of the model: the measured time over the estimate
    if (est > 0.05 && ms > 0.0f) {
        const double ratio = double(ms) / (est / r.modelScale);
        r.modelScale = std::min(8.0, std::max(0.25, 0.7 * r.modelScale + 0.3 * ratio));
    }
    out->deviceMs = ms;
    const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
    const double over = std::max(0.0, wall - double(ms));
    r.overheadMs = std::max(over, 0.97 * r.overheadMs + 0.03 * over);
    return 0;
}

void refineRelease(RefineState& r) {
    if (r.acc) (void)hipFree(r.acc);
    r.acc = nullptr;
    if (r.host) (void)hipHostFree(r.host);
    r.host = nullptr;
    if (r.t0) (void)hipEventDestroy(r.t0);
    if (r.t1) (void)hipEventDestroy(r.t1);
    r.t0 = r.t1 = nullptr;
    r.allocated = false;
}

}  // namespace remo_hip

// Device wide-BVH builder: binned SAH, top-down, then collapsed to 4- or 8-wide nodes in the layout
// the traversal kernels read (WideNode<W>, leaf-order triangles), like the host builder.
//
// Replaces Embree's scene build behind rtcCommitScene (Tracer/Embree/Embree.cpp:648). As with the
// host builder, the hit result does not depend on the tree (closest t, ties to the smaller triangle
// index), so the tree only decides traversal cost.
//
// - Nodes of more than kSmallTask triangles ("large tasks") are split level by level with the whole
//   device: each task's triangles are cut into chunks of kChunk; per chunk a block reduces bounds,
//   bins centroids (kBins per axis, LDS atomics) and merges into the task's bins; one thread per
//   task sweeps the bins; a stable partition (chunk counts, their prefix, in-block prefix) moves
//   the triangle indices. The host sequences the levels and allocates nodes in a fixed order.
// - Every node of at most kSmallTask triangles is finished by one wave: the same binned SAH on an
//   explicit stack in LDS, node indices from a range reserved for its subtree.
// - The binary tree is collapsed top-down one wide level at a time (each wide node takes the
//   largest-area internal child apart until it holds W children), child indices from a prefix sum.
// Every step is deterministic: the same input gives the same tree.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "pg_bvh.hpp"

namespace pg {
namespace {

constexpr uint32_t kSmallTask = 2048;  // triangles a single wave finishes
constexpr uint32_t kChunk = 8192;      // triangles per block in the large-task passes
constexpr uint32_t kMaxBins = 32;
constexpr uint32_t kStack = 256;       // wave-local task stack
constexpr uint32_t kLeafMark = 0x80000000u;

#define PGB_HIP(call)                                                                      \
    do {                                                                                   \
        const hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct TmpBuf {
    void* p = nullptr;
    size_t bytes = 0;
    TmpBuf() = default;
    explicit TmpBuf(size_t n) { reset(n); }
    TmpBuf(const TmpBuf&) = delete;
    TmpBuf& operator=(const TmpBuf&) = delete;
    ~TmpBuf() {
        if (p) (void)hipFree(p);
    }
    void reset(size_t n) {  // at least n bytes (contents undefined)
        n = std::max<size_t>(n, 16);
        if (n <= bytes) return;
        if (p) (void)hipFree(p);
        p = nullptr, bytes = 0;
        PGB_HIP(hipMalloc(&p, n));
        bytes = n;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// Order-preserving float <-> int map, so integer atomics give float min / max.
__device__ __forceinline__ int ford(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float funord(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7FFFFFFF); }

// Binary node: leaf iff a has kLeafMark (a = first triangle position | mark, b = count);
// internal: a, b = child nodes.
struct BNode {
    float lo[3];
    uint32_t a;
    float hi[3];
    uint32_t b;
};

// Bounds of one task as ordered ints: box lo[3], hi[3], centroid lo[3], hi[3].
struct TaskBounds {
    int v[12];
};

// Bins of one task: per axis and bin, count and box (ordered ints).
struct TaskBins {
    int cnt[3][kMaxBins];
    int lo[3][kMaxBins][3];
    int hi[3][kMaxBins][3];
};

struct LargeTask {
    uint32_t begin, count, node, pad;
};

struct Split {
    int axis;         // -1: split by position (median)
    uint32_t bin;     // triangles with bin < this go left
    float scale, lo;  // bin = (centroid[axis] - lo) * scale
};

__device__ __forceinline__ float halfArea(const float lo[3], const float hi[3]) {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return dx < 0.0f ? 0.0f : dx * dy + dy * dz + dz * dx;
}

__device__ __forceinline__ uint32_t binOf(float c, float lo, float scale, uint32_t bins) {
    const float f = fmaxf(0.0f, (c - lo) * scale);
    return min(bins - 1, static_cast<uint32_t>(f));
}

// ---- triangle bounds, identity order, scene extent ----
__global__ void __launch_bounds__(256) kPrimPrep(const float* __restrict__ tv, uint32_t n, float4* plo, float4* phi,
                                                 uint32_t* prim, int* maxAbsOrd) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float m = 1.0f;
    if (i < n) {
        const float* t = tv + 9 * static_cast<size_t>(i);
        float lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(fminf(t[k], t[3 + k]), t[6 + k]);
            hi[k] = fmaxf(fmaxf(t[k], t[3 + k]), t[6 + k]);
            m = fmaxf(m, fmaxf(fabsf(lo[k]), fabsf(hi[k])));
        }
        plo[i] = make_float4(lo[0], lo[1], lo[2], 0.0f);
        phi[i] = make_float4(hi[0], hi[1], hi[2], 0.0f);
        prim[i] = i;
    }
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if (__lane_id() == 0) atomicMax(maxAbsOrd, ford(m));
}

// ---- large tasks: one block per (task, chunk) ----
struct ChunkRef {
    uint32_t task, start, count;  // chunk of task: positions [start, start + count)
};

__device__ __forceinline__ void waveMinMax(float& lo, float& hi) {
    for (int off = 32; off > 0; off >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, off));
        hi = fmaxf(hi, __shfl_xor(hi, off));
    }
}

__global__ void __launch_bounds__(256) kLargeBounds(const ChunkRef* chunks, const uint32_t* prim, const float4* plo,
                                                    const float4* phi, TaskBounds* tb) {
    const ChunkRef c = chunks[blockIdx.x];
    float v[12];
    for (int k = 0; k < 3; ++k) v[k] = v[6 + k] = INFINITY, v[3 + k] = v[9 + k] = -INFINITY;
    for (uint32_t j = threadIdx.x; j < c.count; j += blockDim.x) {
        const uint32_t t = prim[c.start + j];
        const float4 a = plo[t], b = phi[t];
        const float lo[3] = {a.x, a.y, a.z}, hi[3] = {b.x, b.y, b.z};
        for (int k = 0; k < 3; ++k) {
            const float cen = 0.5f * (lo[k] + hi[k]);
            v[k] = fminf(v[k], lo[k]), v[3 + k] = fmaxf(v[3 + k], hi[k]);
            v[6 + k] = fminf(v[6 + k], cen), v[9 + k] = fmaxf(v[9 + k], cen);
        }
    }
    for (int k = 0; k < 3; ++k) {
        waveMinMax(v[k], v[3 + k]);
        waveMinMax(v[6 + k], v[9 + k]);
    }
    if (__lane_id() == 0) {
        int* o = tb[c.task].v;
        for (int k = 0; k < 3; ++k) {
            atomicMin(&o[k], ford(v[k]));
            atomicMax(&o[3 + k], ford(v[3 + k]));
            atomicMin(&o[6 + k], ford(v[6 + k]));
            atomicMax(&o[9 + k], ford(v[9 + k]));
        }
    }
}

__device__ __forceinline__ void taskRange(const TaskBounds& tb, int k, uint32_t bins, float& lo, float& scale) {
    lo = funord(tb.v[6 + k]);
    const float ext = funord(tb.v[9 + k]) - lo;
    scale = ext > 0.0f ? static_cast<float>(bins) / ext : 0.0f;
}

__global__ void __launch_bounds__(256) kLargeBins(const ChunkRef* chunks, const uint32_t* prim, const float4* plo,
                                                  const float4* phi, const TaskBounds* tb, TaskBins* bins,
                                                  uint32_t nb) {
    __shared__ int sCnt[3][kMaxBins];
    __shared__ int sLo[3][kMaxBins][3];
    __shared__ int sHi[3][kMaxBins][3];
    const ChunkRef c = chunks[blockIdx.x];
    for (uint32_t j = threadIdx.x; j < 3 * kMaxBins; j += blockDim.x) {
        (&sCnt[0][0])[j] = 0;
        for (int k = 0; k < 3; ++k) (&sLo[0][0][0])[3 * j + k] = ford(INFINITY), (&sHi[0][0][0])[3 * j + k] = ford(-INFINITY);
    }
    __syncthreads();
    const TaskBounds b = tb[c.task];
    float rlo[3], rsc[3];
    for (int k = 0; k < 3; ++k) taskRange(b, k, nb, rlo[k], rsc[k]);
    for (uint32_t j = threadIdx.x; j < c.count; j += blockDim.x) {
        const uint32_t t = prim[c.start + j];
        const float4 a = plo[t], h = phi[t];
        const float lo[3] = {a.x, a.y, a.z}, hi[3] = {h.x, h.y, h.z};
        for (int k = 0; k < 3; ++k) {
            if (rsc[k] == 0.0f) continue;
            const uint32_t bi = binOf(0.5f * (lo[k] + hi[k]), rlo[k], rsc[k], nb);
            atomicAdd(&sCnt[k][bi], 1);
            for (int e = 0; e < 3; ++e) {
                atomicMin(&sLo[k][bi][e], ford(lo[e]));
                atomicMax(&sHi[k][bi][e], ford(hi[e]));
            }
        }
    }
    __syncthreads();
    TaskBins& o = bins[c.task];
    for (uint32_t j = threadIdx.x; j < 3 * nb; j += blockDim.x) {
        const uint32_t k = j / nb, bi = j % nb;
        if (sCnt[k][bi] == 0) continue;
        atomicAdd(&o.cnt[k][bi], sCnt[k][bi]);
        for (int e = 0; e < 3; ++e) {
            atomicMin(&o.lo[k][bi][e], sLo[k][bi][e]);
            atomicMax(&o.hi[k][bi][e], sHi[k][bi][e]);
        }
    }
}

// SAH sweep over one axis' bins: cost = area(node) + area(L) nL + area(R) nR (a traversal step costs
// about one triangle test), as the host builder.
__device__ void sweepAxis(const int* cnt, const int (*blo)[3], const int (*bhi)[3], uint32_t nb, float nodeArea,
                          float& bestCost, uint32_t& bestBin, bool& found) {
    float rArea[kMaxBins];
    int rCnt[kMaxBins];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    int ac = 0;
    for (uint32_t i = nb - 1; i > 0; --i) {
        if (cnt[i]) {
            for (int e = 0; e < 3; ++e) lo[e] = fminf(lo[e], funord(blo[i][e])), hi[e] = fmaxf(hi[e], funord(bhi[i][e]));
        }
        ac += cnt[i];
        rArea[i] = halfArea(lo, hi);
        rCnt[i] = ac;
    }
    float llo[3] = {INFINITY, INFINITY, INFINITY}, lhi[3] = {-INFINITY, -INFINITY, -INFINITY};
    int lc = 0;
    for (uint32_t i = 0; i + 1 < nb; ++i) {
        if (cnt[i]) {
            for (int e = 0; e < 3; ++e) llo[e] = fminf(llo[e], funord(blo[i][e])), lhi[e] = fmaxf(lhi[e], funord(bhi[i][e]));
        }
        lc += cnt[i];
        if (lc == 0 || rCnt[i + 1] == 0) continue;
        const float cost = nodeArea + halfArea(llo, lhi) * static_cast<float>(lc) + rArea[i + 1] * static_cast<float>(rCnt[i + 1]);
        if (cost < bestCost) bestCost = cost, bestBin = i + 1, found = true;
    }
}

// One thread per large task: node box, best split.
__global__ void kLargeSplit(const LargeTask* tasks, uint32_t numTasks, const TaskBounds* tb, const TaskBins* bins,
                            uint32_t nb, BNode* nodes, Split* splits) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= numTasks) return;
    const TaskBounds b = tb[t];
    BNode& nd = nodes[tasks[t].node];
    float lo[3], hi[3];
    for (int k = 0; k < 3; ++k) lo[k] = funord(b.v[k]), hi[k] = funord(b.v[3 + k]), nd.lo[k] = lo[k], nd.hi[k] = hi[k];
    float best = INFINITY;
    uint32_t bestBin = 0;
    int axis = -1;
    const float area = halfArea(lo, hi);
    for (int k = 0; k < 3; ++k) {
        float rlo, rsc;
        taskRange(b, k, nb, rlo, rsc);
        if (rsc == 0.0f) continue;
        bool found = false;
        uint32_t bin = 0;
        float c = best;
        sweepAxis(bins[t].cnt[k], bins[t].lo[k], bins[t].hi[k], nb, area, c, bin, found);
        if (found) best = c, bestBin = bin, axis = k;
    }
    Split s{axis, bestBin, 0.0f, 0.0f};
    if (axis >= 0) taskRange(b, axis, nb, s.lo, s.scale);
    splits[t] = s;
}

__device__ __forceinline__ bool goesLeft(const Split& s, uint32_t pos, uint32_t begin, uint32_t count, const float4* plo,
                                         const float4* phi, uint32_t t, uint32_t nb) {
    if (s.axis < 0) return pos - begin < count / 2;
    const float4 a = plo[t], h = phi[t];
    const float c = s.axis == 0 ? 0.5f * (a.x + h.x) : s.axis == 1 ? 0.5f * (a.y + h.y) : 0.5f * (a.z + h.z);
    return binOf(c, s.lo, s.scale, nb) < s.bin;
}

// Triangles per chunk that go left.
__global__ void __launch_bounds__(256) kLargeCount(const ChunkRef* chunks, const LargeTask* tasks, const Split* splits,
                                                   const uint32_t* prim, const float4* plo, const float4* phi,
                                                   uint32_t nb, uint32_t* chunkLeft) {
    __shared__ uint32_t s;
    const ChunkRef c = chunks[blockIdx.x];
    const LargeTask tk = tasks[c.task];
    const Split sp = splits[c.task];
    if (threadIdx.x == 0) s = 0;
    __syncthreads();
    uint32_t mine = 0;
    for (uint32_t j = threadIdx.x; j < c.count; j += blockDim.x)
        mine += goesLeft(sp, c.start + j, tk.begin, tk.count, plo, phi, prim[c.start + j], nb) ? 1u : 0u;
    for (int off = 32; off > 0; off >>= 1) mine += __shfl_xor(mine, off);
    if (__lane_id() == 0) atomicAdd(&s, mine);
    __syncthreads();
    if (threadIdx.x == 0) chunkLeft[blockIdx.x] = s;
}

// Stable partition: a chunk's left triangles go to begin + leftBefore[chunk] + (lefts before them
// in the chunk), its right ones to begin + nLeft + rightBefore[chunk] + (rights before them).
__global__ void __launch_bounds__(256) kLargeScatter(const ChunkRef* chunks, const LargeTask* tasks,
                                                     const Split* splits, const uint32_t* prim, const float4* plo,
                                                     const float4* phi, uint32_t nb, const uint32_t* leftBefore,
                                                     const uint32_t* taskLeft, uint32_t* prim2) {
    __shared__ uint32_t waveL[4], waveR[4];
    const ChunkRef c = chunks[blockIdx.x];
    const LargeTask tk = tasks[c.task];
    const Split sp = splits[c.task];
    const uint32_t nLeft = taskLeft[c.task];
    const uint32_t wave = threadIdx.x / 64, lane = __lane_id();
    const uint32_t chunkOff = c.start - tk.begin;
    uint32_t lBase = leftBefore[blockIdx.x], rBase = chunkOff - leftBefore[blockIdx.x];
    const uint64_t lower = (1ull << lane) - 1ull;
    for (uint32_t j0 = 0; j0 < c.count; j0 += blockDim.x) {  // block-uniform
        const uint32_t j = j0 + threadIdx.x;
        const bool act = j < c.count;
        uint32_t t = 0;
        bool left = false;
        if (act) {
            t = prim[c.start + j];
            left = goesLeft(sp, c.start + j, tk.begin, tk.count, plo, phi, t, nb);
        }
        const uint64_t ml = __ballot(act && left), mr = __ballot(act && !left);
        if (lane == 0) waveL[wave] = __popcll(ml), waveR[wave] = __popcll(mr);
        __syncthreads();
        uint32_t lOff = lBase, rOff = rBase, lTot = 0, rTot = 0;
        for (uint32_t w = 0; w < blockDim.x / 64; ++w) {
            if (w < wave) lOff += waveL[w], rOff += waveR[w];
            lTot += waveL[w], rTot += waveR[w];
        }
        if (act) {
            const uint32_t dst = left ? tk.begin + lOff + __popcll(ml & lower)
                                      : tk.begin + nLeft + rOff + __popcll(mr & lower);
            prim2[dst] = t;
        }
        lBase += lTot, rBase += rTot;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) kCopyRange(const ChunkRef* chunks, const uint32_t* src, uint32_t* dst) {
    const ChunkRef c = chunks[blockIdx.x];
    for (uint32_t j = threadIdx.x; j < c.count; j += blockDim.x) dst[c.start + j] = src[c.start + j];
}

// ---- small tasks: one wave builds the whole subtree ----
struct SmallTask {
    uint32_t begin, count, node, reserve;  // descendants take nodes [reserve, reserve + 2 count - 2)
};

struct WaveBins {
    int cnt[3][kMaxBins];
    int lo[3][kMaxBins][3];
    int hi[3][kMaxBins][3];
};

__global__ void __launch_bounds__(64) kSmall(const SmallTask* tasks, uint32_t* prim, uint32_t* tmp, const float4* plo,
                                             const float4* phi, BNode* nodes, uint32_t leafSize, uint32_t nb) {
    __shared__ uint4 stack[kStack];  // begin, count, node
    __shared__ WaveBins wb;
    __shared__ float sCost[3];
    __shared__ uint32_t sBin[3];
    const SmallTask task = tasks[blockIdx.x];
    const uint32_t lane = threadIdx.x;
    const uint64_t lower = (1ull << lane) - 1ull;
    uint32_t sp = 0, alloc = task.reserve;
    if (lane == 0) stack[0] = make_uint4(task.begin, task.count, task.node, 0);
    sp = 1;
    __syncthreads();
    while (sp > 0) {
        const uint4 e = stack[--sp];
        __syncthreads();
        const uint32_t begin = e.x, count = e.y, node = e.z;
        // bounds and centroid bounds
        float v[12];
        for (int k = 0; k < 3; ++k) v[k] = v[6 + k] = INFINITY, v[3 + k] = v[9 + k] = -INFINITY;
        for (uint32_t j = lane; j < count; j += 64) {
            const uint32_t t = prim[begin + j];
            const float4 a = plo[t], h = phi[t];
            const float lo[3] = {a.x, a.y, a.z}, hi[3] = {h.x, h.y, h.z};
            for (int k = 0; k < 3; ++k) {
                const float cen = 0.5f * (lo[k] + hi[k]);
                v[k] = fminf(v[k], lo[k]), v[3 + k] = fmaxf(v[3 + k], hi[k]);
                v[6 + k] = fminf(v[6 + k], cen), v[9 + k] = fmaxf(v[9 + k], cen);
            }
        }
        for (int k = 0; k < 3; ++k) {
            waveMinMax(v[k], v[3 + k]);
            waveMinMax(v[6 + k], v[9 + k]);
        }
        const float nlo[3] = {v[0], v[1], v[2]}, nhi[3] = {v[3], v[4], v[5]};
        const float area = halfArea(nlo, nhi);
        if (lane == 0)
            for (int k = 0; k < 3; ++k) nodes[node].lo[k] = nlo[k], nodes[node].hi[k] = nhi[k];
        float rlo[3], rsc[3];
        for (int k = 0; k < 3; ++k) {
            rlo[k] = v[6 + k];
            const float ext = v[9 + k] - v[6 + k];
            rsc[k] = ext > 0.0f ? static_cast<float>(nb) / ext : 0.0f;
        }
        const float leafCost = static_cast<float>(count) * area;
        int axis = -1;
        uint32_t bin = 0;
        float best = leafCost;
        if (count > 1 && sp + 2 < kStack - 48) {  // near a full stack, split by position (halves the depth)
            for (uint32_t j = lane; j < 3 * kMaxBins; j += 64) {
                (&wb.cnt[0][0])[j] = 0;
                for (int k = 0; k < 3; ++k) (&wb.lo[0][0][0])[3 * j + k] = ford(INFINITY), (&wb.hi[0][0][0])[3 * j + k] = ford(-INFINITY);
            }
            __syncthreads();
            for (uint32_t j = lane; j < count; j += 64) {
                const uint32_t t = prim[begin + j];
                const float4 a = plo[t], h = phi[t];
                const float lo[3] = {a.x, a.y, a.z}, hi[3] = {h.x, h.y, h.z};
                for (int k = 0; k < 3; ++k) {
                    if (rsc[k] == 0.0f) continue;
                    const uint32_t bi = binOf(0.5f * (lo[k] + hi[k]), rlo[k], rsc[k], nb);
                    atomicAdd(&wb.cnt[k][bi], 1);
                    for (int q = 0; q < 3; ++q) atomicMin(&wb.lo[k][bi][q], ford(lo[q])), atomicMax(&wb.hi[k][bi][q], ford(hi[q]));
                }
            }
            __syncthreads();
            if (lane < 3) {
                float c = leafCost;
                uint32_t b = 0;
                bool found = false;
                if (rsc[lane] != 0.0f) sweepAxis(wb.cnt[lane], wb.lo[lane], wb.hi[lane], nb, area, c, b, found);
                sCost[lane] = found ? c : INFINITY;
                sBin[lane] = b;
            }
            __syncthreads();
            for (int k = 0; k < 3; ++k)
                if (sCost[k] < best) best = sCost[k], axis = k, bin = sBin[k];
            __syncthreads();
        }
        const bool leaf = count == 1 || (count <= leafSize && (axis < 0 || best >= leafCost));
        if (leaf) {
            if (lane == 0) nodes[node].a = begin | kLeafMark, nodes[node].b = count;
            __syncthreads();
            continue;
        }
        // stable partition of [begin, begin + count) through tmp
        const Split s{axis, bin, axis >= 0 ? rsc[axis] : 0.0f, axis >= 0 ? rlo[axis] : 0.0f};
        uint32_t nLeft = 0;
        for (uint32_t j = lane; j - lane < count; j += 64) {
            const bool act = j < count;
            const bool l = act && goesLeft(s, begin + j, begin, count, plo, phi, prim[begin + (act ? j : 0)], nb);
            nLeft += __popcll(__ballot(l));
        }
        uint32_t lOff = 0, rOff = 0;
        for (uint32_t j = lane; j - lane < count; j += 64) {
            const bool act = j < count;
            const uint32_t t = act ? prim[begin + j] : 0u;
            const bool l = act && goesLeft(s, begin + j, begin, count, plo, phi, t, nb);
            const uint64_t ml = __ballot(l), mr = __ballot(act && !l);
            if (act) tmp[begin + (l ? lOff + __popcll(ml & lower) : nLeft + rOff + __popcll(mr & lower))] = t;
            lOff += __popcll(ml), rOff += __popcll(mr);
        }
        __syncthreads();  // the lanes' tmp stores are visible before any lane reads them back
        for (uint32_t j = lane; j < count; j += 64) prim[begin + j] = tmp[begin + j];
        const uint32_t left = alloc, right = alloc + 1;
        alloc += 2;
        if (lane == 0) {
            nodes[node].a = left, nodes[node].b = right;
            stack[sp] = make_uint4(begin + nLeft, count - nLeft, right, 0);
            stack[sp + 1] = make_uint4(begin, nLeft, left, 0);
        }
        sp += 2;
        __syncthreads();
    }
}

// ---- collapse to wide nodes, one wide level per launch pair ----
template <int W>
__device__ int gatherKids(const BNode* nodes, uint32_t b, uint32_t kids[W]) {
    if (nodes[b].a & kLeafMark) {
        kids[0] = b;
        return 1;
    }
    kids[0] = nodes[b].a, kids[1] = nodes[b].b;
    int k = 2;
    while (k < W) {
        int best = -1;
        float area = -1.0f;
        for (int i = 0; i < k; ++i) {
            const BNode& c = nodes[kids[i]];
            if (c.a & kLeafMark) continue;
            const float a = halfArea(c.lo, c.hi);
            if (a > area) area = a, best = i;
        }
        if (best < 0) break;
        const BNode c = nodes[kids[best]];
        kids[best] = c.a;
        kids[k++] = c.b;
    }
    return k;
}

template <int W>
__global__ void __launch_bounds__(256) kCollapseCount(const BNode* nodes, const uint32_t* level, uint32_t n,
                                                      uint32_t* internalKids, uint32_t* leafKids) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t kids[W];
    const int k = gatherKids<W>(nodes, level[i], kids);
    uint32_t in = 0;
    for (int j = 0; j < k; ++j) in += (nodes[kids[j]].a & kLeafMark) ? 0u : 1u;
    internalKids[i] = in;
    leafKids[i] = static_cast<uint32_t>(k) - in;
}

template <int W>
__global__ void __launch_bounds__(256) kCollapseWrite(const BNode* nodes, const uint32_t* level, uint32_t n,
                                                      uint32_t levelBase, const uint32_t* kidOffset, uint32_t nextBase,
                                                      float pad, float* wide, uint32_t* nextLevel) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t kids[W];
    const int k = gatherKids<W>(nodes, level[i], kids);
    float* f = wide + static_cast<size_t>(levelBase + i) * 7 * W;
    uint32_t* child = reinterpret_cast<uint32_t*>(f + 6 * W);
    uint32_t next = kidOffset[i];
    for (int c = 0; c < W; ++c) {
        if (c >= k) {
            f[c] = f[W + c] = f[2 * W + c] = 1.0f;
            f[3 * W + c] = f[4 * W + c] = f[5 * W + c] = -1.0f;
            child[c] = kInvalidChild;
            continue;
        }
        const BNode& nd = nodes[kids[c]];
        f[c] = nd.lo[0] - pad, f[W + c] = nd.lo[1] - pad, f[2 * W + c] = nd.lo[2] - pad;
        f[3 * W + c] = nd.hi[0] + pad, f[4 * W + c] = nd.hi[1] + pad, f[5 * W + c] = nd.hi[2] + pad;
        if (nd.a & kLeafMark) {
            child[c] = kLeafBit | ((nd.b - 1u) << kLeafCountShift) | (nd.a & ~kLeafMark);
        } else {
            child[c] = nextBase + next;
            nextLevel[next] = kids[c];
            ++next;
        }
    }
}

// Leaf-order triangles (v0, prim) (v1, 0) (v2, 0) and their materials.
__global__ void __launch_bounds__(256) kLeafTris(const float* __restrict__ tv, const uint32_t* __restrict__ triMat,
                                                 const uint32_t* __restrict__ prim, uint32_t n, float4* tris,
                                                 uint32_t* matLeaf) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t t = prim[i];
    const float* v = tv + 9 * static_cast<size_t>(t);
    tris[3 * static_cast<size_t>(i)] = make_float4(v[0], v[1], v[2], __uint_as_float(t));
    tris[3 * static_cast<size_t>(i) + 1] = make_float4(v[3], v[4], v[5], 0.0f);
    tris[3 * static_cast<size_t>(i) + 2] = make_float4(v[6], v[7], v[8], 0.0f);
    matLeaf[i] = triMat[t];
}

// Empty bounds and bins of numTasks large tasks: counts 0, minima +inf, maxima -inf (ordered ints).
__global__ void __launch_bounds__(256) kInitTasks(TaskBounds* tb, TaskBins* bins, uint32_t numTasks) {
    constexpr int kInf = 0x7F800000, kNegInf = static_cast<int>(0x807FFFFFu);  // ford(+inf), ford(-inf)
    constexpr uint32_t perBins = sizeof(TaskBins) / 4;  // cnt[3][B], lo[3][B][3], hi[3][B][3]
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < numTasks * 12) tb[i / 12].v[i % 12] = (i % 6) < 3 ? kInf : kNegInf;
    if (i < numTasks * perBins) {
        const uint32_t o = i % perBins;
        reinterpret_cast<int*>(bins)[i] = o < 3 * kMaxBins ? 0 : o < 12 * kMaxBins ? kInf : kNegInf;
    }
}

// Links of the nodes split at this level: links[3t] = node, links[3t + 1, 3t + 2] = children.
__global__ void kLink(const uint32_t* links, uint32_t numTasks, BNode* nodes) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= numTasks) return;
    nodes[links[3 * t]].a = links[3 * t + 1];
    nodes[links[3 * t]].b = links[3 * t + 2];
}

uint32_t blocksOf(size_t n, uint32_t bs) { return static_cast<uint32_t>(std::max<size_t>(1, (n + bs - 1) / bs)); }

// Host vectors handed to toDevice must stay unchanged until the stream next synchronises.
template <typename T>
void toDevice(TmpBuf& b, const std::vector<T>& v, hipStream_t st) {
    b.reset(v.size() * sizeof(T));
    if (!v.empty()) PGB_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
}

template <int W>
void collapse(hipStream_t st, const BNode* nodes, uint32_t numTris, float pad, TmpBuf& wide, uint32_t& numNodes,
              uint32_t& numLeaves, uint32_t& depth) {
    // wide nodes <= binary internal nodes + 1 <= numTris
    wide.reset(static_cast<size_t>(std::max<uint32_t>(numTris, 1)) * 7 * W * 4);
    TmpBuf lvA(static_cast<size_t>(numTris + 1) * 4), lvB(static_cast<size_t>(numTris + 1) * 4);
    TmpBuf inK(static_cast<size_t>(numTris + 1) * 4), lfK(static_cast<size_t>(numTris + 1) * 4),
        off(static_cast<size_t>(numTris + 1) * 4);
    uint32_t zero = 0;
    PGB_HIP(hipMemcpyAsync(lvA.p, &zero, 4, hipMemcpyHostToDevice, st));
    uint32_t levelSize = 1, levelBase = 0;
    numNodes = 1, numLeaves = 0, depth = 0;
    uint32_t* cur = lvA.as<uint32_t>();
    uint32_t* nxt = lvB.as<uint32_t>();
    std::vector<uint32_t> in, lf, ofs;
    while (levelSize > 0) {
        ++depth;
        hipLaunchKernelGGL(kCollapseCount<W>, dim3(blocksOf(levelSize, 256)), dim3(256), 0, st, nodes, cur, levelSize,
                           inK.as<uint32_t>(), lfK.as<uint32_t>());
        in.resize(levelSize), lf.resize(levelSize), ofs.resize(levelSize);
        PGB_HIP(hipMemcpyAsync(in.data(), inK.p, levelSize * 4, hipMemcpyDeviceToHost, st));
        PGB_HIP(hipMemcpyAsync(lf.data(), lfK.p, levelSize * 4, hipMemcpyDeviceToHost, st));
        PGB_HIP(hipStreamSynchronize(st));
        uint32_t total = 0;
        for (uint32_t i = 0; i < levelSize; ++i) ofs[i] = total, total += in[i], numLeaves += lf[i];
        PGB_HIP(hipMemcpyAsync(off.p, ofs.data(), levelSize * 4, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(kCollapseWrite<W>, dim3(blocksOf(levelSize, 256)), dim3(256), 0, st, nodes, cur, levelSize,
                           levelBase, off.as<uint32_t>(), numNodes, pad, wide.as<float>(), nxt);
        PGB_HIP(hipGetLastError());
        PGB_HIP(hipStreamSynchronize(st));  // `ofs` is reused next level
        levelBase = numNodes;
        numNodes += total;
        levelSize = total;
        std::swap(cur, nxt);
    }
}

}  // namespace

void buildWideBvhDevice(hipStream_t st, const float* dTriVerts, const uint32_t* dTriMat, uint32_t numTris,
                        const BvhParams& params, DeviceBvh& out) {
    const uint32_t W = params.width;
    if (W != 4 && W != 8) throw std::invalid_argument("BVH width must be 4 or 8");
    if (params.leafSize < 1 || params.leafSize > kMaxLeafSize) throw std::invalid_argument("leaf size must be 1..8");
    if (numTris == 0 || numTris >= (1u << kLeafCountShift))
        throw std::invalid_argument("device BVH build needs 1..2^27-1 triangles");
    const uint32_t nb = std::max(4u, std::min(kMaxBins, params.bins));
    const uint32_t n = numTris;
    TmpBuf plo(static_cast<size_t>(n) * 16), phi(static_cast<size_t>(n) * 16);
    TmpBuf primA(static_cast<size_t>(n) * 4), primB(static_cast<size_t>(n) * 4);
    TmpBuf nodes(static_cast<size_t>(2 * n) * sizeof(BNode)), maxAbs(4);
    const int one = 0x3F800000;  // ford(1.0f)
    PGB_HIP(hipMemcpyAsync(maxAbs.p, &one, 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(kPrimPrep, dim3(blocksOf(n, 256)), dim3(256), 0, st, dTriVerts, n, plo.as<float4>(),
                       phi.as<float4>(), primA.as<uint32_t>(), maxAbs.as<int>());
    PGB_HIP(hipGetLastError());
    uint32_t* prim = primA.as<uint32_t>();
    uint32_t* prim2 = primB.as<uint32_t>();

    // large tasks, level by level
    std::vector<LargeTask> large;
    std::vector<SmallTask> small;
    uint32_t nodeCount = 1;
    auto place = [&](uint32_t begin, uint32_t count, uint32_t node) {
        if (count > kSmallTask) {
            large.push_back({begin, count, node, 0});
        } else {
            small.push_back({begin, count, node, nodeCount});
            nodeCount += 2 * count - 2;
        }
    };
    place(0, n, 0);
    TmpBuf dTasks, dChunks, dBounds, dBins, dSplits, dChunkLeft, dLeftBefore, dTaskLeft, dLinks;
    std::vector<ChunkRef> chunks;
    std::vector<uint32_t> chunkLeft, leftBefore, taskLeft, links;
    while (!large.empty()) {
        const uint32_t T = static_cast<uint32_t>(large.size());
        chunks.clear();
        for (uint32_t t = 0; t < T; ++t)
            for (uint32_t s = 0; s < large[t].count; s += kChunk)
                chunks.push_back({t, large[t].begin + s, std::min(kChunk, large[t].count - s)});
        const uint32_t C = static_cast<uint32_t>(chunks.size());
        toDevice(dTasks, large, st);
        toDevice(dChunks, chunks, st);
        dBounds.reset(static_cast<size_t>(T) * sizeof(TaskBounds));
        dBins.reset(static_cast<size_t>(T) * sizeof(TaskBins));
        dSplits.reset(static_cast<size_t>(T) * sizeof(Split));
        dChunkLeft.reset(static_cast<size_t>(C) * 4);
        hipLaunchKernelGGL(kInitTasks, dim3(blocksOf(static_cast<size_t>(T) * (sizeof(TaskBins) / 4), 256)), dim3(256), 0,
                           st, dBounds.as<TaskBounds>(), dBins.as<TaskBins>(), T);
        hipLaunchKernelGGL(kLargeBounds, dim3(C), dim3(256), 0, st, dChunks.as<ChunkRef>(), prim, plo.as<float4>(),
                           phi.as<float4>(), dBounds.as<TaskBounds>());
        hipLaunchKernelGGL(kLargeBins, dim3(C), dim3(256), 0, st, dChunks.as<ChunkRef>(), prim, plo.as<float4>(),
                           phi.as<float4>(), dBounds.as<TaskBounds>(), dBins.as<TaskBins>(), nb);
        hipLaunchKernelGGL(kLargeSplit, dim3(blocksOf(T, 64)), dim3(64), 0, st, dTasks.as<LargeTask>(), T,
                           dBounds.as<TaskBounds>(), dBins.as<TaskBins>(), nb, nodes.as<BNode>(), dSplits.as<Split>());
        hipLaunchKernelGGL(kLargeCount, dim3(C), dim3(256), 0, st, dChunks.as<ChunkRef>(), dTasks.as<LargeTask>(),
                           dSplits.as<Split>(), prim, plo.as<float4>(), phi.as<float4>(), nb, dChunkLeft.as<uint32_t>());
        PGB_HIP(hipGetLastError());
        chunkLeft.resize(C);
        PGB_HIP(hipMemcpyAsync(chunkLeft.data(), dChunkLeft.p, C * 4, hipMemcpyDeviceToHost, st));
        PGB_HIP(hipStreamSynchronize(st));
        leftBefore.assign(C, 0);
        taskLeft.assign(T, 0);
        for (uint32_t c = 0; c < C; ++c) {
            leftBefore[c] = taskLeft[chunks[c].task];
            taskLeft[chunks[c].task] += chunkLeft[c];
        }
        // children in task order (left, right); a valid split leaves both sides non-empty
        links.assign(3 * static_cast<size_t>(T), 0);
        std::vector<LargeTask> prev;
        prev.swap(large);
        for (uint32_t t = 0; t < T; ++t) {
            const LargeTask& p = prev[t];
            const uint32_t nl = taskLeft[t];
            if (nl == 0 || nl >= p.count) throw std::runtime_error("device BVH build: empty side in a split");
            const uint32_t l = nodeCount, r = nodeCount + 1;
            nodeCount += 2;
            links[3 * t] = p.node, links[3 * t + 1] = l, links[3 * t + 2] = r;
            place(p.begin, nl, l);
            place(p.begin + nl, p.count - nl, r);
        }
        toDevice(dLeftBefore, leftBefore, st);
        toDevice(dTaskLeft, taskLeft, st);
        toDevice(dLinks, links, st);
        hipLaunchKernelGGL(kLargeScatter, dim3(C), dim3(256), 0, st, dChunks.as<ChunkRef>(), dTasks.as<LargeTask>(),
                           dSplits.as<Split>(), prim, plo.as<float4>(), phi.as<float4>(), nb,
                           dLeftBefore.as<uint32_t>(), dTaskLeft.as<uint32_t>(), prim2);
        hipLaunchKernelGGL(kCopyRange, dim3(C), dim3(256), 0, st, dChunks.as<ChunkRef>(), prim2, prim);
        hipLaunchKernelGGL(kLink, dim3(blocksOf(T, 64)), dim3(64), 0, st, dLinks.as<uint32_t>(), T, nodes.as<BNode>());
        PGB_HIP(hipGetLastError());
        PGB_HIP(hipStreamSynchronize(st));  // the uploads above read host vectors reused next level
    }
    // small subtrees
    if (!small.empty()) {
        TmpBuf dSmall;
        toDevice(dSmall, small, st);
        hipLaunchKernelGGL(kSmall, dim3(static_cast<uint32_t>(small.size())), dim3(64), 0, st, dSmall.as<SmallTask>(),
                           prim, prim2, plo.as<float4>(), phi.as<float4>(), nodes.as<BNode>(), params.leafSize, nb);
        PGB_HIP(hipGetLastError());
        PGB_HIP(hipStreamSynchronize(st));
    }
    int maxAbsBits = 0;
    PGB_HIP(hipMemcpyAsync(&maxAbsBits, maxAbs.p, 4, hipMemcpyDeviceToHost, st));
    PGB_HIP(hipStreamSynchronize(st));
    float maxAbsF;
    std::memcpy(&maxAbsF, &maxAbsBits, 4);  // non-negative: the ordered int is the float's bits
    const float pad = 1e-5f * maxAbsF;

    TmpBuf wide;
    uint32_t numNodes = 0, numLeaves = 0, depth = 0;
    if (W == 8) collapse<8>(st, nodes.as<BNode>(), n, pad, wide, numNodes, numLeaves, depth);
    else collapse<4>(st, nodes.as<BNode>(), n, pad, wide, numNodes, numLeaves, depth);

    out.width = W;
    out.numNodes = numNodes;
    out.numLeaves = numLeaves;
    out.maxDepth = depth;
    out.pad = pad;
    void* dNodes = out.alloc(DeviceBvh::kNodes, static_cast<size_t>(numNodes) * 7 * W * 4);
    void* dTris = out.alloc(DeviceBvh::kTris, static_cast<size_t>(n) * 48);
    void* dMat = out.alloc(DeviceBvh::kTriMat, static_cast<size_t>(n) * 4);
    PGB_HIP(hipMemcpyAsync(dNodes, wide.p, static_cast<size_t>(numNodes) * 7 * W * 4, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(kLeafTris, dim3(blocksOf(n, 256)), dim3(256), 0, st, dTriVerts, dTriMat, prim, n,
                       static_cast<float4*>(dTris), static_cast<uint32_t*>(dMat));
    PGB_HIP(hipGetLastError());
    PGB_HIP(hipStreamSynchronize(st));
}

}  // namespace pg

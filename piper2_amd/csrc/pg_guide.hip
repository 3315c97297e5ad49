// Guiding network kernels for gfx950: inference, fused forward+backward, split-K weight gradients
// on the matrix cores, and a deterministic reduce + Adam. Network layout: pg_guide.hpp.
#include "pg_device.hpp"
#include "pg_guide.hpp"
#include "pg_guide_kernels.hpp"

namespace pg {

// ------------------------------------------------------------------------------------------
// Inputs of one sample
// ------------------------------------------------------------------------------------------
// A source hands out each lane's sample in stages so that its dependent loads can be in flight
// while the wave computes the previous samples: begin(i) issues the first loads of entry i (~0u:
// none), stage2 .. stage4 the loads that depend on them, finish() turns them into the inputs.
__device__ __forceinline__ GuideInput noInput() {
    return guideInput(make_float3(0, 0, 0), make_float3(0, 0, 1), make_float3(0, 0, 1));
}
// Inputs at a hit: position in the scene box mapped to [0,1]^3, the geometric normal facing the
// ray (as triNormal) and the outgoing direction.
template <bool MOTION>
__device__ __forceinline__ GuideInput hitInput(float4 oo, float4 dd, float t, float4 a, float4 b, float4 c, float3 lo,
                                               float3 invExt, const DevInstance* inst) {
    const V3 dir{dd.x, dd.y, dd.z};
    const V3 p = V3{oo.x, oo.y, oo.z} + dir * t;
    V3 v0, v1, v2;  // a moving instance's triangle at the path's time (d.w)
    worldTriangle<MOTION>(inst, a, b, c, dd.w, v0, v1, v2);
    const V3 ng = fnormalize(cross(v1 - v0, v2 - v0));
    const V3 n = dot(dir, ng) < 0.0f ? ng : -ng;
    const float3 p01 = make_float3(fminf(fmaxf((p.x - lo.x) * invExt.x, 0.0f), 1.0f),
                                   fminf(fmaxf((p.y - lo.y) * invExt.y, 0.0f), 1.0f),
                                   fminf(fmaxf((p.z - lo.z) * invExt.z, 0.0f), 1.0f));
    return guideInput(p01, make_float3(n.x, n.y, n.z), make_float3(-dir.x, -dir.y, -dir.z));
}

// From an explicit array: 3 float4 per sample = (pos01, -) (normal, -) (wo, -).
struct ArraySource {
    const float4* in;
    struct Pend {
        float4 p, n, w;
        bool valid;
    };
    __device__ __forceinline__ Pend begin(uint32_t i) const {
        Pend q{};
        q.valid = i != ~0u;
        if (q.valid) q.p = in[3 * i], q.n = in[3 * i + 1], q.w = in[3 * i + 2];
        return q;
    }
    __device__ __forceinline__ void stage2(Pend&) const {}
    __device__ __forceinline__ void stage3(Pend&) const {}
    __device__ __forceinline__ void stage4(Pend&) const {}
    __device__ __forceinline__ GuideInput finish(const Pend& q) const {
        if (!q.valid) return noInput();
        return guideInput(make_float3(q.p.x, q.p.y, q.p.z), make_float3(q.n.x, q.n.y, q.n.z),
                          make_float3(q.w.x, q.w.y, q.w.z));
    }
};
// From the path state of an extension-queue entry (hit point, shading normal, -ray direction):
// queue entry -> path record -> hit triangle. MOTION: the scene has moving instances.
template <bool MOTION>
struct PathSource {
    const float4* tris;  // leaf-order triangles of the scene (only the hit triangle's normal is needed)
    const PathRec* paths;
    const uint32_t* queue;
    float3 lo, invExt;
    const DevInstance* inst;  // the scene's moving instances
    struct Pend {
        uint32_t path;
        float4 h, o, d, a, b, c;
        bool valid, tri;
    };
    __device__ __forceinline__ Pend begin(uint32_t i) const {
        Pend q{};
        q.valid = i != ~0u;
        if (q.valid) q.path = queue[i];
        return q;
    }
    __device__ __forceinline__ void stage2(Pend& q) const {
        if (q.valid) q.h = paths[q.path].hit, q.o = paths[q.path].o, q.d = paths[q.path].d;
    }
    __device__ __forceinline__ void stage3(Pend& q) const {
        const uint32_t pos = __float_as_uint(q.h.y);
        q.tri = q.valid && pos != kHitMiss && !(pos & kHitQuad);
        if (q.tri) q.a = tris[3 * pos], q.b = tris[3 * pos + 1], q.c = tris[3 * pos + 2];
    }
    __device__ __forceinline__ void stage4(Pend&) const {}
    __device__ __forceinline__ GuideInput finish(const Pend& q) const {
        if (!q.tri) return noInput();
        return hitInput<MOTION>(q.o, q.d, q.h.x, q.a, q.b, q.c, lo, invExt, inst);
    }
};

// Mixture of lobes 2g, 2g+1 of this lane's sample, decoded from the lane's own output accumulators
// (output slot u = 4 mt + j is row 16 mt + 4 g + j; slots 5q..5q+4 are lobe 2g + q's mu xyz, log
// kappa and weight logit). Only the softmax over the 8 weights crosses lanes (s, s+16, s+32, s+48).
struct LobePair {
    float mx[2], my[2], mz[2], mn[2], r[2], kappa[2], pi[2];
    float z;  // logit of the learned selection probability (lane group 0's slot 10, on every lane)
};
__device__ __forceinline__ float outSlot(const f32x4 (&out)[3], int u) { return out[u >> 2][u & 3]; }
__device__ __forceinline__ LobePair lobesFromTile(const f32x4 (&out)[3]) {
    LobePair L;
    float a[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const float ox = outSlot(out, 5 * q), oy = outSlot(out, 5 * q + 1), oz = outSlot(out, 5 * q + 2);
        const float n = fmaxf(__builtin_amdgcn_sqrtf(ox * ox + oy * oy + oz * oz), 1e-6f);
        const float inv = __builtin_amdgcn_rcpf(n);  // 1 ulp: the stored lobe is what sampling and pdf use
        L.mn[q] = n;
        L.mx[q] = ox * inv, L.my[q] = oy * inv, L.mz[q] = oz * inv;
        L.r[q] = outSlot(out, 5 * q + 3);
        L.kappa[q] = __expf(fminf(fmaxf(L.r[q], kGKappaLo), kGKappaHi));
        a[q] = outSlot(out, 5 * q + 4);
    }
    float m = fmaxf(a[0], a[1]);
    m = fmaxf(m, __shfl_xor(m, 16));
    m = fmaxf(m, __shfl_xor(m, 32));
    const float e0 = __expf(a[0] - m), e1 = __expf(a[1] - m);
    float sum = e0 + e1;
    sum += __shfl_xor(sum, 16);
    sum += __shfl_xor(sum, 32);
    const float rs = __builtin_amdgcn_rcpf(sum);
    L.pi[0] = e0 * rs, L.pi[1] = e1 * rs;
    L.z = __shfl(outSlot(out, kGAlphaSlot), static_cast<int>(__lane_id() & 15u));
    return L;
}

// Orders one wave's LDS stores before its lanes' later LDS loads (no block barrier: other waves of
// the block keep running).
__device__ __forceinline__ void waveLdsSync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------------------------------------
// inference: mixture records for n samples
// ------------------------------------------------------------------------------------------
// Each wave takes 64 samples at a time: every lane gathers one sample's inputs (one dependent load
// chain per 64 samples) into the wave's LDS stage, then the wave runs four 16-sample MFMA tiles.
constexpr int kInferStageBytes = 64 * 3 * 16;  // per wave: 64 samples x 3 float4
// 8 waves share one LDS copy of the weights: 54 KB + 8 x 3 KB, two blocks per CU (4 waves per SIMD)
constexpr int kInferBlock = 512;
static_assert(kInferBlock == static_cast<int>(kGroupUnit), "a grouped unit is one block's 8 x 64 samples");

// Stores the mixture of this lane's lobes (2g, 2g + 1) of sample `i` (~0u: none).
__device__ __forceinline__ void storeLobes(uint32_t i, const LobePair& lp, int grp, float* __restrict__ mix) {
    if (i == ~0u) return;
    float* o = mix + static_cast<size_t>(i) * kGMixFloats;
    reinterpret_cast<float4*>(o)[2 * grp] = make_float4(lp.mx[0], lp.my[0], lp.mz[0], lp.kappa[0]);
    reinterpret_cast<float4*>(o)[2 * grp + 1] = make_float4(lp.mx[1], lp.my[1], lp.mz[1], lp.kappa[1]);
    reinterpret_cast<float2*>(o + 32)[grp] = make_float2(lp.pi[0], lp.pi[1]);
    reinterpret_cast<float2*>(o + 40)[grp] =
        make_float2(lp.pi[0] * vmfNormSampled(lp.kappa[0]), lp.pi[1] * vmfNormSampled(lp.kappa[1]));
    if (grp == 0) reinterpret_cast<float4*>(o)[12] = make_float4(alphaFromLogit(lp.z), 0.0f, 0.0f, 0.0f);
}

// Mixtures of one wave's 64 samples: lane l's sample is queue/array entry `entry` (~0u: none) with
// inputs `in`. The samples go through the wave's LDS stage so that tile t (samples 16t..16t+15) is
// spread over all 64 lanes (lane group g computes features 4g.. of sample 16t + l % 16); per pass
// of T tiles (4 / T passes) each tile's features feed its layer 0 at once, then layers 1..4 run the
// T tiles side by side. While they run, the lane's next entry
// `next` goes through the source's load stages; returns its inputs. Wave-uniform; the stage is the
// wave's own LDS slice.
template <int T, typename Src>
__device__ __forceinline__ GuideInput inferWave64(const Src& src, uint32_t entry, const GuideInput& in, uint32_t next,
                                                  const GuideLdsX& L, float4* stage, float* __restrict__ mix) {
    static_assert(T == 1 || T == 2 || T == 4, "tiles per pass");
    const int lane = __lane_id(), s = lane & 15, grp = lane >> 4;
    stage[3 * lane] = make_float4(in.px, in.py, in.pz, in.nx);
    stage[3 * lane + 1] = make_float4(in.ny, in.nz, in.wx, in.wy);
    stage[3 * lane + 2] = make_float4(in.wz, __uint_as_float(entry), 0.0f, 0.0f);
    const uint64_t live = __ballot(entry != ~0u);
    waveLdsSync();  // the wave's own stage: written and read by its lanes only
    typename Src::Pend p = src.begin(next);
    // The next samples' dependent loads are spread over the work: stage 2 after the first tile's
    // layer 0, stage 3 after the third's, stage 4 after the last hidden layers.
    constexpr uint64_t kPassLanes = T == 4 ? ~0ull : (1ull << (16 * T)) - 1ull;  // a pass's samples in `live`
#pragma unroll
    for (int pass = 0; pass < 4 / T; ++pass) {
        const bool run = (live >> (16 * T * pass)) & kPassLanes;  // wave-uniform
        f32x4 acc[T][4];
        uint32_t idx[T];
#pragma unroll
        for (int t = 0; t < T; ++t) {  // features and layer 0 tile by tile (one tile's features live)
            const int tile = T * pass + t;
            if (run) {
                const int k = 16 * tile + s;
                const float4 a = stage[3 * k], b = stage[3 * k + 1], c = stage[3 * k + 2];
                idx[t] = __float_as_uint(c.y);
                bf16x4 x0[9];
                guideFeatureTiles(GuideInput{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x}, grp, x0);
                guideLayer0X32(L, x0, acc[t]);
            }
            if (tile == 0) src.stage2(p);
            if (tile == 2) src.stage3(p);
        }
        f32x4 out[T][3];
        if (run) guideHiddenX32<T>(L, acc, out);
        if (pass == 4 / T - 1) src.stage4(p);
        if (run)
#pragma unroll
            for (int t = 0; t < T; ++t) storeLobes(idx[t], lobesFromTile(out[t]), grp, mix);
    }
    const GuideInput nin = src.finish(p);
    waveLdsSync();  // stage reads done before the next samples overwrite it
    return nin;
}

// Tiles per pass of a block of `waves` waves: 4 (about 150 registers) where the block size allows;
// a 16-wave block (1024 threads) caps the registers at 128, which one tile at a time fits.
constexpr int inferTiles(int waves) { return waves >= 16 ? 1 : 4; }

// WAVES: waves per block, sharing one LDS copy of the weights (a launch setting).
template <typename Src, int WAVES>
__global__ void __launch_bounds__(WAVES * 64) kGuideInfer(Src src, uint32_t nHost, const uint32_t* nDev,
                                                           const uint16_t* __restrict__ wb, const float* __restrict__ theta,
                                                           float* __restrict__ mix) {
    extern __shared__ unsigned char smem[];
    const uint32_t n = nDev ? *nDev : nHost;
    const GuideLdsX L = guideLdsXLayout(smem);
    float4* stage = reinterpret_cast<float4*>(smem + kGLdsXWeightBytes + (threadIdx.x / 64) * kInferStageBytes);
    guideLoadLdsX(L, wb, theta);
    __syncthreads();  // the weights are shared by the block's waves; after this each wave runs alone
    const uint32_t lane = __lane_id();
    const uint32_t wave = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64, waves = gridDim.x * (blockDim.x / 64);
    auto entryAt = [&](uint32_t base) { return base < n && base + lane < n ? base + lane : ~0u; };
    uint32_t base = wave * 64, entry = entryAt(base);
    typename Src::Pend p = src.begin(entry);
    src.stage2(p), src.stage3(p), src.stage4(p);
    GuideInput in = src.finish(p);
    for (; base < n; base += waves * 64) {  // wave-uniform
        const uint32_t next = entryAt(base + waves * 64);
        in = inferWave64<inferTiles(WAVES)>(src, entry, in, next, L, stage, mix);
        entry = next;
    }
}

// ---- guided batches: inference grouped by guide --------------------------------------------------
// Queue entries of a guided batch: path slot -> item -> its scene's triangles and box.
template <bool MOTION>
struct BatchPathSource {
    const DevScene* scenes;
    const ItemGuide* guide;
    const PathRec* paths;
    const uint32_t* queue;
    __device__ __forceinline__ uint32_t key(uint32_t i) const { return guide[paths[queue[i]].meta.z >> 16].key; }
    // queue entry -> path record (with its item) -> the item's scene and box -> hit triangle
    struct Pend {
        uint32_t path, item;
        float4 h, o, d, a, b, c;
        const float4* tris;
        const DevInstance* inst;
        float3 lo, invExt;
        bool valid, tri;
    };
    __device__ __forceinline__ Pend begin(uint32_t i) const {
        Pend q{};
        q.valid = i != ~0u;
        if (q.valid) q.path = queue[i];
        return q;
    }
    __device__ __forceinline__ void stage2(Pend& q) const {
        if (q.valid)
            q.h = paths[q.path].hit, q.o = paths[q.path].o, q.d = paths[q.path].d, q.item = paths[q.path].meta.z >> 16;
    }
    __device__ __forceinline__ void stage3(Pend& q) const {
        if (q.valid)
            q.tris = scenes[q.item].tris, q.inst = scenes[q.item].inst, q.lo = guide[q.item].lo,
            q.invExt = guide[q.item].invExt;
    }
    __device__ __forceinline__ void stage4(Pend& q) const {
        const uint32_t pos = __float_as_uint(q.h.y);
        q.tri = q.valid && pos != kHitMiss && !(pos & kHitQuad);
        if (q.tri) q.a = q.tris[3 * pos], q.b = q.tris[3 * pos + 1], q.c = q.tris[3 * pos + 2];
    }
    __device__ __forceinline__ GuideInput finish(const Pend& q) const {
        if (!q.tri) return noInput();
        return hitInput<MOTION>(q.o, q.d, q.h.x, q.a, q.b, q.c, q.lo, q.invExt, q.inst);
    }
};

// Contiguous share [lo, hi) of the n queue entries for sort block blockIdx.x (of kSortBlocks).
__device__ __forceinline__ void sortShare(uint32_t n, uint32_t& lo, uint32_t& hi) {
    const uint32_t chunk = (n + kSortBlocks - 1) / kSortBlocks;
    lo = min(n, blockIdx.x * chunk);
    hi = min(n, lo + chunk);
}

__global__ void __launch_bounds__(256) kGroupCount(BatchPathSource<false> src, const uint32_t* count, uint32_t nKeys,
                                                   GroupBuffers gb) {
    __shared__ uint32_t hist[kMaxBatchGuides];
    for (uint32_t k = threadIdx.x; k < nKeys; k += blockDim.x) hist[k] = 0;
    __syncthreads();
    uint32_t lo, hi;
    sortShare(*count, lo, hi);
    for (uint32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) atomicAdd(&hist[src.key(i)], 1u);  // LDS
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < nKeys; k += blockDim.x) gb.blockCount[k * kSortBlocks + blockIdx.x] = hist[k];
}

// One block of kSortBlocks threads (thread b = sort block b): per key, the exclusive scan of the
// blocks' counts (wave scans, then the waves' totals); keys laid out one after another, each group
// padded to kGroupUnit entries.
__global__ void __launch_bounds__(kSortBlocks) kGroupScan(uint32_t nKeys, GroupBuffers gb) {
    static_assert(kSortBlocks == 256 && kMaxBatchGuides == 256, "one thread per sort block and per key");
    __shared__ uint32_t waveSum[4][kMaxBatchGuides];
    __shared__ uint32_t keyBase[kMaxBatchGuides];
    const uint32_t b = threadIdx.x, lane = __lane_id(), wave = b / 64;
    for (uint32_t k = 0; k < nKeys; ++k) {
        const uint32_t c = gb.blockCount[k * kSortBlocks + b];
        uint32_t x = c;  // inclusive scan within the wave
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if (lane >= static_cast<uint32_t>(off)) x += y;
        }
        gb.blockBase[k * kSortBlocks + b] = x - c;
        if (lane == 63) waveSum[wave][k] = x;
    }
    __syncthreads();
    if (b < nKeys) {  // thread k: the waves' offsets within key k, and key k's total
        uint32_t acc = 0;
        for (int w = 0; w < 4; ++w) {
            const uint32_t t = waveSum[w][b];
            waveSum[w][b] = acc;
            acc += t;
        }
        keyBase[b] = acc;
    }
    __syncthreads();
    if (b == 0) {  // group starts, each group padded to whole units
        uint32_t start = 0;
        for (uint32_t k = 0; k < nKeys; ++k) {
            const uint32_t total = keyBase[k];
            gb.keyStart[k] = start;
            gb.keyEnd[k] = start + total;
            keyBase[k] = start;
            start += (total + kGroupUnit - 1) / kGroupUnit * kGroupUnit;
        }
        gb.keyStart[nKeys] = start;
    }
    __syncthreads();
    for (uint32_t k = 0; k < nKeys; ++k) gb.blockBase[k * kSortBlocks + b] += keyBase[k] + waveSum[wave][k];
}

// The same shares as kGroupCount; inside a block the entries of one key take consecutive slots
// from the block's base (LDS atomics: the order inside a group does not matter).
__global__ void __launch_bounds__(256) kGroupScatter(BatchPathSource<false> src, const uint32_t* count, uint32_t nKeys,
                                                     GroupBuffers gb) {
    __shared__ uint32_t next[kMaxBatchGuides];
    for (uint32_t k = threadIdx.x; k < nKeys; k += blockDim.x) next[k] = gb.blockBase[k * kSortBlocks + blockIdx.x];
    __syncthreads();
    uint32_t lo, hi;
    sortShare(*count, lo, hi);
    for (uint32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) gb.perm[atomicAdd(&next[src.key(i)], 1u)] = i;
}

// Each block takes a contiguous range of units (kGroupUnit entries of one guide each) and reloads
// the weights only when the guide changes; the order of entries inside a group does not matter
// (every mixture depends on its own entry only).
template <bool MOTION>
__global__ void __launch_bounds__(kInferBlock) kGuideInferGrouped(BatchPathSource<MOTION> src, uint32_t nKeys, GroupBuffers gb,
                                                                  const uint16_t* const* __restrict__ keyWb,
                                                                  const float* const* __restrict__ keyTheta,
                                                                  float* __restrict__ mix) {
    extern __shared__ unsigned char smem[];
    const GuideLdsX L = guideLdsXLayout(smem);
    float4* stage = reinterpret_cast<float4*>(smem + kGLdsXWeightBytes + (threadIdx.x / 64) * kInferStageBytes);
    const uint32_t units = gb.keyStart[nKeys] / kGroupUnit;
    const uint32_t per = (units + gridDim.x - 1) / gridDim.x;
    const uint32_t u0 = blockIdx.x * per, u1 = min(units, u0 + per);
    const uint32_t lane = __lane_id(), wave = threadIdx.x / 64;
    // the guide of unit u (keys in unit order, so one forward scan per block)
    uint32_t scanKey = 0;
    auto keyOf = [&](uint32_t u) {
        while (scanKey + 1 < nKeys && gb.keyStart[scanKey + 1] <= u * kGroupUnit) ++scanKey;
        return scanKey;
    };
    auto entryOf = [&](uint32_t u, uint32_t k) {
        const uint32_t pos = u * kGroupUnit + wave * 64 + lane;
        return pos < gb.keyEnd[k] ? gb.perm[pos] : ~0u;
    };
    uint32_t loaded = ~0u;
    uint32_t key = u0 < u1 ? keyOf(u0) : 0, entry = u0 < u1 ? entryOf(u0, key) : ~0u;
    typename BatchPathSource<MOTION>::Pend p = src.begin(entry);
    src.stage2(p), src.stage3(p), src.stage4(p);
    GuideInput in = src.finish(p);
    for (uint32_t u = u0; u < u1; ++u) {  // block-uniform
        if (key != loaded) {
            __syncthreads();  // every wave is done with the previous weights
            guideLoadLdsX(L, keyWb[key], keyTheta[key]);
            __syncthreads();
            loaded = key;
        }
        const uint32_t nextKey = u + 1 < u1 ? keyOf(u + 1) : key;
        const uint32_t next = u + 1 < u1 ? entryOf(u + 1, nextKey) : ~0u;
        in = inferWave64<inferTiles(kInferBlock / 64)>(src, entry, in, next, L, stage, mix);
        key = nextKey, entry = next;
    }
}

// ------------------------------------------------------------------------------------------
// training, part 1: forward with saved activations, loss gradient, backward data path.
// Training record (4 float4): (pos01, c) (normal, -) (wo, -) (sampled direction, -).
// Stores X_l (layer inputs) and dZ_l (pre-activation gradients) feature-major, BF16, [rows][npad].
// ------------------------------------------------------------------------------------------
struct TrainBuffers {
    uint16_t* x0;  // 144 rows
    uint16_t* xh;  // 4 x 64 rows: inputs of layers 1..4
    uint16_t* dz;  // 4 x 64 rows (layers 0..3) + 48 rows (layer 4)
    uint32_t npad;
    uint32_t learnAlpha;  // loss of the one-sample MIS density with the learned selection probability
};

__device__ __forceinline__ void storeRows(uint16_t* dst, uint32_t npad, int row0, int grp, uint32_t col, bf16x4 v) {
#pragma unroll
    for (int j = 0; j < 4; ++j) dst[static_cast<size_t>(row0 + 4 * grp + j) * npad + col] = static_cast<uint16_t>(v[j]);
}

// LDS: forward weights, then transposed W1..W4 ([in = 64][out + 4]).
constexpr int kGLdsTransBytes = (3 * kGHid * kGLdH + kGHid * kGLdOT) * 2;

__global__ void __launch_bounds__(256) kGuideTrainFB(const float4* __restrict__ rec, uint32_t n,
                                                     const uint16_t* __restrict__ wb, const float* __restrict__ theta,
                                                     const float* __restrict__ csum, TrainBuffers tb, float* lossAcc) {
    extern __shared__ unsigned char smem[];
    const GuideLds L = guideLdsLayout(smem);
    uint16_t* wt = reinterpret_cast<uint16_t*>(smem + kGLdsWeightBytes);  // W1T, W2T, W3T [64][kGLdH], W4T [64][kGLdOT]
    guideLoadLds(L, wb, theta);
    for (int l = 0; l < 3; ++l) {
        const int off = l == 0 ? kGW1 : (l == 1 ? kGW2 : kGW3);
        for (int i = threadIdx.x; i < kGHid * kGHid; i += blockDim.x) {  // W[o][k] -> WT[k][o]
            const int o = i / kGHid, k = i % kGHid;
            wt[l * kGHid * kGLdH + k * kGLdH + o] = wb[off + i];
        }
    }
    for (int i = threadIdx.x; i < kGOut * kGHid; i += blockDim.x) {
        const int o = i / kGHid, k = i % kGHid;
        wt[3 * kGHid * kGLdH + k * kGLdOT + o] = wb[kGW4 + i];
    }
    __syncthreads();
    const float invC = 1.0f / fmaxf(*csum, 1e-30f);
    const int lane = __lane_id(), s = lane & 15, grp = lane >> 4;
    float lossPart = 0.0f;
    for (uint32_t base = blockIdx.x * 64; base < tb.npad; base += gridDim.x * 64) {  // block-uniform
        const uint32_t i = base + (threadIdx.x / 64) * 16 + s;
        const bool valid = i < n;
        float4 r0 = make_float4(0, 0, 0, 0), r1 = make_float4(0, 0, 1, 0), r2 = make_float4(0, 0, 1, 0),
               r3 = make_float4(0, 0, 1, 0);
        if (valid) r0 = rec[4 * i], r1 = rec[4 * i + 1], r2 = rec[4 * i + 2], r3 = rec[4 * i + 3];
        const GuideInput in = guideInput(make_float3(r0.x, r0.y, r0.z), make_float3(r1.x, r1.y, r1.z),
                                         make_float3(r2.x, r2.y, r2.z));
        const float c = valid ? r0.w * invC : 0.0f;
        bf16x4 x0[9];
        guideFeatureTiles(in, grp, x0);
#pragma unroll
        for (int kt = 0; kt < 9; ++kt) storeRows(tb.x0, tb.npad, kt * 16, grp, i, x0[kt]);
        bf16x4 h[4][4];
        f32x4 out[3];
        guideForwardTile(L, x0, out, h);
#pragma unroll
        for (int l = 0; l < 4; ++l)
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) storeRows(tb.xh + static_cast<size_t>(l) * kGHid * tb.npad, tb.npad, mt * 16, grp, i, h[l][mt]);
        // mixture, loss and gradient with respect to this lane's output slots (lobes 2g, 2g+1)
        const LobePair lp = lobesFromTile(out);
        const float wx = r3.x, wy = r3.y, wz = r3.z;
        float v[2], cs[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            cs[q] = lp.mx[q] * wx + lp.my[q] * wy + lp.mz[q] * wz;
            v[q] = vmfPdf(lp.kappa[q], cs[q]);
        }
        float p = lp.pi[0] * v[0] + lp.pi[1] * v[1];
        p += __shfl_xor(p, 16);
        p += __shfl_xor(p, 32);
        p = fmaxf(p, 1e-30f);
        // learned selection: loss of q = alpha p + (1 - alpha) p_bsdf, alpha = alphaFromLogit(z)
        float dLdp, dz10 = 0.0f;
        if (tb.learnAlpha) {
            const float sg = __builtin_amdgcn_rcpf(1.0f + __expf(-lp.z));
            const float al = kGAlphaLo + (kGAlphaHi - kGAlphaLo) * sg;
            const float pb = fmaxf(r1.w, 0.0f);
            const float qd = fmaxf(al * p + (1.0f - al) * pb, 1e-30f);
            if (grp == 0) lossPart += -c * logf(qd);
            const float dLdq = -c / qd;
            dLdp = dLdq * al;
            dz10 = dLdq * (p - pb) * (kGAlphaHi - kGAlphaLo) * sg * (1.0f - sg);
        } else {
            if (grp == 0) lossPart += -c * logf(p);
            dLdp = -c / p;
        }
        // softmax: dL/da_i = pi_i (g_i - sum_j pi_j g_j), g_i = dL/dp v_i
        float gpi = lp.pi[0] * dLdp * v[0] + lp.pi[1] * dLdp * v[1];
        gpi += __shfl_xor(gpi, 16);
        gpi += __shfl_xor(gpi, 32);
        float gs[12];  // dL/d(output slot u); slot 10 of lane group 0 is the selection logit, the rest zero rows
        gs[10] = grp == 0 ? dz10 : 0.0f;
        gs[11] = 0.0f;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const float dv = dLdp * lp.pi[q];
            const float k = lp.kappa[q];
            const float dlogC = 1.0f / k - 2.0f / expm1f(2.0f * k);
            const float dk = dv * v[q] * (dlogC + cs[q] - 1.0f);
            const float dr = (lp.r[q] > kGKappaLo && lp.r[q] < kGKappaHi) ? dk * k : 0.0f;
            const float sc = dv * v[q] * k;  // dL/dmu = sc * w
            const float dmx = sc * wx, dmy = sc * wy, dmz = sc * wz;
            const float proj = lp.mx[q] * dmx + lp.my[q] * dmy + lp.mz[q] * dmz;
            float* o = gs + 5 * q;
            o[0] = (dmx - lp.mx[q] * proj) / lp.mn[q];
            o[1] = (dmy - lp.my[q] * proj) / lp.mn[q];
            o[2] = (dmz - lp.mz[q] * proj) / lp.mn[q];
            o[3] = dr;
            o[4] = lp.pi[q] * (dLdp * v[q] - gpi);
        }
        // backward data path: dZ4 -> dZ3 ... dZ0
        bf16x4 dz[4];
#pragma unroll
        for (int mt = 0; mt < 3; ++mt) {
            const f32x4 t{gs[4 * mt], gs[4 * mt + 1], gs[4 * mt + 2], gs[4 * mt + 3]};
            dz[mt] = pack(t);
            storeRows(tb.dz + 4ull * kGHid * tb.npad, tb.npad, mt * 16, grp, i, dz[mt]);
        }
#pragma unroll
        for (int l = 3; l >= 0; --l) {
            // dH_l (64 x 16) = W_{l+1}^T dZ_{l+1}; W^T rows = this layer's units
            const uint16_t* w = wt + l * kGHid * kGLdH;
            const int ld = l == 3 ? kGLdOT : kGLdH;
            const int kts = l == 3 ? 3 : 4;
            f32x4 acc[4];
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                acc[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int kt = 0; kt < 4; ++kt)
                    if (kt < kts) acc[mt] = mfma16(ldsA(w, ld, mt * 16 + s, kt * 16 + 4 * grp), dz[kt], acc[mt]);
            }
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                f32x4 t;
#pragma unroll
                for (int j = 0; j < 4; ++j)  // ReLU'(z) from h = relu(z): nonzero magnitude
                    t[j] = (static_cast<uint16_t>(h[l][mt][j]) & 0x7FFFu) != 0 ? acc[mt][j] : 0.0f;
                dz[mt] = pack(t);
                storeRows(tb.dz + static_cast<size_t>(l) * kGHid * tb.npad, tb.npad, mt * 16, grp, i, dz[mt]);
            }
        }
    }
    if (lossAcc) {
        float t = lossPart;
        for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
        if (lane == 0) atomicAdd(lossAcc, t);
    }
}

// ------------------------------------------------------------------------------------------
// training, part 2: dW_l = dZ_l X_l^T over the batch (split K = samples), bias sums; partials per
// sample chunk. One wave = one 16x16 output tile over one chunk.
// ------------------------------------------------------------------------------------------

__global__ void __launch_bounds__(64) kGuideWGrad(const WGradTile* __restrict__ tiles, uint32_t npad, uint32_t chunk,
                                                  float* __restrict__ partial) {
    const WGradTile t = tiles[blockIdx.x];
    const uint32_t c = blockIdx.y;
    const int lane = threadIdx.x, s = lane & 15, grp = lane >> 4;
    const uint16_t* dzRow = t.dz + static_cast<size_t>(t.mo * 16 + s) * npad;
    const uint16_t* xRow = t.x + static_cast<size_t>(t.ki * 16 + s) * npad;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    float bsum = 0.0f;
    const uint32_t n0 = c * chunk, n1 = n0 + chunk;
    for (uint32_t n = n0; n < n1; n += 16) {
        const bf16x4 a = *reinterpret_cast<const bf16x4*>(dzRow + n + 4 * grp);
        const bf16x4 b = *reinterpret_cast<const bf16x4*>(xRow + n + 4 * grp);
        acc = mfma16(a, b, acc);
        if (t.ki == 0)
#pragma unroll
            for (int j = 0; j < 4; ++j) bsum += fromBf16(static_cast<uint16_t>(a[j]));
    }
    float* out = partial + static_cast<size_t>(c) * kGParams;
#pragma unroll
    for (int j = 0; j < 4; ++j) out[t.wOff + (t.mo * 16 + 4 * grp + j) * t.cols + t.ki * 16 + s] = acc[j];
    if (t.ki == 0) {
        bsum += __shfl_xor(bsum, 16);
        bsum += __shfl_xor(bsum, 32);
        if (grp == 0) out[t.bOff + t.mo * 16 + s] = bsum;
    }
}

// ------------------------------------------------------------------------------------------
// training, part 3: sum the chunk partials in a fixed order; optionally Adam-update the FP32
// masters and refresh the BF16 copy.
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) kGuideAdam(const float* __restrict__ partial, uint32_t chunks, float* theta,
                                                  float* m, float* v, uint16_t* wb, float* gradOut, AdamParams ap) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kGParams) return;
    float g = 0.0f;
    for (uint32_t c = 0; c < chunks; ++c) g += partial[static_cast<size_t>(c) * kGParams + p];
    if (gradOut) gradOut[p] = g;
    if (!ap.update) return;
    const float mm = ap.b1 * m[p] + (1.0f - ap.b1) * g;
    const float vv = ap.b2 * v[p] + (1.0f - ap.b2) * g * g;
    m[p] = mm, v[p] = vv;
    const float mh = mm / ap.c1, vh = vv / ap.c2;
    const float th = theta[p] - ap.lr * mh / (sqrtf(vh) + ap.eps);
    theta[p] = th;
    wb[p] = toBf16(th);
}

__global__ void kGuideToBf16(const float* theta, uint16_t* wb) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kGParams) wb[p] = toBf16(theta[p]);
}

// Deterministic sum of the records' weights c (one block).
__global__ void __launch_bounds__(256) kGuideSumC(const float4* rec, uint32_t n, float* out) {
    __shared__ float part[256];
    float t = 0.0f;
    for (uint32_t i = threadIdx.x; i < n; i += 256) t += rec[4 * i].w;
    part[threadIdx.x] = t;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = part[0];
}

// ------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------
constexpr size_t inferLds(int waves) { return kGLdsXWeightBytes + static_cast<size_t>(waves) * kInferStageBytes; }
constexpr size_t kInferLds = inferLds(kInferBlock / 64);
static_assert(2 * kInferLds <= 160 * 1024, "two inference blocks (16 waves) fit one CU's LDS");
static_assert(inferLds(16) <= 160 * 1024, "a 16-wave inference block fits one CU's LDS");
// blocks of `waves` waves one CU's 160 KB of LDS holds
constexpr uint32_t inferBlocksPerCu(int waves) { return static_cast<uint32_t>((160 * 1024) / inferLds(waves)); }
constexpr size_t kTrainLds = kGLdsWeightBytes + kGLdsTransBytes;

static uint32_t gridFor(uint32_t n, uint32_t per, uint32_t cap) {
    const uint32_t b = (n + per - 1) / per;
    return b < 1 ? 1 : (b > cap ? cap : b);
}

void launchGuideInferArray(hipStream_t st, const float* in, uint32_t n, const uint16_t* wb, const float* theta,
                           float* mix, uint32_t maxBlocks) {
    hipLaunchKernelGGL((kGuideInfer<ArraySource, kInferBlock / 64>), dim3(gridFor(n, kInferBlock, maxBlocks)),
                       dim3(kInferBlock), kInferLds, st, ArraySource{reinterpret_cast<const float4*>(in)}, n, nullptr,
                       wb, theta, mix);
}

template <int WAVES, typename Src>
static void inferPaths(hipStream_t st, const Src& src, const uint32_t* count, uint32_t capacity, const uint16_t* wb,
                       const float* theta, float* mix, uint32_t cus) {
    const dim3 grid(gridFor(capacity, WAVES * 64, cus * inferBlocksPerCu(WAVES)));
    hipLaunchKernelGGL((kGuideInfer<Src, WAVES>), grid, dim3(WAVES * 64), inferLds(WAVES), st, src, capacity, count, wb,
                       theta, mix);
}
template <int WAVES>
static void inferPathsW(hipStream_t st, const DevScene& s, const PathState& ps, const uint32_t* queue,
                        const uint32_t* count, uint32_t capacity, float3 lo, float3 invExt, const uint16_t* wb,
                        const float* theta, float* mix, uint32_t cus) {
    if (s.numInst)
        inferPaths<WAVES>(st, PathSource<true>{s.tris, ps.p, queue, lo, invExt, s.inst}, count, capacity, wb, theta, mix,
                          cus);
    else
        inferPaths<WAVES>(st, PathSource<false>{s.tris, ps.p, queue, lo, invExt, s.inst}, count, capacity, wb, theta,
                          mix, cus);
}
void launchGuideInferPaths(hipStream_t st, const DevScene& s, const PathState& ps, const uint32_t* queue,
                           const uint32_t* count, uint32_t capacity, float3 lo, float3 invExt, const uint16_t* wb,
                           const float* theta, float* mix, uint32_t cus, uint32_t waves) {
    if (waves == 4) inferPathsW<4>(st, s, ps, queue, count, capacity, lo, invExt, wb, theta, mix, cus);
    else if (waves == 16) inferPathsW<16>(st, s, ps, queue, count, capacity, lo, invExt, wb, theta, mix, cus);
    else inferPathsW<8>(st, s, ps, queue, count, capacity, lo, invExt, wb, theta, mix, cus);
}

size_t groupBufferBytes(uint32_t capacity) {
    return (2 * static_cast<size_t>(kMaxBatchGuides) * kSortBlocks + 2 * kMaxBatchGuides + 1 + capacity +
            static_cast<size_t>(kMaxBatchGuides) * kGroupUnit) * 4;
}

GroupBuffers groupBuffers(void* base, uint32_t capacity) {
    uint32_t* p = static_cast<uint32_t*>(base);
    (void)capacity;
    const size_t table = static_cast<size_t>(kMaxBatchGuides) * kSortBlocks;
    return GroupBuffers{p, p + table, p + 2 * table, p + 2 * table + kMaxBatchGuides + 1,
                        p + 2 * table + 2 * kMaxBatchGuides + 1};
}

void launchGuideInferGrouped(hipStream_t st, const BatchTable& bt, const PathState& ps, const uint32_t* queue,
                             const uint32_t* count, uint32_t capacity, uint32_t nKeys, const GroupBuffers& gb,
                             const uint16_t* const* keyWb, const float* const* keyTheta, float* mix,
                             uint32_t maxBlocks) {
    const BatchPathSource<false> src{bt.scenes, bt.guide, ps.p, queue};
    hipLaunchKernelGGL(kGroupCount, dim3(kSortBlocks), dim3(256), 0, st, src, count, nKeys, gb);
    hipLaunchKernelGGL(kGroupScan, dim3(1), dim3(kSortBlocks), 0, st, nKeys, gb);
    hipLaunchKernelGGL(kGroupScatter, dim3(kSortBlocks), dim3(256), 0, st, src, count, nKeys, gb);
    const dim3 grid(gridFor(capacity, kInferBlock, maxBlocks));
    if (bt.motion) {
        const BatchPathSource<true> msrc{bt.scenes, bt.guide, ps.p, queue};
        hipLaunchKernelGGL(kGuideInferGrouped<true>, grid, dim3(kInferBlock), kInferLds, st, msrc, nKeys, gb, keyWb,
                           keyTheta, mix);
    } else {
        hipLaunchKernelGGL(kGuideInferGrouped<false>, grid, dim3(kInferBlock), kInferLds, st, src, nKeys, gb, keyWb,
                           keyTheta, mix);
    }
}

GuideBatchLayout guideBatchLayout(uint32_t n) {
    const uint32_t n64 = (n + 63) / 64 * 64;
    const uint32_t perChunk = (n64 + kGMaxChunks - 1) / kGMaxChunks;
    GuideBatchLayout b;
    b.chunk = (perChunk + 63) / 64 * 64;
    if (b.chunk == 0) b.chunk = 64;
    b.chunks = (n64 + b.chunk - 1) / b.chunk;
    if (b.chunks == 0) b.chunks = 1;
    b.npad = b.chunks * b.chunk;
    return b;
}

size_t guideTrainScratchBytes(const GuideBatchLayout& b) {
    const size_t rows = kGIn + 4 * kGHid + 4 * kGHid + kGOut;
    return rows * b.npad * 2;
}

void buildGuideTiles(void* scratch, const GuideBatchLayout& b, WGradTile* host) {
    const uint32_t npad = b.npad;
    uint16_t* x0 = static_cast<uint16_t*>(scratch);
    uint16_t* xh = x0 + static_cast<size_t>(kGIn) * npad;
    uint16_t* dz = xh + static_cast<size_t>(4 * kGHid) * npad;
    int t = 0;
    const uint32_t wOff[5] = {kGW0, kGW1, kGW2, kGW3, kGW4}, bOff[5] = {kGB0, kGB1, kGB2, kGB3, kGB4};
    for (int l = 0; l < 5; ++l) {
        const int rows = l == 4 ? kGOut : kGHid, cols = l == 0 ? kGIn : kGHid;
        const uint16_t* x = l == 0 ? x0 : xh + static_cast<size_t>(l - 1) * kGHid * npad;
        const uint16_t* d = dz + static_cast<size_t>(l) * kGHid * npad;
        for (int mo = 0; mo < rows / 16; ++mo)
            for (int ki = 0; ki < cols / 16; ++ki)
                host[t++] = WGradTile{static_cast<uint16_t>(rows), static_cast<uint16_t>(cols), static_cast<uint16_t>(mo),
                                      static_cast<uint16_t>(ki), wOff[l], bOff[l], d, x};
    }
}

void launchGuideGradients(hipStream_t st, const float* rec, uint32_t n, const GuideBatchLayout& b, const uint16_t* wb,
                          const float* theta, void* scratch, float* csum, float* partial, float* lossAcc,
                          const WGradTile* tilesDev, uint32_t maxBlocks, bool learnAlpha) {
    uint16_t* base = static_cast<uint16_t*>(scratch);
    TrainBuffers tb{base, base + static_cast<size_t>(kGIn) * b.npad, base + static_cast<size_t>(kGIn + 4 * kGHid) * b.npad,
                    b.npad, learnAlpha ? 1u : 0u};
    hipLaunchKernelGGL(kGuideSumC, dim3(1), dim3(256), 0, st, reinterpret_cast<const float4*>(rec), n, csum);
    hipLaunchKernelGGL(kGuideTrainFB, dim3(gridFor(b.npad, 64, maxBlocks)), dim3(256), kTrainLds, st,
                       reinterpret_cast<const float4*>(rec), n, wb, theta, csum, tb, lossAcc);
    hipLaunchKernelGGL(kGuideWGrad, dim3(kGTiles, b.chunks), dim3(64), 0, st, tilesDev, b.npad, b.chunk, partial);
}

void launchGuideAdam(hipStream_t st, const float* partial, uint32_t chunks, float* theta, float* m, float* v,
                     uint16_t* wb, float* gradOut, const AdamParams& ap) {
    hipLaunchKernelGGL(kGuideAdam, dim3((kGParams + 255) / 256), dim3(256), 0, st, partial, chunks, theta, m, v, wb,
                       gradOut, ap);
}

void guideConfigureLds() {
    static bool done = false;
    if (done) return;
    auto setLds = [](const void* k, size_t bytes) {
        (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
    };
    setLds(reinterpret_cast<const void*>(&kGuideInfer<ArraySource, kInferBlock / 64>), kInferLds);
    setLds(reinterpret_cast<const void*>(&kGuideInferGrouped<false>), kInferLds);
    setLds(reinterpret_cast<const void*>(&kGuideInferGrouped<true>), kInferLds);
    setLds(reinterpret_cast<const void*>(&kGuideInfer<PathSource<false>, 4>), inferLds(4));
    setLds(reinterpret_cast<const void*>(&kGuideInfer<PathSource<true>, 4>), inferLds(4));
    setLds(reinterpret_cast<const void*>(&kGuideInfer<PathSource<false>, 8>), inferLds(8));
    setLds(reinterpret_cast<const void*>(&kGuideInfer<PathSource<true>, 8>), inferLds(8));
    setLds(reinterpret_cast<const void*>(&kGuideInfer<PathSource<false>, 16>), inferLds(16));
    setLds(reinterpret_cast<const void*>(&kGuideInfer<PathSource<true>, 16>), inferLds(16));
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&kGuideTrainFB), hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(kTrainLds));
    done = true;
}

void launchGuideToBf16(hipStream_t st, const float* theta, uint16_t* wb) {
    hipLaunchKernelGGL(kGuideToBf16, dim3((kGParams + 255) / 256), dim3(256), 0, st, theta, wb);
}

}  // namespace pg

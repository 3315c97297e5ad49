// gfx950 kernels of the wavefront path tracer and their launchers.
//
// One render call processes the (pixel, sample) paths of a rect in waves of `capacity` paths.
// Per wave: raygen, then per bounce extend (closest hit) -> shade (emission, NEE, MIS-BSDF ray,
// continuation, Russian roulette) -> shadow (any hit) -> mis (closest hit, selected light test)
// -> film (splat of finished paths). The stages communicate through queues compacted with one
// atomic per wave (ballot + popcount prefix). The per-bounce structure restates trace() and
// multipleImportanceSampling() of Integrator/PathTracing/Kernel.cpp:23-193; the per-sample driver
// restates piperMain (Tracer/Embree/Kernel.cpp:47-100).
#include <algorithm>
#include <map>
#include <tuple>
#include <mutex>
#include <type_traits>

#include "pg_device.hpp"
#include "pg_guide.hpp"
#include "pg_kernels.hpp"

namespace pg {

// ------------------------------------------------------------------------------------------
// raygen: piperMain up to rayGen (Tracer/Embree/Kernel.cpp:47-90)
// ------------------------------------------------------------------------------------------
struct FilmPoint {
    float x, y;
};

// Path g of a segment: pixel-major over the rect, the segment's samples of one pixel adjacent.
__device__ __forceinline__ void pathPixel(const RenderConsts& rc, uint32_t sppBegin, uint32_t sppCount, uint32_t g,
                                          uint32_t& px, uint32_t& py, uint32_t& si) {
    si = sppBegin + g % sppCount;
    const uint32_t pix = g / sppCount;
    px = rc.rectX + pix % rc.rectW;
    py = rc.rectY + pix / rc.rectW;
}
__device__ __forceinline__ void pathPixel(const RenderConsts& rc, uint32_t g, uint32_t& px, uint32_t& py,
                                          uint32_t& si) {
    pathPixel(rc, rc.sppBegin, rc.sppCount, g, px, py, si);
}

// Paths of one segment of a single-scene render.
__device__ __forceinline__ uint64_t segmentPaths(const RenderConsts& rc) {
    return static_cast<uint64_t>(rc.rectW) * rc.rectH * rc.segSpp;
}

__device__ __forceinline__ Ray primaryRay(const DevScene& s, const RenderConsts& rc, uint32_t px, uint32_t py,
                                          uint32_t si, uint32_t sampleBits, uint64_t& idx, FilmPoint& fp,
                                          float& time) {
    idx = intervalToIndex(s, rc.log2Res, si, px, py, sampleBits);
    const float res = static_cast<float>(1u << rc.log2Res);
    float u[8];
    sobolWindow(s, idx, 0, u);
    // sobolStart (ScrambledSobolSampler/Kernel.cpp:52-62)
    float fx = fminf(fmaxf(u[0] * res - static_cast<float>(px), 0.0f), 1.0f);
    float fy = fminf(fmaxf(u[1] * res - static_cast<float>(py), 0.0f), 1.0f);
    fx += static_cast<float>(px);
    fy += static_cast<float>(py);
    fp = FilmPoint{fx, fy};
    const float ndcX = rc.ndc[0] + rc.ndc[2] * fx / static_cast<float>(rc.width);
    const float ndcY = rc.ndc[1] + rc.ndc[3] * fy / static_cast<float>(rc.height);
    // dim 2 is the path's time (Embree/Kernel.cpp:77), read by moving instances; dims 3, 4 drive the lens
    time = u[2];
    return cameraRay(s.cam, ndcX, ndcY, u[3], u[4]);
}

// Work item of a path (meta.z bits 16..31).
__device__ __forceinline__ uint32_t sceneOf(uint32_t metaZ) { return metaZ >> 16; }

// Path pool. The slots of finished paths form a stack (done[0, counts[kCntDone])); the shade
// stage pushes finished paths on top, the record stage writes their film records, and the next
// iteration restarts the slots with the next path numbers. The extension queue therefore stays
// full until the last paths of the render drain, and no wave has to finish before the next starts.
__global__ void __launch_bounds__(256) kPoolInit(Queues q, uint32_t capacity) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < capacity) q.done[i] = i;
    if (i == 0) {
        q.counts[kCntExtCur] = q.counts[kCntExtNext] = q.counts[kCntShadow] = q.counts[kCntMis] = 0;
        q.counts[kCntDone] = q.counts[kCntRecorded] = capacity;
        q.counts[kCntRegenBase] = q.counts[kCntRegenCount] = 0;
        q.pool->next = 0;
        q.pool->regenFirst = 0;
    }
}

// One thread at the start of an iteration: new path numbers for the free slots (up to the limit
// the host keeps in device memory), and the counter resets of the bounce (a kernel, not a memset,
// so the sequence can be captured in a graph).
__global__ void kIterBegin(Queues q, uint32_t curSlot) {
    const uint32_t top = q.counts[kCntDone];
    const unsigned long long next = q.pool->next;
    const unsigned long long limit = q.pool->limit;
    const unsigned long long avail = limit > next ? limit - next : 0ull;
    const uint32_t m = avail < top ? static_cast<uint32_t>(avail) : top;
    q.counts[kCntRegenBase] = top - m;
    q.counts[kCntRegenCount] = m;
    q.counts[kCntRegenExt] = q.counts[curSlot];  // behind the paths that continue
    q.counts[curSlot] += m;
    q.counts[kCntDone] = top - m;
    q.counts[kCntRecorded] = top - m;
    q.pool->regenFirst = next;
    q.pool->next = next + m;
    q.counts[curSlot ^ 1u] = 0;
    q.counts[kCntShadow] = q.counts[kCntMis] = 0;
}

// raygen for the new paths (piperMain up to rayGen, Tracer/Embree/Kernel.cpp:47-90); their slots
// join the current extension queue behind the paths that continue from the last bounce.
template <bool BATCH>
__global__ void __launch_bounds__(256) kRegen(DevScene s0, PathState ps, Queues q, RenderConsts rc0, uint32_t curSlot,
                                              BatchTable bt) {
    const uint32_t m = q.counts[kCntRegenCount];
    const uint32_t base = q.counts[kCntRegenBase];
    const uint64_t first = q.pool->regenFirst;
    uint32_t* ext = q.ext[curSlot] + q.counts[kCntRegenExt];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        {
            const uint32_t slot = q.done[base + i];
            ext[i] = slot;
            const uint64_t G = first + i;
            uint32_t item, g, sppBegin, sppCount;
            if (BATCH) {  // last item whose first path number is <= G
                uint32_t lo = 0, hi = bt.n;
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (bt.pathOffsets[mid] <= G) lo = mid;
                    else hi = mid;
                }
                item = lo;
                g = static_cast<uint32_t>(G - bt.pathOffsets[item]);
                sppBegin = bt.rc[item].sppBegin;
                sppCount = bt.rc[item].sppCount;
            } else {
                // one render's path numbers stay below 2^32 (checked on the host): 32-bit division
                const uint32_t segPaths = static_cast<uint32_t>(segmentPaths(rc0)), G32 = static_cast<uint32_t>(G);
                item = G32 / segPaths;
                g = G32 - item * segPaths;
                sppBegin = rc0.sppBegin + item * rc0.segSpp;
                sppCount = min(rc0.segSpp, rc0.sppBegin + rc0.sppCount - sppBegin);
            }
            const DevScene& s = BATCH ? bt.scenes[item] : s0;
            const RenderConsts& rc = BATCH ? bt.rc[item] : rc0;
            uint32_t px, py, si;
            pathPixel(rc, sppBegin, sppCount, g, px, py, si);
            uint64_t idx;
            FilmPoint fp;
            float time;
            // si < rc.spp, a power of two; a batch's renders may differ in it
            const Ray r = primaryRay(s, rc, px, py, si, BATCH ? 32u : __builtin_ctz(rc0.spp), idx, fp, time);
            ps.p[slot].o = make_float4(r.o.x, r.o.y, r.o.z, 0.0f);
            ps.p[slot].d = make_float4(r.d.x, r.d.y, r.d.z, time);  // d.w: the path's time, kept by every bounce
            ps.p[slot].thr = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
            ps.p[slot].L = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            ps.p[slot].Lmis = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            ps.p[slot].hit = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // the whole 128 B line is written
            ps.p[slot].idx = idx;
            // sample dims start at 5 (Embree/Kernel.cpp:68); a camera ray counts as specular
            ps.p[slot].meta = make_uint4(5u, 0u, 1u | (item << 16), g);
            ps.p[slot].film = make_float2(fp.x, fp.y);
        }
    }
}

// ------------------------------------------------------------------------------------------
// extend: closest hit for every queued path
// ------------------------------------------------------------------------------------------
template <int W, bool COUNT, bool BATCH, bool MOTION>
__global__ void __launch_bounds__(64) kExtend(DevScene s0, PathState ps, const uint32_t* __restrict__ queue,
                                              const uint32_t* __restrict__ count, uint2* spill, uint32_t spillPerLane,
                                              unsigned long long* travCounts, BatchTable bt, int ldsDepth) {
    extern __shared__ uint2 lstack[];  // ldsDepth x 64 entries
    const uint32_t n = *count;
    const uint32_t lane = threadIdx.x;
    const uint32_t gtid = blockIdx.x * 64 + lane;
    uint2* gs = spill + static_cast<size_t>(gtid) * spillPerLane;
    TravCount cnt;
    for (uint32_t i = gtid; i < n; i += gridDim.x * 64) {
        const uint32_t path = queue[i];
        const DevScene& s = BATCH ? bt.scenes[sceneOf(ps.p[path].meta.z)] : s0;
        const float4 o = ps.p[path].o, d = ps.p[path].d;
        const Ray r{xyz(o), xyz(d)};
        const HitRec h = closestHit<W, COUNT, MOTION>(s, r, 1e-3f, 1e5f, TravStack{lstack + lane, gs, ldsDepth}, cnt, d.w);
        ps.p[path].hit = make_float4(h.t, __uint_as_float(h.pos), h.u, h.v);
    }
    if (COUNT) {
        atomicAdd(&travCounts[0], static_cast<unsigned long long>(cnt.nodes));
        atomicAdd(&travCounts[1], static_cast<unsigned long long>(cnt.tris));
    }
}

// ------------------------------------------------------------------------------------------
// shade: one vertex of trace() (PathTracing/Kernel.cpp:115-190) including
// multipleImportanceSampling (:23-107); shadow and MIS-BSDF rays go to their queues.
// ------------------------------------------------------------------------------------------

// Training history record of one guided vertex (6 float4): (pos01, mixture pdf) (normal) (wo)
// (sampled direction) (throughput after the bounce) (L once the vertex's direct light is in).
constexpr uint32_t kHistFloat4 = 6;

__device__ __forceinline__ bool isTrainPath(const GuideShade& gs, uint32_t g) {
    return gs.train && (g % gs.trainStride) == gs.trainPhase;
}
// The same for path g of an item of a guided batch.
__device__ __forceinline__ bool isTrainPath(const ItemGuide& ig, uint32_t g) {
    return ig.tr.capacity > 0 && (g % ig.trainStride) == ig.trainPhase;
}

// GGX lobes narrower than this (alpha = roughness^2 < 0.04) are sampled better by the BSDF than by
// any 8-lobe mixture the network can express: such vertices draw from the BSDF alone (the guide
// fraction is 0 there, which keeps the one-sample MIS estimator unbiased) and record no history.
constexpr float kGuideMinAlpha = 0.04f;
// alpha < 0: the network's learned selection probability, the mixture record's float 48.
__device__ __forceinline__ float guideFraction(float alpha, const Mat& m, const float* mix) {
    if (m.type() == kDevRoughGlass) return 0.0f;  // the guide proposes reflected directions only
    if (m.type() == 1u && m.alpha < kGuideMinAlpha) return 0.0f;
    return alpha < 0.0f ? mix[48] : alpha;
}

// Queue records of one shade block, staged in LDS until the block knows where its records go.
// Device-scope atomics on one counter serialise across the whole chip, so the shade stage takes
// one per block and queue (not one per wave); the staged records then leave as coalesced stores.
template <uint32_t BLOCK>
struct ShadeLds {
    ShadowRec nee[BLOCK];   // NEE shadow ray
    ShadowRec rec2[BLOCK];  // MIS-BSDF occlusion query (ShadowRec) or MIS ray (MisRec)
    uint32_t cnt[5][BLOCK / 64];  // per wave: ext, done, nee, MIS query, MIS ray
    uint32_t base[4];                   // block's first entry: ext, done, shadow, mis queue
};
static_assert(sizeof(MisRec) == sizeof(ShadowRec), "MIS rays share the staging slots");

// Waves per SIMD: 4 (<= 128 VGPRs) for the plain and batched kernels: at 5 (<= 96 VGPRs) the
// plastic / Oren-Nayar lobes spill 14 VGPRs and the stage takes 18.9 instead of 17.0 ms on the
// 1080p64 terrain frame. The guided ones keep 3 for their extra state (the untextured guided kernel
// also fits 4 without spilling, and shades no faster).
template <bool BATCH, bool GUIDED, bool TEX, bool MOTION, uint32_t BLOCK>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(GUIDED ? 3 : 4)))
kShade(DevScene s0, PathState ps, Queues q, RenderConsts rc0, uint32_t curSlot, uint32_t nextSlot, uint64_t* totals,
       GuideShade gs, BatchTable bt) {
    __shared__ ShadeLds<BLOCK> lds;
    const uint32_t n = q.counts[curSlot];
    const uint32_t* qcur = q.ext[curSlot];
    uint32_t* qnext = q.ext[nextSlot];
    const uint32_t tid = threadIdx.x, wave = tid / 64, lane = __lane_id();
    const uint64_t lowerLanes = (1ull << lane) - 1ull;
    uint64_t totExt = 0, totNee = 0, totMis = 0;  // thread 0: this block's ray counts
    const uint32_t stride = gridDim.x * BLOCK;
    for (uint32_t i0 = blockIdx.x * BLOCK; i0 < n; i0 += stride) {  // block-uniform loop (barriers)
        const uint32_t i = i0 + tid;
        bool pushExt = false, pushShadow = false, pushMis = false, pushMisQuery = false, pushDone = false;
        uint32_t path = 0;
        if (i < n) {
            path = qcur[i];
            const float4 hit4 = ps.p[path].hit;
            const uint32_t pos = __float_as_uint(hit4.y);
            uint4 meta = ps.p[path].meta;
            const DevScene& s = BATCH ? bt.scenes[sceneOf(meta.z)] : s0;
            const RenderConsts& rc = BATCH ? bt.rc[sceneOf(meta.z)] : rc0;
            const bool d1 = rc.flags & kCompatD1, d3 = rc.flags & kCompatD3, d9 = rc.flags & kCompatD9,
                       d13 = rc.flags & kCompatD13;
            const bool envBlack = !valid(ld3(s.env));
            const float4 d4 = ps.p[path].d;
            const float time = MOTION ? d4.w : 0.0f;  // the path's time matters to moving instances only
            Ray ray{xyz(ps.p[path].o), xyz(d4)};
            V3 thr = xyz(ps.p[path].thr);
            // fold in last bounce's MIS-BSDF light hit (after its NEE term, as the reference adds them);
            // L and Lmis are written back only when they change (adding a zero leaves L's bits)
            const float4 lmis = ps.p[path].Lmis;
            bool lDirty = lmis.x != 0.0f || lmis.y != 0.0f || lmis.z != 0.0f;
            V3 L = xyz(ps.p[path].L) + xyz(lmis);
            if (lDirty) ps.p[path].Lmis = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (GUIDED && (meta.z & kMetaPending)) {  // the previous guided vertex's direct light is now in L
                const uint32_t r = ((meta.z >> 8) & 0xFFu) - 1u;
                gs.hist[(static_cast<size_t>(path) * gs.histDepth + r) * kHistFloat4 + 5] = make_float4(L.x, L.y, L.z, 0.0f);
                meta.z &= ~kMetaPending;
            }
            const bool specular = meta.z & 1u;
            const bool isLight = pos != kHitMiss && (pos & kHitQuad);
            if (specular && (isLight || pos == kHitMiss)) {  // :119-140, two-sided emission
                const V3 rad = isLight ? ld3(s.quads[pos & ~kHitQuad].radiance) : ld3(s.env);
                L = L + mul(thr, rad);
                lDirty = true;
            }
            if (isLight) {  // :142-146 pass through the light
                if (d13) meta.y -= 1u;
                ray.o = ray.o + ray.d * (hit4.x + 0.001f);
                ps.p[path].o = make_float4(ray.o.x, ray.o.y, ray.o.z, 0.0f);
                pushExt = true;
            } else if (pos == kHitMiss || meta.y >= rc.maxDepth) {  // :148
                pushDone = true;
            } else {
                bool front;
                const V3 nrm = triNormal<MOTION>(s, pos, ray.d, &front, time);
                const Mat mat = loadMat<TEX>(s, s.triMat[pos], front, pos, hit4.z, hit4.w);
                const Frame fr = shadingFrame(nrm, d3);
                const V3 wo = fr.toLocal(-ray.d);
                const V3 ng = fr.toLocal(nrm);
                const V3 hp = ray.o + ray.d * hit4.x;
                // The vertex draws at most 8 consecutive dimensions: light select, light (2),
                // MIS BSDF (2), continuation (2), Russian roulette (1), in that order. A specular
                // vertex samples no light and draws only the last three.
                const bool specV = isSpecular(mat);
                float u[8];
                sobolWindow(s, ps.p[path].idx, meta.x, u);
                uint32_t used = specV ? 2u : 7u;

                // A point light is a delta light: its sample is not MIS-weighted and the BSDF branch
                // draws nothing (select.delta, :48-54, :61-62), so the continuation follows the light
                // sample's dimensions directly.
                bool deltaSel = false;
                if (!specV) {  // multipleImportanceSampling (:23-107) runs only if noSpecular
                    // light selection: uniformSelect over {environment, area lights, point lights}
                    const uint32_t count = 1u + s.numLights + s.numPoints;
                    const uint32_t sel = min(static_cast<uint32_t>(floorf(u[0] * static_cast<float>(count))), count - 1u);
                    const float selInv = __builtin_amdgcn_rcpf(__builtin_amdgcn_rcpf(static_cast<float>(count)));
                    deltaSel = sel > s.numLights;
                    if (deltaSel) used = 5u;
                    const LightSample ls = sampleLight(s, sel, hp, d9 ? u[2] : u[1], d9 ? u[1] : u[2], d1);
                    if (ls.pdf > 0.0f && valid(ls.rad)) {
                        const V3 wi = fr.toLocal(ls.dir);
                        const V3 f = bsdfEval(mat, wo, wi, ng);
                        if (valid(f)) {
                            float w = fdiv(fabsf(wi.z), ls.pdf);
                            if (!deltaSel) w = w * powerHeuristic(ls.pdf, bsdfPdf(mat, wo, wi));
                            const V3 c = mul(thr, mul(f, ls.rad) * w * selInv);
                            // staged where it is made, so its registers are free for the rest of the vertex
                            lds.nee[tid] = ShadowRec{hp.x, hp.y, hp.z, fminf(1e-3f, ls.dist), ls.dir.x, ls.dir.y,
                                                     ls.dir.z, ls.dist, c.x, c.y, c.z, path};
                            pushShadow = true;
                        }
                    }
                    BsdfSample bs{{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, 0.0f};
                    if (!deltaSel) bs = bsdfSample(mat, wo, ng, d9 ? u[4] : u[3], d9 ? u[3] : u[4]);
                    if (bs.pdf > 0.0f && valid(bs.f)) {
                        const float weight = fdiv(fabsf(bs.wi.z), bs.pdf);
                        const V3 dir = fr.toWorld(bs.wi);
                        const V3 c = mul(thr, bs.f * weight * selInv);
                        if (!envBlack) {
                            *reinterpret_cast<MisRec*>(&lds.rec2[tid]) =
                                MisRec{hp.x, hp.y, hp.z, bs.pdf, dir.x, dir.y, dir.z, sel, c.x, c.y, c.z, path};
                            pushMis = true;
                        } else if (sel != 0u) {
                            // With a black environment the BSDF ray contributes only if a primitive
                            // of the selected light is its closest hit: that primitive must be the
                            // light's first along the ray, beat every other light's quads in (t, index)
                            // order (closestHit's quad loop), and no triangle may lie in (tmin, tq]
                            // (triangles win ties). The first two are decided here; the triangle test
                            // becomes an occlusion query on the shadow queue.
                            const Ray mray{hp, dir};
                            float tq;
                            uint32_t qi = 0;
                            if (firstLightPrimitive(s, sel - 1u, mray, tq, qi)) {
                                const DevQuad& qd = s.quads[qi];
                                if (noOtherLightFirst(s, sel - 1u, mray, tq, qi)) {
                                    const V3 qn = quadNormal(qd, dir);
                                    const V3 rad = dot(qn, dir) < 0.0f ? ld3(qd.radiance) : V3{0.0f, 0.0f, 0.0f};
                                    const float lpdf = fdiv(tq * tq, fabsf(dot(qn, dir)) * qd.area);
                                    if (valid(rad) && lpdf > 0.0f) {
                                        const V3 cm = mul(c, rad) * powerHeuristic(bs.pdf, lpdf);
                                        // tq > 0: the next float up makes the range (tmin, tq] inclusive
                                        const float tqUp = __uint_as_float(__float_as_uint(tq) + 1u);
                                        lds.rec2[tid] = ShadowRec{hp.x, hp.y, hp.z, 1e-3f, dir.x, dir.y, dir.z, tqUp,
                                                                  cm.x, cm.y, cm.z, path | kMisFlag};
                                        pushMisQuery = true;
                                    }
                                }
                            }
                        }
                    }
                }

                // continuation (:164-190). With a guide: one-sample MIS that draws from the guide
                // mixture with probability alpha, else from the BSDF, and weights by the combined
                // density alpha p_guide + (1 - alpha) p_bsdf, so the estimator stays unbiased.
                const float c0 = specV ? u[0] : (deltaSel ? u[3] : u[5]), c1 = specV ? u[1] : (deltaSel ? u[4] : u[6]);
                const float uc1 = d9 ? c1 : c0, uc2 = d9 ? c0 : c1;
                BsdfSample ss;
                float pdfMix = 0.0f, a = 0.0f;
                float pb = 0.0f;  // the BSDF's density of the drawn direction (training with a learned alpha)
                if (!GUIDED || specV) {
                    ss = bsdfSample(mat, wo, ng, uc1, uc2);
                } else {
                    const float* mix = gs.mix + static_cast<size_t>(i) * kGMixFloats;
                    a = guideFraction(BATCH ? bt.guide[sceneOf(meta.z)].alpha : gs.alpha, mat, mix);
                    bool ok = true;
                    const bool fromGuide = uc1 < a;
                    if (fromGuide) {
                        const float3 w = mixtureSample(mix, fdiv(uc1, a), uc2);
                        ss.wi = fr.toLocal(V3{w.x, w.y, w.z});
                    } else {  // the BSDF's own sample carries its value and density
                        const BsdfSample b = bsdfSample(mat, wo, ng, fdiv(uc1 - a, 1.0f - a), uc2);
                        ss.wi = b.wi;
                        ss.f = b.f;
                        pb = b.pdf;
                        ok = b.pdf > 0.0f;
                    }
                    if (ok) {
                        if (fromGuide) pb = bsdfPdf(mat, wo, ss.wi), ss.f = bsdfEval(mat, wo, ss.wi, ng);
                        pdfMix = (1.0f - a) * pb;
                        if (a > 0.0f) {
                            const V3 ww = fr.toWorld(ss.wi);
                            pdfMix += a * mixturePdf(mix, make_float3(ww.x, ww.y, ww.z));
                        }
                        ss.pdf = pdfMix;
                    } else {
                        ss.f = V3{0.0f, 0.0f, 0.0f};
                        ss.pdf = 0.0f;
                    }
                }
                if (!(ss.pdf > 0.0f && valid(ss.f))) {
                    pushDone = true;
                } else {
                    thr = mul(thr, ss.f * fdiv(fabsf(ss.wi.z), ss.pdf));
                    bool alive = true;
                    if (meta.y > rc.minRR) {
                        const float qrr = fmaxf(0.05f, 1.0f - maxc(thr));
                        used += 1u;
                        if ((specV ? u[2] : (deltaSel ? u[5] : u[7])) < qrr) alive = false;
                        else thr = thr * __builtin_amdgcn_rcpf(1.0f - qrr);
                    }
                    if (alive) {
                        const V3 nd = fr.toWorld(ss.wi);
                        // :184-187: offset along +Ng for reflection, -Ng for transmission
                        const V3 no = hp + nrm * (ss.wi.z < 0.0f ? -0.001f : 0.001f);
                        ps.p[path].o = make_float4(no.x, no.y, no.z, 0.0f);
                        ps.p[path].d = make_float4(nd.x, nd.y, nd.z, time);
                        ps.p[path].thr = make_float4(thr.x, thr.y, thr.z, 0.0f);
                        meta.y += 1u;
                        meta.z = specV ? (meta.z | 1u) : (meta.z & ~1u);  // :189 specular = part & Specular
                        pushExt = true;
                        const uint32_t r = (meta.z >> 8) & 0xFFu;
                        const bool train = BATCH ? isTrainPath(bt.guide[sceneOf(meta.z)], meta.w) : isTrainPath(gs, meta.w);
                        if (GUIDED && a > 0.0f && train && r < gs.histDepth) {
                            float4* h = gs.hist + (static_cast<size_t>(path) * gs.histDepth + r) * kHistFloat4;
                            const float3 lo = BATCH ? bt.guide[sceneOf(meta.z)].lo : gs.lo;
                            const float3 ie = BATCH ? bt.guide[sceneOf(meta.z)].invExt : gs.invExt;
                            const float px = fminf(fmaxf((hp.x - lo.x) * ie.x, 0.0f), 1.0f);
                            const float py = fminf(fmaxf((hp.y - lo.y) * ie.y, 0.0f), 1.0f);
                            const float pz = fminf(fmaxf((hp.z - lo.z) * ie.z, 0.0f), 1.0f);
                            h[0] = make_float4(px, py, pz, pdfMix);
                            h[1] = make_float4(nrm.x, nrm.y, nrm.z, pb);
                            h[2] = make_float4(-ray.d.x, -ray.d.y, -ray.d.z, 0.0f);
                            h[3] = make_float4(nd.x, nd.y, nd.z, 0.0f);
                            h[4] = make_float4(thr.x, thr.y, thr.z, 0.0f);
                            meta.z = (meta.z & ~0xFF00u) | ((r + 1u) << 8) | kMetaPending;
                        }
                    } else {
                        pushDone = true;
                    }
                }
                meta.x += used;
            }
            if (lDirty) ps.p[path].L = make_float4(L.x, L.y, L.z, 0.0f);
            ps.p[path].meta = meta;
        }
        // block-wide compaction: per-wave counts, one atomic per queue, then every lane's slot
        const uint64_t mExt = __ballot(pushExt), mDone = __ballot(pushDone), mNee = __ballot(pushShadow),
                       mMq = __ballot(pushMisQuery), mMis = __ballot(pushMis);
        if (lane == 0) {
            lds.cnt[0][wave] = __popcll(mExt);
            lds.cnt[1][wave] = __popcll(mDone);
            lds.cnt[2][wave] = __popcll(mNee);
            lds.cnt[3][wave] = __popcll(mMq);
            lds.cnt[4][wave] = __popcll(mMis);
        }
        __syncthreads();
        if (tid < 4) {  // queue tid: ext, done, shadow (NEE + MIS queries), mis
            uint32_t t = 0;
            for (uint32_t w = 0; w < BLOCK / 64; ++w)
                t += tid == 2 ? lds.cnt[2][w] + lds.cnt[3][w] : lds.cnt[tid == 3 ? 4 : tid][w];
            uint32_t* ctr = tid == 0 ? &q.counts[nextSlot]
                            : tid == 1 ? &q.counts[kCntDone]
                            : tid == 2 ? &q.counts[kCntShadow]
                                       : &q.counts[kCntMis];
            lds.base[tid] = t ? atomicAdd(ctr, t) : 0u;
        }
        if (tid == 0)
            for (uint32_t w = 0; w < BLOCK / 64; ++w) {
                totExt += lds.cnt[0][w];
                totNee += lds.cnt[2][w];
                totMis += lds.cnt[3][w] + lds.cnt[4][w];
            }
        __syncthreads();
        uint32_t oExt = lds.base[0], oDone = lds.base[1], oSh = lds.base[2], oMis = lds.base[3];
        for (uint32_t w = 0; w < wave; ++w) {
            oExt += lds.cnt[0][w];
            oDone += lds.cnt[1][w];
            oSh += lds.cnt[2][w] + lds.cnt[3][w];
            oMis += lds.cnt[4][w];
        }
        if (pushExt) qnext[oExt + __popcll(mExt & lowerLanes)] = path;
        if (pushDone) q.done[oDone + __popcll(mDone & lowerLanes)] = path;
        if (pushShadow) q.shadow[oSh + __popcll(mNee & lowerLanes)] = lds.nee[tid];
        if (pushMisQuery) q.shadow[oSh + __popcll(mNee) + __popcll(mMq & lowerLanes)] = lds.rec2[tid];
        if (pushMis) q.mis[oMis + __popcll(mMis & lowerLanes)] = *reinterpret_cast<const MisRec*>(&lds.rec2[tid]);
        __syncthreads();  // the staging slots and counts are reused by the next iteration
    }
    if (totals && tid == 0) {
        if (totExt) atomicAdd(reinterpret_cast<unsigned long long*>(&totals[kTotExt]), totExt);
        if (totNee) atomicAdd(reinterpret_cast<unsigned long long*>(&totals[kTotShadow]), totNee);
        if (totMis) atomicAdd(reinterpret_cast<unsigned long long*>(&totals[kTotMis]), totMis);
    }
}

// ------------------------------------------------------------------------------------------
// shadow: occlusion test of NEE rays (sampleLightSource, PathTracing/Kernel.cpp:40-62)
// ------------------------------------------------------------------------------------------
template <int W, bool BATCH, bool MOTION>
__global__ void __launch_bounds__(64) kShadow(DevScene s0, PathState ps, Queues q, uint2* spill, uint32_t spillPerLane,
                                              BatchTable bt, int ldsDepth) {
    extern __shared__ uint2 lstack[];  // ldsDepth x 64 entries
    const uint32_t n = q.counts[kCntShadow];
    const uint32_t lane = threadIdx.x;
    const uint32_t gtid = blockIdx.x * 64 + lane;
    uint2* gs = spill + static_cast<size_t>(gtid) * spillPerLane;
    for (uint32_t i = gtid; i < n; i += gridDim.x * 64) {
        const ShadowRec r = q.shadow[i];
        const DevScene& s = BATCH ? bt.scenes[sceneOf(ps.p[r.path & ~kMisFlag].meta.z)] : s0;
        const Ray ray{{r.ox, r.oy, r.oz}, {r.dx, r.dy, r.dz}};
        const float time = s.numInst ? ps.p[r.path & ~kMisFlag].d.w : 0.0f;
        if (!occluded<W, MOTION>(s, ray, r.tmin, r.tmax, TravStack{lstack + lane, gs, ldsDepth}, time)) {
            if (r.path & kMisFlag) {  // MIS-BSDF light hit: kept apart, folded in after the NEE term
                ps.p[r.path & ~kMisFlag].Lmis = make_float4(r.cr, r.cg, r.cb, 0.0f);
            } else {
                float4 L = ps.p[r.path].L;
                L.x += r.cr, L.y += r.cg, L.z += r.cb;
                ps.p[r.path].L = L;
            }
        }
    }
}

// ------------------------------------------------------------------------------------------
// mis: the separate BSDF-sampled ray of sampleBSDF (PathTracing/Kernel.cpp:63-102); counts only
// if it hits the selected light (or escapes to the environment)
// ------------------------------------------------------------------------------------------
template <int W, bool BATCH, bool MOTION>
__global__ void __launch_bounds__(64) kMis(DevScene s0, PathState ps, Queues q, uint2* spill, uint32_t spillPerLane,
                                           BatchTable bt, int ldsDepth) {
    extern __shared__ uint2 lstack[];  // ldsDepth x 64 entries
    const uint32_t n = q.counts[kCntMis];
    const uint32_t lane = threadIdx.x;
    const uint32_t gtid = blockIdx.x * 64 + lane;
    uint2* gs = spill + static_cast<size_t>(gtid) * spillPerLane;
    for (uint32_t i = gtid; i < n; i += gridDim.x * 64) {
        const MisRec r = q.mis[i];
        const DevScene& s = BATCH ? bt.scenes[sceneOf(ps.p[r.path].meta.z)] : s0;
        const bool envBlack = !valid(ld3(s.env));
        const Ray ray{{r.ox, r.oy, r.oz}, {r.dx, r.dy, r.dz}};
        const float time = s.numInst ? ps.p[r.path].d.w : 0.0f;
        V3 rad{0.0f, 0.0f, 0.0f};
        float lpdf = 0.0f;
        if (envBlack && r.light != 0u) {
            // With a black environment the ray contributes only if a primitive of the selected
            // light is its closest hit. Same decision as closestHit(): the light's first primitive
            // along the ray must beat every other light's quads in (t, index) order, and no triangle
            // may lie in (tmin, t_q] (triangles win ties).
            float tq;
            uint32_t qi = 0;
            if (firstLightPrimitive(s, r.light - 1u, ray, tq, qi)) {
                const DevQuad& qd = s.quads[qi];
                const bool first = noOtherLightFirst(s, r.light - 1u, ray, tq, qi);
                // tq > 0: the next float up makes the occlusion range (tmin, tq] inclusive
                const float tqUp = __uint_as_float(__float_as_uint(tq) + 1u);
                if (first && !occluded<W, MOTION>(s, ray, 1e-3f, tqUp, TravStack{lstack + lane, gs, ldsDepth}, time)) {
                    const V3 qn = quadNormal(qd, ray.d);
                    rad = dot(qn, ray.d) < 0.0f ? ld3(qd.radiance) : V3{0.0f, 0.0f, 0.0f};
                    lpdf = fdiv(tq * tq, fabsf(dot(qn, ray.d)) * qd.area);
                }
            }
        } else {
            TravCount tc;
            const HitRec h = closestHit<W, false, MOTION>(s, ray, 1e-3f, 1e5f, TravStack{lstack + lane, gs, ldsDepth}, tc, time);
            if (h.pos == kHitMiss) {
                rad = ld3(s.env);
                lpdf = envPdf(ray.d);
            } else if ((h.pos & kHitQuad) && r.light == 1u + s.quads[h.pos & ~kHitQuad].light) {
                const DevQuad& qd = s.quads[h.pos & ~kHitQuad];
                const V3 qn = quadNormal(qd, ray.d);
                rad = dot(qn, ray.d) < 0.0f ? ld3(qd.radiance) : V3{0.0f, 0.0f, 0.0f};
                lpdf = fdiv(h.t * h.t, fabsf(dot(qn, ray.d)) * qd.area);
            }
        }
        if (valid(rad) && lpdf > 0.0f) {
            const float w = powerHeuristic(r.bsdfPdf, lpdf);
            const V3 c = mul(V3{r.cr, r.cg, r.cb}, rad) * w;
            float4 L = ps.p[r.path].L;
            L.x += c.x, L.y += c.y, L.z += c.z;
            ps.p[r.path].L = L;
        }
    }
}

// ------------------------------------------------------------------------------------------
// film: accumulate with the box filter, a 2x2 splat into RGBW (FixedSampler/Kernel.cpp:30-58)
// ------------------------------------------------------------------------------------------
// Film: accumulate() with the box filter (RenderDriver/FixedSampler/Kernel.cpp:30-58) adds (L, 1)
// to the 2x2 pixels starting at (int)(p - 0.5) for every sample. Instead of atomics, each finished
// path writes a record (L, film point) at its path number (in a ring of record slots); once all
// paths of a film segment are recorded, kQuadSums reduces each source pixel's samples into the
// four 2x2 footprints they can start at, and kResolve has every film pixel gather the footprints
// that cover it from its 3x3 source neighbourhood. Every sum runs in a fixed order, so the film is
// bit-reproducible whatever order the paths finish in.
template <bool BATCH>
__global__ void __launch_bounds__(256) kRecord(PathState ps, Queues q, RenderConsts rc0, RecordRing ring,
                                               GuideShade gs, GuideRecordOut tr, BatchTable bt) {
    const uint32_t lo = q.counts[kCntRecorded];
    const uint32_t n = q.counts[kCntDone] - lo;
    const uint32_t lane = __lane_id(), wave = threadIdx.x / 64;
    uint32_t wItem = 0, wCount = 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < n; i0 += stride) {  // block-uniform loop (ballots, barrier)
        const uint32_t i = i0 + threadIdx.x;
        const bool act = i < n;
        uint32_t item = 0;
        if (act) {
            const uint32_t path = q.done[lo + i];
            const V3 L = xyz(ps.p[path].L) + xyz(ps.p[path].Lmis);
            const float2 fp = ps.p[path].film;
            const uint4 meta = ps.p[path].meta;
            const uint32_t g = meta.w;
            item = sceneOf(meta.z);
            const uint64_t G = (BATCH ? bt.pathOffsets[item] : static_cast<uint64_t>(item) * segmentPaths(rc0)) + g;
            // a single render's ring and path numbers stay below 2^32 (checked on the host)
            const uint64_t slot = BATCH ? G % ring.size : static_cast<uint32_t>(G) % static_cast<uint32_t>(ring.size);
            ring.L[slot] = make_float4(L.x, L.y, L.z, fp.x);
            ring.y[slot] = fp.y;
            const bool train = BATCH ? (bt.guide != nullptr && isTrainPath(bt.guide[item], g)) : isTrainPath(gs, g);
            if (train) {
                const GuideRecordOut out = BATCH ? bt.guide[item].tr : tr;
                // Each recorded vertex learns from the radiance that reached it along its sampled
                // direction: (L_final - L after its direct light) / throughput after the bounce.
                // The loss weights every record by that radiance over its pdf, so records that
                // found no light carry no gradient and are not written.
                const uint32_t cnt = (meta.z >> 8) & 0xFFu;
                for (uint32_t r = 0; r < cnt; ++r) {
                    const float4* h = gs.hist + (static_cast<size_t>(path) * gs.histDepth + r) * kHistFloat4;
                    const float4 h0 = h[0], t = h[4], snap = h[5];
                    const float lr = t.x > 0.0f ? (L.x - snap.x) / t.x : 0.0f;
                    const float lg = t.y > 0.0f ? (L.y - snap.y) / t.y : 0.0f;
                    const float lb = t.z > 0.0f ? (L.z - snap.z) / t.z : 0.0f;
                    const float c = fmaxf((lr + lg + lb) * (1.0f / 3.0f), 0.0f) / h0.w;
                    if (!(c > 0.0f)) continue;
                    const uint32_t k = atomicAdd(out.count, 1u);
                    if (k >= out.capacity) break;
                    out.rec[4 * k] = make_float4(h0.x, h0.y, h0.z, c);
                    out.rec[4 * k + 1] = h[1];
                    out.rec[4 * k + 2] = h[2];
                    out.rec[4 * k + 3] = h[3];
                }
            }
        }
        // paths recorded per item, counted per wave (wave-uniform); a change of item flushes
        uint64_t pending = __ballot(act);
        while (pending) {
            const int leader = __builtin_ctzll(pending);
            const uint32_t it = __shfl(item, leader);
            const uint64_t same = __ballot(act && item == it);
            const uint32_t c = static_cast<uint32_t>(__popcll(same));
            if (wCount != 0 && it != wItem) {
                if (lane == 0) atomicAdd(&ring.itemDone[wItem], wCount);
                wCount = 0;
            }
            wItem = it;
            wCount += c;
            pending &= ~same;
        }
    }
    // one device atomic per block and item (atomics on one counter serialise across the chip)
    __shared__ uint32_t sItem[4], sCount[4];
    if (lane == 0) sItem[wave] = wItem, sCount[wave] = wCount;
    __syncthreads();
    if (threadIdx.x == 0)
        for (uint32_t w = 0; w < blockDim.x / 64; ++w) {
            if (sCount[w] == 0) continue;
            uint32_t c = sCount[w];
            for (uint32_t w2 = w + 1; w2 < blockDim.x / 64; ++w2)
                if (sCount[w2] != 0 && sItem[w2] == sItem[w]) c += sCount[w2], sCount[w2] = 0;
            atomicAdd(&ring.itemDone[sItem[w]], c);
        }
}

// Per source pixel of the rect: sums over its samples for each footprint origin
// (bx, by) in {px-1, px} x {py-1, py}; quad[4 * pixel + 2 * (by - py + 1) + (bx - px + 1)] = (rgb, count).
// A group of G lanes (G = sppCount rounded up to a power of two, at most 64) reduces one pixel:
// lane l takes samples l, l + G, ... (the group reads consecutive records, so each load is
// coalesced), then a butterfly over the group's lanes adds the partial sums in a fixed order.
template <typename F>
__global__ void __launch_bounds__(256) kQuadSums(RenderConsts rc, const float4* __restrict__ recL,
                                                 const float* __restrict__ recY, uint64_t start, uint64_t ringSize,
                                                 uint32_t groupLanes, F* quad) {
    const uint32_t G = groupLanes;
    const uint32_t lane = __lane_id(), l = lane % G;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const uint32_t pix = wave * (64 / G) + lane / G;
    const bool act = pix < rc.rectW * rc.rectH;  // no early exit: every lane joins the butterfly
    F acc[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) acc[k] = F(0);
    if (act) {
        const int32_t px = static_cast<int32_t>(rc.rectX + pix % rc.rectW);
        const int32_t py = static_cast<int32_t>(rc.rectY + pix / rc.rectW);
        const uint64_t base = start + static_cast<uint64_t>(pix) * rc.sppCount;  // start < ringSize
        for (uint32_t si = l; si < rc.sppCount; si += G) {
            uint64_t j = base + si;
            if (j >= ringSize) j -= ringSize;  // a segment is never longer than the ring
            const float4 r = recL[j];
            const int32_t bx = static_cast<int32_t>(r.w - 0.5f), by = static_cast<int32_t>(recY[j] - 0.5f);
            const int q = 2 * (by - py + 1) + (bx - px + 1);  // 0..3 (bx in {px-1, px}, by in {py-1, py})
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (q == k) {
                    acc[4 * k] += static_cast<F>(r.x);
                    acc[4 * k + 1] += static_cast<F>(r.y);
                    acc[4 * k + 2] += static_cast<F>(r.z);
                    acc[4 * k + 3] += F(1);
                }
        }
    }
    for (uint32_t off = G / 2; off > 0; off >>= 1)
#pragma unroll
        for (int k = 0; k < 16; ++k) acc[k] += __shfl_xor(acc[k], static_cast<int>(off));
    if (act && l == 0) {
        F* o = quad + 16 * static_cast<size_t>(pix);
#pragma unroll
        for (int k = 0; k < 16; ++k) o[k] = acc[k];
    }
}

// Filter weight of a sample at offset (dx, dy) from a pixel centre (FixedSampler/Kernel.cpp:55-70).
__device__ __forceinline__ float filterWeight(const RenderConsts& rc, float dx, float dy) {
    if (rc.filter == kFilterTriangle) return fmaxf(0.0f, 1.0f - fabsf(dx)) * fmaxf(0.0f, 1.0f - fabsf(dy));
    const float wx = fmaxf(expf(rc.negAlpha * dx * dx) - rc.sub, 0.0f);
    const float wy = fmaxf(expf(rc.negAlpha * dy * dy) - rc.sub, 0.0f);
    return wx * wy;
}

// Weighted filters (triangle, gaussian): per source pixel, the weighted sums (rgb w, w) its samples
// give each pixel of its 3x3 neighbourhood: bins[36 * pixel + 4 * (3 (ty - py + 1) + tx - px + 1)].
// Lane groups and the butterfly as in kQuadSums.
template <typename F>
__global__ void __launch_bounds__(256) kBinSums(RenderConsts rc, const float4* __restrict__ recL,
                                                const float* __restrict__ recY, uint64_t start, uint64_t ringSize,
                                                uint32_t groupLanes, F* bins) {
    const uint32_t G = groupLanes;
    const uint32_t lane = __lane_id(), l = lane % G;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const uint32_t pix = wave * (64 / G) + lane / G;
    const bool act = pix < rc.rectW * rc.rectH;  // no early exit: every lane joins the butterfly
    F acc[36];
#pragma unroll
    for (int k = 0; k < 36; ++k) acc[k] = F(0);
    if (act) {
        const int32_t px = static_cast<int32_t>(rc.rectX + pix % rc.rectW);
        const int32_t py = static_cast<int32_t>(rc.rectY + pix / rc.rectW);
        const uint64_t base = start + static_cast<uint64_t>(pix) * rc.sppCount;
        for (uint32_t si = l; si < rc.sppCount; si += G) {
            uint64_t j = base + si;
            if (j >= ringSize) j -= ringSize;
            const float4 r = recL[j];
            const float fx = r.w, fy = recY[j];
            const int32_t bx = static_cast<int32_t>(fx - 0.5f), by = static_cast<int32_t>(fy - 0.5f);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int32_t tx = bx + (k & 1), ty = by + (k >> 1);
                const float w = filterWeight(rc, fx - (static_cast<float>(tx) + 0.5f), fy - (static_cast<float>(ty) + 0.5f));
                const int bin = 3 * (ty - py + 1) + (tx - px + 1);  // 0..8
#pragma unroll
                for (int b = 0; b < 9; ++b)
                    if (bin == b) {
                        acc[4 * b] += static_cast<F>(r.x * w);
                        acc[4 * b + 1] += static_cast<F>(r.y * w);
                        acc[4 * b + 2] += static_cast<F>(r.z * w);
                        acc[4 * b + 3] += static_cast<F>(w);
                    }
            }
        }
    }
    for (uint32_t off = G / 2; off > 0; off >>= 1)
#pragma unroll
        for (int k = 0; k < 36; ++k) acc[k] += __shfl_xor(acc[k], static_cast<int>(off));
    if (act && l == 0) {
        F* o = bins + 36 * static_cast<size_t>(pix);
#pragma unroll
        for (int k = 0; k < 36; ++k) o[k] = acc[k];
    }
}

// Per film pixel around the rect: the bins its 3x3 source neighbourhood holds for it.
template <typename F>
__global__ void __launch_bounds__(256) kResolveBins(RenderConsts rc, const F* __restrict__ bins, F* film) {
    const int32_t x0 = max(static_cast<int32_t>(rc.rectX) - 1, 0), y0 = max(static_cast<int32_t>(rc.rectY) - 1, 0);
    const int32_t x1 = min(static_cast<int32_t>(rc.rectX + rc.rectW) + 1, static_cast<int32_t>(rc.width));
    const int32_t y1 = min(static_cast<int32_t>(rc.rectY + rc.rectH) + 1, static_cast<int32_t>(rc.height));
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t w = static_cast<uint32_t>(x1 - x0);
    if (i >= w * static_cast<uint32_t>(y1 - y0)) return;
    const int32_t tx = x0 + static_cast<int32_t>(i % w), ty = y0 + static_cast<int32_t>(i / w);
    F r = F(0), g = F(0), b = F(0), n = F(0);
    for (int32_t sy = ty - 1; sy <= ty + 1; ++sy) {
        if (sy < static_cast<int32_t>(rc.rectY) || sy >= static_cast<int32_t>(rc.rectY + rc.rectH)) continue;
        for (int32_t sx = tx - 1; sx <= tx + 1; ++sx) {
            if (sx < static_cast<int32_t>(rc.rectX) || sx >= static_cast<int32_t>(rc.rectX + rc.rectW)) continue;
            const F* q = bins + 36 * (static_cast<size_t>(sy - rc.rectY) * rc.rectW + (sx - rc.rectX)) +
                         4 * (3 * (ty - sy + 1) + (tx - sx + 1));
            r += q[0], g += q[1], b += q[2], n += q[3];
        }
    }
    F* p = film + 4 * (static_cast<size_t>(ty) * rc.width + tx);
    p[0] += r, p[1] += g, p[2] += b, p[3] += n;
}

// Per film pixel around the rect (rect grown by one pixel, clipped to the film).
template <typename F>
__global__ void __launch_bounds__(256) kResolve(RenderConsts rc, const F* __restrict__ quad, F* film) {
    const int32_t x0 = max(static_cast<int32_t>(rc.rectX) - 1, 0), y0 = max(static_cast<int32_t>(rc.rectY) - 1, 0);
    const int32_t x1 = min(static_cast<int32_t>(rc.rectX + rc.rectW) + 1, static_cast<int32_t>(rc.width));
    const int32_t y1 = min(static_cast<int32_t>(rc.rectY + rc.rectH) + 1, static_cast<int32_t>(rc.height));
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t w = static_cast<uint32_t>(x1 - x0);
    if (i >= w * static_cast<uint32_t>(y1 - y0)) return;
    const int32_t tx = x0 + static_cast<int32_t>(i % w), ty = y0 + static_cast<int32_t>(i / w);
    F r = F(0), g = F(0), b = F(0), n = F(0);
    for (int32_t sy = ty - 1; sy <= ty + 1; ++sy) {
        if (sy < static_cast<int32_t>(rc.rectY) || sy >= static_cast<int32_t>(rc.rectY + rc.rectH)) continue;
        for (int32_t sx = tx - 1; sx <= tx + 1; ++sx) {
            if (sx < static_cast<int32_t>(rc.rectX) || sx >= static_cast<int32_t>(rc.rectX + rc.rectW)) continue;
            const F* q = quad + 16 * (static_cast<size_t>(sy - rc.rectY) * rc.rectW + (sx - rc.rectX));
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int32_t bx = sx - 1 + (k & 1), by = sy - 1 + (k >> 1);
                if (tx >= bx && tx <= bx + 1 && ty >= by && ty <= by + 1) {
                    r += q[4 * k], g += q[4 * k + 1], b += q[4 * k + 2], n += q[4 * k + 3];
                }
            }
        }
    }
    F* p = film + 4 * (static_cast<size_t>(ty) * rc.width + tx);
    p[0] += r, p[1] += g, p[2] += b, p[3] += n;
}

// Plain copy kernel of known size, used to scale HBM byte counters.
__global__ void __launch_bounds__(256) kCopy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

// ------------------------------------------------------------------------------------------
// parity kernels
// ------------------------------------------------------------------------------------------
template <int W, bool MOTION>
__global__ void __launch_bounds__(64) kTrace(DevScene s, const float* rays, uint32_t n, float* t, uint32_t* prim,
                                             float* uv, uint8_t* occ, uint2* spill, uint32_t spillPerLane,
                                             const float* times, int ldsDepth) {
    extern __shared__ uint2 lstack[];  // ldsDepth x 64 entries
    const uint32_t lane = threadIdx.x;
    const uint32_t gtid = blockIdx.x * 64 + lane;
    uint2* gs = spill + static_cast<size_t>(gtid) * spillPerLane;
    for (uint32_t i = gtid; i < n; i += gridDim.x * 64) {
        const float* r = rays + 8 * static_cast<size_t>(i);
        const Ray ray{{r[0], r[1], r[2]}, {r[3], r[4], r[5]}};
        const float time = times ? times[i] : 0.0f;
        if (occ) {
            occ[i] = occluded<W, MOTION>(s, ray, r[6], r[7], TravStack{lstack + lane, gs, ldsDepth}, time) ? 1 : 0;
        } else {
            TravCount tc;
            const HitRec h = closestHit<W, false, MOTION>(s, ray, r[6], r[7], TravStack{lstack + lane, gs, ldsDepth}, tc, time);
            t[i] = h.t;
            prim[i] = h.prim;
            uv[2 * i] = h.u;
            uv[2 * i + 1] = h.v;
        }
    }
}

__global__ void kCameraRays(DevScene s, RenderConsts rc, const uint32_t* pix, uint32_t n, float* rays, float* film) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t idx;
    FilmPoint fp;
    float time;
    const Ray r = primaryRay(s, rc, pix[3 * i], pix[3 * i + 1], pix[3 * i + 2], __builtin_ctz(rc.spp), idx, fp, time);
    float* o = rays + 8 * static_cast<size_t>(i);
    o[0] = r.o.x, o[1] = r.o.y, o[2] = r.o.z, o[3] = r.d.x, o[4] = r.d.y, o[5] = r.d.z, o[6] = 1e-3f, o[7] = 1e5f;
    film[2 * i] = fp.x, film[2 * i + 1] = fp.y;
}

__global__ void kSobol(DevScene s, uint32_t m, uint32_t dims, const uint32_t* pix, uint32_t n, float* out,
                       uint64_t* indexOut) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t idx = intervalToIndex(s, m, pix[3 * i + 2], pix[3 * i], pix[3 * i + 1], 32u);
    if (indexOut) indexOut[i] = idx;
    for (uint32_t d = 0; d < dims; ++d) out[static_cast<size_t>(i) * dims + d] = sobolSample(s.sobol, idx, d);
}

// ------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------
static inline uint32_t blocksFor(uint64_t n, uint32_t bs, uint32_t cap) {
    const uint64_t b = (n + bs - 1) / bs;
    return static_cast<uint32_t>(b < 1 ? 1 : (b > cap ? cap : b));
}

// Persistent (grid-stride) kernels launch at most the blocks the device keeps resident at once
// for the kernel (its occupancy, set by VGPRs or LDS): a larger grid leaves a tail of blocks that
// start only as others end, which for traversal on a 1920x1080 64-spp frame cost 7-12 ms of ~91.
// gridLanes (for traversal: the lanes the spill region is sized for) stays the upper bound.
// dynLds: the launch's dynamic LDS bytes per block; perCuCap > 0 caps the blocks per CU.
static dim3 residentGrid(const void* kernel, uint32_t gridLanes, uint32_t blockSize, size_t dynLds, uint32_t perCuCap) {
    struct Fit {
        uint32_t cus, perCu;  // 0: unknown, keep the caller's grid
    };
    static std::mutex mu;
    static std::map<std::tuple<int, const void*, size_t>, Fit> cache;
    uint32_t blocks = gridLanes / blockSize;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return dim3(std::max(blocks, 1u));
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find({dev, kernel, dynLds});
    if (it == cache.end()) {
        int cus = 0, per = 0;
        Fit f{0, 0};
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, static_cast<int>(blockSize), dynLds) == hipSuccess &&
            cus > 0 && per > 0)
            f = Fit{static_cast<uint32_t>(cus), static_cast<uint32_t>(per)};
        it = cache.emplace(std::make_tuple(dev, kernel, dynLds), f).first;
    }
    const Fit& f = it->second;
    if (f.perCu) blocks = std::min(blocks, f.cus * (perCuCap ? std::min(perCuCap, f.perCu) : f.perCu));
    return dim3(std::max(blocks, 1u));
}
template <typename F>
static dim3 fit(F* kernel, uint32_t gridLanes, uint32_t blockSize = 64, size_t dynLds = 0, uint32_t perCuCap = 0) {
    return residentGrid(reinterpret_cast<const void*>(kernel), gridLanes, blockSize, dynLds, perCuCap);
}

void launchPoolInit(hipStream_t st, const Queues& q, uint32_t capacity) {
    hipLaunchKernelGGL(kPoolInit, dim3(blocksFor(capacity, 256, 1u << 30)), dim3(256), 0, st, q, capacity);
}
void launchIterBegin(hipStream_t st, const Queues& q, uint32_t curSlot) {
    hipLaunchKernelGGL(kIterBegin, dim3(1), dim3(1), 0, st, q, curSlot);
}
void launchRegen(hipStream_t st, const DevScene& s, const PathState& ps, const Queues& q, const RenderConsts& rc,
                 uint32_t curSlot, uint32_t gridThreads, const BatchTable* bt) {
    if (bt) hipLaunchKernelGGL(kRegen<true>, fit(kRegen<true>, gridThreads, 256), dim3(256), 0, st,
            s, ps, q, rc, curSlot, *bt);
    else hipLaunchKernelGGL(kRegen<false>, fit(kRegen<false>, gridThreads, 256), dim3(256), 0, st,
            s, ps, q, rc, curSlot, BatchTable{});
}

// Runtime flags to template arguments: f(std::integral_constant<bool, b>).
template <typename F>
static void withFlag(bool b, F&& f) {
    if (b) f(std::true_type{});
    else f(std::false_type{});
}
template <typename F>
static void withWidth(uint32_t w, F&& f) {
    if (w == 8) f(std::integral_constant<int, 8>{});
    else f(std::integral_constant<int, 4>{});
}
// Kernels with the moving-instance code run only for scenes (or batches) that have moving instances.
static bool anyMotion(const DevScene& s, const BatchTable* bt) { return bt ? bt->motion != 0 : s.numInst != 0; }

static size_t stackLds(const LaunchTune& lt) { return static_cast<size_t>(lt.ldsStack) * 64 * sizeof(uint2); }

void launchExtend(hipStream_t st, const DevScene& s, const PathState& ps, const Queues& q, uint32_t curSlot,
                  uint32_t gridLanes, uint2* spill, uint32_t spillPerLane, uint32_t* travCounts,
                  const BatchTable* bt, const LaunchTune& lt) {
    auto* tc = reinterpret_cast<unsigned long long*>(travCounts);
    const uint32_t* queue = q.ext[curSlot];
    const uint32_t* count = q.counts + curSlot;
    const BatchTable b = bt ? *bt : BatchTable{};
    const size_t lds = stackLds(lt);
    withWidth(s.width, [&](auto w) {
        withFlag(tc != nullptr, [&](auto cnt) {
            withFlag(bt != nullptr, [&](auto batch) {
                withFlag(anyMotion(s, bt), [&](auto mot) {
                    constexpr int W = decltype(w)::value;
                    constexpr bool C = decltype(cnt)::value, B = decltype(batch)::value, M = decltype(mot)::value;
                    hipLaunchKernelGGL((kExtend<W, C, B, M>), fit(kExtend<W, C, B, M>, gridLanes, 64, lds, lt.travWaves),
                                       dim3(64), lds, st, s, ps, queue, count, spill, spillPerLane, tc, b,
                                       static_cast<int>(lt.ldsStack));
                });
            });
        });
    });
}
void launchShade(hipStream_t st, const DevScene& s, const PathState& ps, const Queues& q, const RenderConsts& rc,
                 uint32_t curSlot, uint32_t nextSlot, uint32_t gridThreads, uint64_t* totals, const GuideShade& gs,
                 const BatchTable* bt, bool textured, const LaunchTune& lt) {
    const BatchTable b = bt ? *bt : BatchTable{};
    auto go = [&](auto blk) {
        withFlag(bt != nullptr, [&](auto batch) {  // a guided batch reads its per-item guides in bt.guide
            withFlag(gs.enabled != 0, [&](auto guided) {
                withFlag(textured, [&](auto tex) {
                    withFlag(anyMotion(s, bt), [&](auto mot) {
                        constexpr bool B = decltype(batch)::value, G = decltype(guided)::value,
                                       T = decltype(tex)::value, M = decltype(mot)::value;
                        constexpr uint32_t K = decltype(blk)::value;
                        hipLaunchKernelGGL((kShade<B, G, T, M, K>), fit(kShade<B, G, T, M, K>, gridThreads, K), dim3(K),
                                           0, st, s, ps, q, rc, curSlot, nextSlot, totals, gs, b);
                    });
                });
            });
        });
    };
    if (lt.shadeBlock == 128) go(std::integral_constant<uint32_t, 128>{});
    else go(std::integral_constant<uint32_t, 256>{});
}
void launchShadow(hipStream_t st, const DevScene& s, const PathState& ps, const Queues& q, uint32_t gridLanes,
                  uint2* spill, uint32_t spillPerLane, const BatchTable* bt, const LaunchTune& lt) {
    const BatchTable b = bt ? *bt : BatchTable{};
    const size_t lds = stackLds(lt);
    withWidth(s.width, [&](auto w) {
        withFlag(bt != nullptr, [&](auto batch) {
            withFlag(anyMotion(s, bt), [&](auto mot) {
                constexpr int W = decltype(w)::value;
                constexpr bool B = decltype(batch)::value, M = decltype(mot)::value;
                hipLaunchKernelGGL((kShadow<W, B, M>), fit(kShadow<W, B, M>, gridLanes, 64, lds, lt.travWaves), dim3(64),
                                   lds, st, s, ps, q, spill, spillPerLane, b, static_cast<int>(lt.ldsStack));
            });
        });
    });
}
void launchMis(hipStream_t st, const DevScene& s, const PathState& ps, const Queues& q, uint32_t gridLanes,
               uint2* spill, uint32_t spillPerLane, const BatchTable* bt, const LaunchTune& lt) {
    const BatchTable b = bt ? *bt : BatchTable{};
    const size_t lds = stackLds(lt);
    withWidth(s.width, [&](auto w) {
        withFlag(bt != nullptr, [&](auto batch) {
            withFlag(anyMotion(s, bt), [&](auto mot) {
                constexpr int W = decltype(w)::value;
                constexpr bool B = decltype(batch)::value, M = decltype(mot)::value;
                hipLaunchKernelGGL((kMis<W, B, M>), fit(kMis<W, B, M>, gridLanes, 64, lds, lt.travWaves), dim3(64), lds,
                                   st, s, ps, q, spill, spillPerLane, b, static_cast<int>(lt.ldsStack));
            });
        });
    });
}
void launchRecord(hipStream_t st, const PathState& ps, const Queues& q, const RenderConsts& rc, uint32_t gridThreads,
                  const RecordRing& ring, const GuideShade& gs, const GuideRecordOut& tr, const BatchTable* bt) {
    if (bt) hipLaunchKernelGGL(kRecord<true>, fit(kRecord<true>, gridThreads, 256), dim3(256), 0, st,
            ps, q, rc, ring, gs, tr, *bt);
    else hipLaunchKernelGGL(kRecord<false>, fit(kRecord<false>, gridThreads, 256), dim3(256), 0, st,
            ps, q, rc, ring, gs, tr, BatchTable{});
}
void launchResolve(hipStream_t st, const RenderConsts& rc, const RecordRing& ring, uint64_t first, void* quad,
                   float* film32, double* film64) {
    const uint32_t pixels = rc.rectW * rc.rectH;
    const uint32_t targets = (rc.rectW + 2) * (rc.rectH + 2);
    const uint64_t start = first % ring.size;
    uint32_t g = 1;  // lanes per pixel in kQuadSums
    while (g < 64 && g < rc.sppCount) g <<= 1;
    const uint32_t sumBlocks = static_cast<uint32_t>((static_cast<uint64_t>(pixels) * g + 255) / 256);
    if (rc.filter != kFilterBox) {  // weighted footprints
        if (film64) {
            hipLaunchKernelGGL(kBinSums<double>, dim3(sumBlocks), dim3(256), 0, st, rc, ring.L, ring.y, start, ring.size,
                               g, static_cast<double*>(quad));
            hipLaunchKernelGGL(kResolveBins<double>, dim3(blocksFor(targets, 256, 1u << 30)), dim3(256), 0, st, rc,
                               static_cast<const double*>(quad), film64);
        } else {
            hipLaunchKernelGGL(kBinSums<float>, dim3(sumBlocks), dim3(256), 0, st, rc, ring.L, ring.y, start, ring.size,
                               g, static_cast<float*>(quad));
            hipLaunchKernelGGL(kResolveBins<float>, dim3(blocksFor(targets, 256, 1u << 30)), dim3(256), 0, st, rc,
                               static_cast<const float*>(quad), film32);
        }
        return;
    }
    if (film64) {
        hipLaunchKernelGGL(kQuadSums<double>, dim3(sumBlocks), dim3(256), 0, st, rc, ring.L, ring.y, start, ring.size, g,
                           static_cast<double*>(quad));
        hipLaunchKernelGGL(kResolve<double>, dim3(blocksFor(targets, 256, 1u << 30)), dim3(256), 0, st, rc,
                           static_cast<const double*>(quad), film64);
    } else {
        hipLaunchKernelGGL(kQuadSums<float>, dim3(sumBlocks), dim3(256), 0, st, rc, ring.L, ring.y, start, ring.size, g,
                           static_cast<float*>(quad));
        hipLaunchKernelGGL(kResolve<float>, dim3(blocksFor(targets, 256, 1u << 30)), dim3(256), 0, st, rc,
                           static_cast<const float*>(quad), film32);
    }
}
void launchTrace(hipStream_t st, const DevScene& s, const float* rays, uint32_t n, float* t, uint32_t* prim, float* uv,
                 uint8_t* occ, uint32_t gridLanes, uint2* spill, uint32_t spillPerLane, const LaunchTune& lt,
                 const float* times) {
    const size_t lds = stackLds(lt);
    withWidth(s.width, [&](auto w) {
        withFlag(s.numInst != 0, [&](auto mot) {
            constexpr int W = decltype(w)::value;
            constexpr bool M = decltype(mot)::value;
            hipLaunchKernelGGL((kTrace<W, M>), fit(kTrace<W, M>, gridLanes, 64, lds, lt.travWaves), dim3(64), lds, st, s,
                               rays, n, t, prim, uv, occ, spill, spillPerLane, times, static_cast<int>(lt.ldsStack));
        });
    });
}
void launchCameraRays(hipStream_t st, const DevScene& s, const RenderConsts& rc, const uint32_t* pix, uint32_t n,
                      float* rays, float* film) {
    hipLaunchKernelGGL(kCameraRays, dim3(blocksFor(n, 256, 1u << 30)), dim3(256), 0, st, s, rc, pix, n, rays, film);
}
void launchCopy(hipStream_t st, const void* src, void* dst, size_t bytes) {
    const size_t n = bytes / 16;
    hipLaunchKernelGGL(kCopy, dim3(blocksFor(n, 256, 8192)), dim3(256), 0, st, static_cast<const uint4*>(src),
                       static_cast<uint4*>(dst), n);
}
void launchSobol(hipStream_t st, const DevScene& s, uint32_t m, uint32_t dims, const uint32_t* pix, uint32_t n,
                 float* out, uint64_t* indexOut) {
    hipLaunchKernelGGL(kSobol, dim3(blocksFor(n, 256, 1u << 30)), dim3(256), 0, st, s, m, dims, pix, n, out, indexOut);
}

}  // namespace pg

// gfx950 kernels of the wavefront path tracer and their launchers.
//
// One render call processes the (pixel, sample) paths of a rect in waves of `capacity` paths.
// Per wave: raygen, then per bounce extend (closest hit) -> shade (emission, NEE, MIS-BSDF ray,
// continuation, Russian roulette) -> shadow (any hit) -> mis (closest hit, selected light test)
// -> film (splat of finished paths). The stages communicate through queues compacted with one
// atomic per wave (ballot + popcount prefix). The per-bounce structure restates trace() and
// multipleImportanceSampling() of Integrator/PathTracing/Kernel.cpp:23-193; the per-sample driver
// restates piperMain (Tracer/Embree/Kernel.cpp:47-100).
#include "pg_device.hpp"
#include "pg_kernels.hpp"

namespace pg {

// ------------------------------------------------------------------------------------------
// raygen: piperMain up to rayGen (Tracer/Embree/Kernel.cpp:47-90)
// ------------------------------------------------------------------------------------------
struct FilmPoint {
    float x, y;
};

__device__ __forceinline__ void pathPixel(const RenderConsts& rc, uint32_t g, uint32_t& px, uint32_t& py,
                                          uint32_t& si) {
    si = rc.sppBegin + g % rc.sppCount;
    const uint32_t pix = g / rc.sppCount;
    px = rc.rectX + pix % rc.rectW;
    py = rc.rectY + pix / rc.rectW;
}

__device__ __forceinline__ Ray primaryRay(const DevScene& s, const RenderConsts& rc, uint32_t px, uint32_t py,
                                          uint32_t si, uint64_t& idx, FilmPoint& fp) {
    idx = intervalToIndex(s, rc.log2Res, si, px, py);
    const float res = static_cast<float>(1u << rc.log2Res);
    // sobolStart (ScrambledSobolSampler/Kernel.cpp:52-62)
    float fx = fminf(fmaxf(sobolSample(s.sobol, idx, 0) * res - static_cast<float>(px), 0.0f), 1.0f);
    float fy = fminf(fmaxf(sobolSample(s.sobol, idx, 1) * res - static_cast<float>(py), 0.0f), 1.0f);
    fx += static_cast<float>(px);
    fy += static_cast<float>(py);
    fp = FilmPoint{fx, fy};
    const float ndcX = rc.ndc[0] + rc.ndc[2] * fx / static_cast<float>(rc.width);
    const float ndcY = rc.ndc[1] + rc.ndc[3] * fy / static_cast<float>(rc.height);
    // dim 2 is the time sample (unused by static scenes); dims 3, 4 drive the lens
    return cameraRay(s.cam, ndcX, ndcY, sobolSample(s.sobol, idx, 3), sobolSample(s.sobol, idx, 4));
}

__global__ void __launch_bounds__(256) kRaygen(DevScene s, PathState ps, Queues q, RenderConsts rc, uint64_t first,
                                               uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        q.counts[kCntExtCur] = n;
        q.counts[kCntExtNext] = 0;
        q.counts[kCntShadow] = q.counts[kCntMis] = q.counts[kCntDone] = 0;
    }
    if (i >= n) return;
    const uint32_t g = static_cast<uint32_t>(first + i);
    uint32_t px, py, si;
    pathPixel(rc, g, px, py, si);
    uint64_t idx;
    FilmPoint fp;
    const Ray r = primaryRay(s, rc, px, py, si, idx, fp);
    ps.o[i] = make_float4(r.o.x, r.o.y, r.o.z, 0.0f);
    ps.d[i] = make_float4(r.d.x, r.d.y, r.d.z, 0.0f);
    ps.thr[i] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    ps.L[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    ps.Lmis[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    ps.idx[i] = idx;
    ps.meta[i] = make_uint4(5u, 0u, 1u, g);  // sample dims start at 5 (Embree/Kernel.cpp:68); specular
    ps.film[i] = make_float2(fp.x, fp.y);
    q.ext[0][i] = i;
}

// Counter reset at the start of a bounce (a kernel, not a memset, so the sequence can be captured).
__global__ void kBounceBegin(Queues q, uint32_t nextSlot) {
    q.counts[nextSlot] = 0;
    q.counts[kCntShadow] = q.counts[kCntMis] = q.counts[kCntDone] = 0;
}

// ------------------------------------------------------------------------------------------
// extend: closest hit for every queued path
// ------------------------------------------------------------------------------------------
template <int W, bool COUNT>
__global__ void __launch_bounds__(64) kExtend(DevScene s, PathState ps, const uint32_t* __restrict__ queue,
                                              const uint32_t* __restrict__ count, uint2* spill, uint32_t spillPerLane,
                                              unsigned long long* travCounts) {
    __shared__ uint2 lstack[kLdsStack * 64];
    const uint32_t n = *count;
    const uint32_t lane = threadIdx.x;
    const uint32_t gtid = blockIdx.x * 64 + lane;
    uint2* gs = spill + static_cast<size_t>(gtid) * spillPerLane;
    TravCount cnt;
    for (uint32_t i = gtid; i < n; i += gridDim.x * 64) {
        const uint32_t path = queue[i];
        const float4 o = ps.o[path], d = ps.d[path];
        const Ray r{xyz(o), xyz(d)};
        const HitRec h = closestHit<W, COUNT>(s, r, 1e-3f, 1e5f, lstack + lane, gs, cnt);
        ps.hit[path] = make_float4(h.t, __uint_as_float(h.pos), h.u, h.v);
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

__global__ void __launch_bounds__(256) kShade(DevScene s, PathState ps, Queues q, RenderConsts rc, uint32_t curSlot,
                                              uint32_t nextSlot, uint64_t* totals) {
    const uint32_t n = q.counts[curSlot];
    const uint32_t* qcur = q.ext[curSlot];
    uint32_t* qnext = q.ext[nextSlot];
    const bool d1 = rc.flags & kCompatD1, d3 = rc.flags & kCompatD3, d9 = rc.flags & kCompatD9,
               d13 = rc.flags & kCompatD13;
    const bool envBlack = !valid(ld3(s.env));
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < n; i0 += stride) {  // block-uniform loop
        const uint32_t i = i0 + threadIdx.x;
        bool pushExt = false, pushShadow = false, pushMis = false, pushMisQuery = false, pushDone = false;
        uint32_t path = 0;
        ShadowRec sr, mq;
        MisRec mr;
        if (i < n) {
            path = qcur[i];
            const float4 hit4 = ps.hit[path];
            const uint32_t pos = __float_as_uint(hit4.y);
            uint4 meta = ps.meta[path];
            Ray ray{xyz(ps.o[path]), xyz(ps.d[path])};
            V3 thr = xyz(ps.thr[path]);
            // fold in last bounce's MIS-BSDF light hit (after its NEE term, as the reference adds them)
            V3 L = xyz(ps.L[path]) + xyz(ps.Lmis[path]);
            ps.Lmis[path] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            const bool specular = meta.z & 1u;
            const bool isLight = pos != kHitMiss && (pos & kHitQuad);
            if (specular && (isLight || pos == kHitMiss)) {  // :119-140, two-sided emission
                const V3 rad = isLight ? ld3(s.quads[pos & ~kHitQuad].radiance) : ld3(s.env);
                L = L + mul(thr, rad);
            }
            if (isLight) {  // :142-146 pass through the light
                if (d13) meta.y -= 1u;
                ray.o = ray.o + ray.d * (hit4.x + 0.001f);
                ps.o[path] = make_float4(ray.o.x, ray.o.y, ray.o.z, 0.0f);
                pushExt = true;
            } else if (pos == kHitMiss || meta.y >= rc.maxDepth) {  // :148
                pushDone = true;
            } else {
                const Mat mat = loadMat(s, s.triMat[pos]);
                const V3 nrm = triNormal(s, pos, ray.d);
                const Frame fr = shadingFrame(nrm, d3);
                const V3 wo = fr.toLocal(-ray.d);
                const V3 ng = fr.toLocal(nrm);
                const V3 hp = ray.o + ray.d * hit4.x;
                // The vertex draws at most 8 consecutive dimensions: light select, light (2),
                // MIS BSDF (2), continuation (2), Russian roulette (1), in that order.
                float u[8];
                sobolWindow(s, ps.idx[path], meta.x, u);
                uint32_t used = 7;

                // light selection: uniformSelect over {environment, quads}
                const uint32_t count = 1u + s.numQuads;
                const uint32_t sel = min(static_cast<uint32_t>(floorf(u[0] * static_cast<float>(count))), count - 1u);
                const float selInv = 1.0f / (1.0f / static_cast<float>(count));
                const LightSample ls = sampleLight(s, sel, hp, d9 ? u[2] : u[1], d9 ? u[1] : u[2], d1);
                if (ls.pdf > 0.0f && valid(ls.rad)) {
                    const V3 wi = fr.toLocal(ls.dir);
                    const V3 f = bsdfEval(mat, wo, wi, ng);
                    if (valid(f)) {
                        float w = fabsf(wi.z) / ls.pdf;
                        w = w * powerHeuristic(ls.pdf, bsdfPdf(mat, wo, wi));
                        const V3 c = mul(thr, mul(f, ls.rad) * w * selInv);
                        sr = ShadowRec{hp.x, hp.y, hp.z, fminf(1e-3f, ls.dist), ls.dir.x, ls.dir.y, ls.dir.z, ls.dist,
                                       c.x, c.y, c.z, path};
                        pushShadow = true;
                    }
                }
                const BsdfSample bs = bsdfSample(mat, wo, d9 ? u[4] : u[3], d9 ? u[3] : u[4]);
                if (bs.pdf > 0.0f && valid(bs.f)) {
                    const float weight = fabsf(bs.wi.z) / bs.pdf;
                    const V3 dir = fr.toWorld(bs.wi);
                    const V3 c = mul(thr, bs.f * weight * selInv);
                    if (!envBlack) {
                        mr = MisRec{hp.x, hp.y, hp.z, bs.pdf, dir.x, dir.y, dir.z, sel, c.x, c.y, c.z, path};
                        pushMis = true;
                    } else if (sel != 0u) {
                        // With a black environment the BSDF ray contributes only if the selected
                        // quad is its closest hit: the quad must be hit, beat every other quad in
                        // (t, index) order (closestHit's quad loop), and no triangle may lie in
                        // (tmin, tq] (triangles win ties). The first two are decided here; the
                        // triangle test becomes an occlusion query on the shadow queue.
                        const Ray mray{hp, dir};
                        const uint32_t qi = sel - 1u;
                        const DevQuad& qd = s.quads[qi];
                        float tq;
                        if (intersectQuad(mray, qd, 1e-3f, 1e5f, tq)) {
                            bool first = true;
                            for (uint32_t k = 0; k < s.numQuads && first; ++k) {
                                float t2;
                                if (k != qi && intersectQuad(mray, s.quads[k], 1e-3f, 1e5f, t2) &&
                                    (t2 < tq || (t2 == tq && k < qi)))
                                    first = false;
                            }
                            if (first) {
                                const V3 qn = quadNormal(qd, dir);
                                const V3 rad = dot(qn, dir) < 0.0f ? ld3(qd.radiance) : V3{0.0f, 0.0f, 0.0f};
                                const float lpdf = tq * tq / (fabsf(dot(qn, dir)) * qd.area);
                                if (valid(rad) && lpdf > 0.0f) {
                                    const V3 cm = mul(c, rad) * powerHeuristic(bs.pdf, lpdf);
                                    // tq > 0: the next float up makes the range (tmin, tq] inclusive
                                    const float tqUp = __uint_as_float(__float_as_uint(tq) + 1u);
                                    mq = ShadowRec{hp.x, hp.y, hp.z, 1e-3f, dir.x, dir.y, dir.z, tqUp,
                                                   cm.x, cm.y, cm.z, path | kMisFlag};
                                    pushMisQuery = true;
                                }
                            }
                        }
                    }
                }

                const BsdfSample ss = bsdfSample(mat, wo, d9 ? u[6] : u[5], d9 ? u[5] : u[6]);  // :164-190
                if (!(ss.pdf > 0.0f && valid(ss.f))) {
                    pushDone = true;
                } else {
                    thr = mul(thr, ss.f * (fabsf(ss.wi.z) / ss.pdf));
                    bool alive = true;
                    if (meta.y > rc.minRR) {
                        const float qrr = fmaxf(0.05f, 1.0f - maxc(thr));
                        used = 8;
                        if (u[7] < qrr) alive = false;
                        else thr = thr * (1.0f / (1.0f - qrr));
                    }
                    if (alive) {
                        const V3 nd = fr.toWorld(ss.wi);
                        const V3 no = hp + nrm * 0.001f;
                        ps.o[path] = make_float4(no.x, no.y, no.z, 0.0f);
                        ps.d[path] = make_float4(nd.x, nd.y, nd.z, 0.0f);
                        ps.thr[path] = make_float4(thr.x, thr.y, thr.z, 0.0f);
                        meta.y += 1u;
                        meta.z = 0u;
                        pushExt = true;
                    } else {
                        pushDone = true;
                    }
                }
                meta.x += used;
            }
            ps.L[path] = make_float4(L.x, L.y, L.z, 0.0f);
            ps.meta[path] = meta;
        }
        const uint32_t se = wavePush(&q.counts[nextSlot], pushExt);
        if (pushExt) qnext[se] = path;
        const uint32_t ss = wavePush(&q.counts[kCntShadow], pushShadow);
        if (pushShadow) q.shadow[ss] = sr;
        const uint32_t sq = wavePush(&q.counts[kCntShadow], pushMisQuery);
        if (pushMisQuery) q.shadow[sq] = mq;
        const uint32_t sm = wavePush(&q.counts[kCntMis], pushMis);
        if (pushMis) q.mis[sm] = mr;
        const uint32_t sd = wavePush(&q.counts[kCntDone], pushDone);
        if (pushDone) q.done[sd] = path;
        if (totals) {
            const uint64_t me = __ballot(pushExt), msh = __ballot(pushShadow), mm = __ballot(pushMis || pushMisQuery);
            if (__lane_id() == 0) {
                if (me) atomicAdd(reinterpret_cast<unsigned long long*>(&totals[kTotExt]), __popcll(me));
                if (msh) atomicAdd(reinterpret_cast<unsigned long long*>(&totals[kTotShadow]), __popcll(msh));
                if (mm) atomicAdd(reinterpret_cast<unsigned long long*>(&totals[kTotMis]), __popcll(mm));
            }
        }
    }
}

// ------------------------------------------------------------------------------------------
// shadow: occlusion test of NEE rays (sampleLightSource, PathTracing/Kernel.cpp:40-62)
// ------------------------------------------------------------------------------------------
template <int W>
__global__ void __launch_bounds__(64) kShadow(DevScene s, PathState ps, Queues q, uint2* spill, uint32_t spillPerLane) {
    __shared__ uint2 lstack[kLdsStack * 64];
    const uint32_t n = q.counts[kCntShadow];
    const uint32_t lane = threadIdx.x;
    const uint32_t gtid = blockIdx.x * 64 + lane;
    uint2* gs = spill + static_cast<size_t>(gtid) * spillPerLane;
    for (uint32_t i = gtid; i < n; i += gridDim.x * 64) {
        const ShadowRec r = q.shadow[i];
        const Ray ray{{r.ox, r.oy, r.oz}, {r.dx, r.dy, r.dz}};
        if (!occluded<W>(s, ray, r.tmin, r.tmax, lstack + lane, gs)) {
            if (r.path & kMisFlag) {  // MIS-BSDF light hit: kept apart, folded in after the NEE term
                ps.Lmis[r.path & ~kMisFlag] = make_float4(r.cr, r.cg, r.cb, 0.0f);
            } else {
                float4 L = ps.L[r.path];
                L.x += r.cr, L.y += r.cg, L.z += r.cb;
                ps.L[r.path] = L;
            }
        }
    }
}

// ------------------------------------------------------------------------------------------
// mis: the separate BSDF-sampled ray of sampleBSDF (PathTracing/Kernel.cpp:63-102); counts only
// if it hits the selected light (or escapes to the environment)
// ------------------------------------------------------------------------------------------
template <int W>
__global__ void __launch_bounds__(64) kMis(DevScene s, PathState ps, Queues q, uint2* spill, uint32_t spillPerLane) {
    __shared__ uint2 lstack[kLdsStack * 64];
    const uint32_t n = q.counts[kCntMis];
    const uint32_t lane = threadIdx.x;
    const uint32_t gtid = blockIdx.x * 64 + lane;
    uint2* gs = spill + static_cast<size_t>(gtid) * spillPerLane;
    const bool envBlack = !valid(ld3(s.env));
    for (uint32_t i = gtid; i < n; i += gridDim.x * 64) {
        const MisRec r = q.mis[i];
        const Ray ray{{r.ox, r.oy, r.oz}, {r.dx, r.dy, r.dz}};
        V3 rad{0.0f, 0.0f, 0.0f};
        float lpdf = 0.0f;
        if (envBlack && r.light != 0u) {
            // With a black environment the ray contributes only if the selected quad is its
            // closest hit. Same decision as closestHit(): the quad must be hit, beat every other
            // quad in (t, index) order, and no triangle may lie in (tmin, t_q] (triangles win ties).
            const uint32_t sel = r.light - 1u;
            const DevQuad& qd = s.quads[sel];
            float tq;
            if (intersectQuad(ray, qd, 1e-3f, 1e5f, tq)) {
                bool first = true;
                for (uint32_t k = 0; k < s.numQuads && first; ++k) {
                    float t2;
                    if (k != sel && intersectQuad(ray, s.quads[k], 1e-3f, 1e5f, t2) && (t2 < tq || (t2 == tq && k < sel)))
                        first = false;
                }
                // tq > 0: the next float up makes the occlusion range (tmin, tq] inclusive
                const float tqUp = __uint_as_float(__float_as_uint(tq) + 1u);
                if (first && !occluded<W>(s, ray, 1e-3f, tqUp, lstack + lane, gs)) {
                    const V3 qn = quadNormal(qd, ray.d);
                    rad = dot(qn, ray.d) < 0.0f ? ld3(qd.radiance) : V3{0.0f, 0.0f, 0.0f};
                    lpdf = tq * tq / (fabsf(dot(qn, ray.d)) * qd.area);
                }
            }
        } else {
            TravCount tc;
            const HitRec h = closestHit<W>(s, ray, 1e-3f, 1e5f, lstack + lane, gs, tc);
            if (h.pos == kHitMiss) {
                rad = ld3(s.env);
                lpdf = envPdf(ray.d);
            } else if ((h.pos & kHitQuad) && r.light == 1u + (h.pos & ~kHitQuad)) {
                const DevQuad& qd = s.quads[h.pos & ~kHitQuad];
                const V3 qn = quadNormal(qd, ray.d);
                rad = dot(qn, ray.d) < 0.0f ? ld3(qd.radiance) : V3{0.0f, 0.0f, 0.0f};
                lpdf = h.t * h.t / (fabsf(dot(qn, ray.d)) * qd.area);
            }
        }
        if (valid(rad) && lpdf > 0.0f) {
            const float w = powerHeuristic(r.bsdfPdf, lpdf);
            const V3 c = mul(V3{r.cr, r.cg, r.cb}, rad) * w;
            float4 L = ps.L[r.path];
            L.x += c.x, L.y += c.y, L.z += c.z;
            ps.L[r.path] = L;
        }
    }
}

// ------------------------------------------------------------------------------------------
// film: accumulate with the box filter, a 2x2 splat into RGBW (FixedSampler/Kernel.cpp:30-58)
// ------------------------------------------------------------------------------------------
// Box-filter splat of an aggregated group: `n` samples whose 2x2 footprint starts at (bx, by)
// add their summed radiance and a weight of n to each covered pixel.
template <typename F>
__device__ __forceinline__ void splatGroup(F* rgbw, uint32_t w, uint32_t h, int32_t bx, int32_t by, F r, F g, F b,
                                           F n) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t x = static_cast<uint32_t>(bx + (k & 1)), y = static_cast<uint32_t>(by + (k >> 1));
        if (!(x < w && y < h)) continue;
        F* p = rgbw + 4 * (static_cast<size_t>(y) * w + x);
        unsafeAtomicAdd(p + 0, r);
        unsafeAtomicAdd(p + 1, g);
        unsafeAtomicAdd(p + 2, b);
        unsafeAtomicAdd(p + 3, n);
    }
}

template <typename F>
__device__ __forceinline__ F waveSum(F v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// accumulate() with the box filter (FixedSampler/Kernel.cpp:30-58): every sample adds (L, 1) to the
// 2x2 pixels from (int)(p - 0.5). Samples of one pixel sit in neighbouring lanes, so lanes that
// share a footprint are summed across the wave first and one lane issues the atomics (at most
// kFilmGroups rounds per wave; any lanes left after that add their own sample).
constexpr int kFilmGroups = 8;

template <typename F>
__device__ __forceinline__ void filmAccumulate(F* rgbw, uint32_t w, uint32_t h, bool active, float fx, float fy, V3 L) {
    const int32_t bx = static_cast<int32_t>(fx - 0.5f), by = static_cast<int32_t>(fy - 0.5f);
    const uint32_t key = active ? static_cast<uint32_t>(by) * (w + 1u) + static_cast<uint32_t>(bx) : 0xFFFFFFFFu;
    const uint32_t lane = __lane_id();
    uint64_t pending = __ballot(active);
    for (int it = 0; it < kFilmGroups && pending; ++it) {
        const int leader = __builtin_ctzll(pending);
        const uint32_t lk = __shfl(key, leader);
        const bool match = active && key == lk;
        const uint64_t mask = __ballot(match);
        const F r = waveSum<F>(match ? static_cast<F>(L.x) : F(0));
        const F g = waveSum<F>(match ? static_cast<F>(L.y) : F(0));
        const F b = waveSum<F>(match ? static_cast<F>(L.z) : F(0));
        if (lane == static_cast<uint32_t>(leader))
            splatGroup<F>(rgbw, w, h, bx, by, r, g, b, static_cast<F>(__popcll(mask)));
        pending &= ~mask;
    }
    if ((pending >> lane) & 1ull)
        splatGroup<F>(rgbw, w, h, bx, by, static_cast<F>(L.x), static_cast<F>(L.y), static_cast<F>(L.z), F(1));
}

__global__ void __launch_bounds__(256) kFilm(PathState ps, Queues q, RenderConsts rc, float* film32, double* film64,
                                             float* pathL) {
    const uint32_t n = q.counts[kCntDone];
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x) {  // block-uniform loop
        const uint32_t i = i0 + threadIdx.x;
        const bool active = i < n;
        uint32_t path = 0;
        V3 L{0.0f, 0.0f, 0.0f};
        float2 fp = make_float2(0.0f, 0.0f);
        if (active) {
            path = q.done[i];
            L = xyz(ps.L[path]) + xyz(ps.Lmis[path]);
            fp = ps.film[path];
            if (pathL) {
                const uint32_t g = ps.meta[path].w;
                pathL[3 * static_cast<size_t>(g)] = L.x;
                pathL[3 * static_cast<size_t>(g) + 1] = L.y;
                pathL[3 * static_cast<size_t>(g) + 2] = L.z;
            }
        }
        if (film64) filmAccumulate<double>(film64, rc.width, rc.height, active, fp.x, fp.y, L);
        else if (film32) filmAccumulate<float>(film32, rc.width, rc.height, active, fp.x, fp.y, L);
    }
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
template <int W>
__global__ void __launch_bounds__(64) kTrace(DevScene s, const float* rays, uint32_t n, float* t, uint32_t* prim,
                                             float* uv, uint8_t* occ, uint2* spill, uint32_t spillPerLane) {
    __shared__ uint2 lstack[kLdsStack * 64];
    const uint32_t lane = threadIdx.x;
    const uint32_t gtid = blockIdx.x * 64 + lane;
    uint2* gs = spill + static_cast<size_t>(gtid) * spillPerLane;
    for (uint32_t i = gtid; i < n; i += gridDim.x * 64) {
        const float* r = rays + 8 * static_cast<size_t>(i);
        const Ray ray{{r[0], r[1], r[2]}, {r[3], r[4], r[5]}};
        if (occ) {
            occ[i] = occluded<W>(s, ray, r[6], r[7], lstack + lane, gs) ? 1 : 0;
        } else {
            TravCount tc;
            const HitRec h = closestHit<W>(s, ray, r[6], r[7], lstack + lane, gs, tc);
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
    const Ray r = primaryRay(s, rc, pix[3 * i], pix[3 * i + 1], pix[3 * i + 2], idx, fp);
    float* o = rays + 8 * static_cast<size_t>(i);
    o[0] = r.o.x, o[1] = r.o.y, o[2] = r.o.z, o[3] = r.d.x, o[4] = r.d.y, o[5] = r.d.z, o[6] = 1e-3f, o[7] = 1e5f;
    film[2 * i] = fp.x, film[2 * i + 1] = fp.y;
}

__global__ void kSobol(DevScene s, uint32_t m, uint32_t dims, const uint32_t* pix, uint32_t n, float* out,
                       uint64_t* indexOut) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t idx = intervalToIndex(s, m, pix[3 * i + 2], pix[3 * i], pix[3 * i + 1]);
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

void launchRaygen(hipStream_t st, const DevScene& s, const PathState& ps, const Queues& q, const RenderConsts& rc,
                  uint64_t first, uint32_t n) {
    hipLaunchKernelGGL(kRaygen, dim3(blocksFor(n, 256, 1u << 30)), dim3(256), 0, st, s, ps, q, rc, first, n);
}
void launchBounceBegin(hipStream_t st, const Queues& q, uint32_t nextSlot) {
    hipLaunchKernelGGL(kBounceBegin, dim3(1), dim3(1), 0, st, q, nextSlot);
}
void launchExtend(hipStream_t st, const DevScene& s, const PathState& ps, const Queues& q, uint32_t curSlot,
                  uint32_t gridLanes, uint2* spill, uint32_t spillPerLane, uint32_t* travCounts) {
    const dim3 grid(gridLanes / 64);
    auto* tc = reinterpret_cast<unsigned long long*>(travCounts);
    const uint32_t* queue = q.ext[curSlot];
    const uint32_t* count = q.counts + curSlot;
    if (s.width == 8) {
        if (tc) hipLaunchKernelGGL((kExtend<8, true>), grid, dim3(64), 0, st, s, ps, queue, count, spill, spillPerLane, tc);
        else hipLaunchKernelGGL((kExtend<8, false>), grid, dim3(64), 0, st, s, ps, queue, count, spill, spillPerLane, tc);
    } else {
        if (tc) hipLaunchKernelGGL((kExtend<4, true>), grid, dim3(64), 0, st, s, ps, queue, count, spill, spillPerLane, tc);
        else hipLaunchKernelGGL((kExtend<4, false>), grid, dim3(64), 0, st, s, ps, queue, count, spill, spillPerLane, tc);
    }
}
void launchShade(hipStream_t st, const DevScene& s, const PathState& ps, const Queues& q, const RenderConsts& rc,
                 uint32_t curSlot, uint32_t nextSlot, uint32_t gridThreads, uint64_t* totals) {
    hipLaunchKernelGGL(kShade, dim3(gridThreads / 256), dim3(256), 0, st, s, ps, q, rc, curSlot, nextSlot, totals);
}
void launchShadow(hipStream_t st, const DevScene& s, const PathState& ps, const Queues& q, uint32_t gridLanes,
                  uint2* spill, uint32_t spillPerLane) {
    const dim3 grid(gridLanes / 64);
    if (s.width == 8)
        hipLaunchKernelGGL(kShadow<8>, grid, dim3(64), 0, st, s, ps, q, spill, spillPerLane);
    else
        hipLaunchKernelGGL(kShadow<4>, grid, dim3(64), 0, st, s, ps, q, spill, spillPerLane);
}
void launchMis(hipStream_t st, const DevScene& s, const PathState& ps, const Queues& q, uint32_t gridLanes,
               uint2* spill, uint32_t spillPerLane) {
    const dim3 grid(gridLanes / 64);
    if (s.width == 8)
        hipLaunchKernelGGL(kMis<8>, grid, dim3(64), 0, st, s, ps, q, spill, spillPerLane);
    else
        hipLaunchKernelGGL(kMis<4>, grid, dim3(64), 0, st, s, ps, q, spill, spillPerLane);
}
void launchFilm(hipStream_t st, const PathState& ps, const Queues& q, const RenderConsts& rc, uint32_t gridThreads,
                float* film32, double* film64, float* pathL) {
    hipLaunchKernelGGL(kFilm, dim3(gridThreads / 256), dim3(256), 0, st, ps, q, rc, film32, film64, pathL);
}
void launchTrace(hipStream_t st, const DevScene& s, const float* rays, uint32_t n, float* t, uint32_t* prim, float* uv,
                 uint8_t* occ, uint32_t gridLanes, uint2* spill, uint32_t spillPerLane) {
    const dim3 grid(gridLanes / 64);
    if (s.width == 8)
        hipLaunchKernelGGL(kTrace<8>, grid, dim3(64), 0, st, s, rays, n, t, prim, uv, occ, spill, spillPerLane);
    else
        hipLaunchKernelGGL(kTrace<4>, grid, dim3(64), 0, st, s, rays, n, t, prim, uv, occ, spill, spillPerLane);
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

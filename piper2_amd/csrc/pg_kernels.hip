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
template <int W>
__global__ void __launch_bounds__(64) kExtend(DevScene s, PathState ps, const uint32_t* __restrict__ queue,
                                              const uint32_t* __restrict__ count, uint2* spill, uint32_t spillPerLane,
                                              unsigned long long* travCounts) {
    __shared__ uint2 lstack[kLdsStack * 64];
    const uint32_t n = *count;
    const uint32_t lane = threadIdx.x;
    const uint32_t gtid = blockIdx.x * 64 + lane;
    uint2* gs = spill + static_cast<size_t>(gtid) * spillPerLane;
    uint32_t nodes = 0, tris = 0;
    for (uint32_t i = gtid; i < n; i += gridDim.x * 64) {
        const uint32_t path = queue[i];
        const float4 o = ps.o[path], d = ps.d[path];
        const Ray r{xyz(o), xyz(d)};
        const HitRec h = closestHit<W>(s, r, 1e-3f, 1e5f, lstack + lane, gs, travCounts ? &nodes : nullptr,
                                       travCounts ? &tris : nullptr);
        ps.hit[path] = make_float4(h.t, __uint_as_float(h.pos), h.u, h.v);
    }
    if (travCounts) {
        atomicAdd(&travCounts[0], static_cast<unsigned long long>(nodes));
        atomicAdd(&travCounts[1], static_cast<unsigned long long>(tris));
    }
}

// ------------------------------------------------------------------------------------------
// shade: one vertex of trace() (PathTracing/Kernel.cpp:115-190) including
// multipleImportanceSampling (:23-107); shadow and MIS-BSDF rays go to their queues.
// ------------------------------------------------------------------------------------------
struct SobolStream {
    const uint32_t* m;
    uint64_t idx;
    uint32_t dim;
    __device__ __forceinline__ float next() { return sobolSample(m, idx, dim++); }
    __device__ __forceinline__ void next2(bool d9, float& u1, float& u2) {
        const float a = next(), b = next();
        if (d9) u1 = b, u2 = a;
        else u1 = a, u2 = b;
    }
};

__global__ void __launch_bounds__(256) kShade(DevScene s, PathState ps, Queues q, RenderConsts rc, uint32_t curSlot,
                                              uint32_t nextSlot, uint64_t* totals) {
    const uint32_t n = q.counts[curSlot];
    const uint32_t* qcur = q.ext[curSlot];
    uint32_t* qnext = q.ext[nextSlot];
    const bool d1 = rc.flags & kCompatD1, d3 = rc.flags & kCompatD3, d9 = rc.flags & kCompatD9,
               d13 = rc.flags & kCompatD13;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < n; i0 += stride) {  // block-uniform loop
        const uint32_t i = i0 + threadIdx.x;
        bool pushExt = false, pushShadow = false, pushMis = false, pushDone = false;
        uint32_t path = 0;
        ShadowRec sr;
        MisRec mr;
        if (i < n) {
            path = qcur[i];
            const float4 hit4 = ps.hit[path];
            const uint32_t pos = __float_as_uint(hit4.y);
            uint4 meta = ps.meta[path];
            Ray ray{xyz(ps.o[path]), xyz(ps.d[path])};
            V3 thr = xyz(ps.thr[path]);
            V3 L = xyz(ps.L[path]);
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
                SobolStream smp{s.sobol, ps.idx[path], meta.x};

                // light selection: uniformSelect over {environment, quads}
                const uint32_t count = 1u + s.numQuads;
                const uint32_t sel =
                    min(static_cast<uint32_t>(floorf(smp.next() * static_cast<float>(count))), count - 1u);
                const float selInv = 1.0f / (1.0f / static_cast<float>(count));
                float u1, u2;
                smp.next2(d9, u1, u2);
                const LightSample ls = sampleLight(s, sel, hp, u1, u2, d1);
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
                smp.next2(d9, u1, u2);
                const BsdfSample bs = bsdfSample(mat, wo, u1, u2);
                if (bs.pdf > 0.0f && valid(bs.f)) {
                    const float weight = fabsf(bs.wi.z) / bs.pdf;
                    const V3 dir = fr.toWorld(bs.wi);
                    const V3 c = mul(thr, bs.f * weight * selInv);
                    mr = MisRec{hp.x, hp.y, hp.z, bs.pdf, dir.x, dir.y, dir.z, sel, c.x, c.y, c.z, path};
                    pushMis = true;
                }

                smp.next2(d9, u1, u2);  // continuation, :164-190
                const BsdfSample ss = bsdfSample(mat, wo, u1, u2);
                if (!(ss.pdf > 0.0f && valid(ss.f))) {
                    pushDone = true;
                } else {
                    thr = mul(thr, ss.f * (fabsf(ss.wi.z) / ss.pdf));
                    bool alive = true;
                    if (meta.y > rc.minRR) {
                        const float qrr = fmaxf(0.05f, 1.0f - maxc(thr));
                        if (smp.next() < qrr) alive = false;
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
                meta.x = smp.dim;
            }
            ps.L[path] = make_float4(L.x, L.y, L.z, 0.0f);
            ps.meta[path] = meta;
        }
        const uint32_t se = wavePush(&q.counts[nextSlot], pushExt);
        if (pushExt) qnext[se] = path;
        const uint32_t ss = wavePush(&q.counts[kCntShadow], pushShadow);
        if (pushShadow) q.shadow[ss] = sr;
        const uint32_t sm = wavePush(&q.counts[kCntMis], pushMis);
        if (pushMis) q.mis[sm] = mr;
        const uint32_t sd = wavePush(&q.counts[kCntDone], pushDone);
        if (pushDone) q.done[sd] = path;
        if (totals) {
            const uint64_t me = __ballot(pushExt), msh = __ballot(pushShadow), mm = __ballot(pushMis);
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
            float4 L = ps.L[r.path];
            L.x += r.cr, L.y += r.cg, L.z += r.cb;
            ps.L[r.path] = L;
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
    for (uint32_t i = gtid; i < n; i += gridDim.x * 64) {
        const MisRec r = q.mis[i];
        const Ray ray{{r.ox, r.oy, r.oz}, {r.dx, r.dy, r.dz}};
        const HitRec h = closestHit<W>(s, ray, 1e-3f, 1e5f, lstack + lane, gs);
        V3 rad{0.0f, 0.0f, 0.0f};
        float lpdf = 0.0f;
        if (h.pos == kHitMiss) {
            rad = ld3(s.env);
            lpdf = envPdf(ray.d);
        } else if ((h.pos & kHitQuad) && r.light == 1u + (h.pos & ~kHitQuad)) {
            const DevQuad& qd = s.quads[h.pos & ~kHitQuad];
            const V3 qn = quadNormal(qd, ray.d);
            rad = dot(qn, ray.d) < 0.0f ? ld3(qd.radiance) : V3{0.0f, 0.0f, 0.0f};
            lpdf = h.t * h.t / (fabsf(dot(qn, ray.d)) * qd.area);
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
template <typename F>
__device__ __forceinline__ void splat(F* rgbw, uint32_t w, uint32_t h, float fx, float fy, V3 L) {
    const int32_t bx = static_cast<int32_t>(fx - 0.5f), by = static_cast<int32_t>(fy - 0.5f);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t x = static_cast<uint32_t>(bx + (k & 1)), y = static_cast<uint32_t>(by + (k >> 1));
        if (!(x < w && y < h)) continue;
        F* p = rgbw + 4 * (static_cast<size_t>(y) * w + x);
        unsafeAtomicAdd(p + 0, static_cast<F>(L.x));
        unsafeAtomicAdd(p + 1, static_cast<F>(L.y));
        unsafeAtomicAdd(p + 2, static_cast<F>(L.z));
        unsafeAtomicAdd(p + 3, static_cast<F>(1));
    }
}

__global__ void __launch_bounds__(256) kFilm(PathState ps, Queues q, RenderConsts rc, float* film32, double* film64,
                                             float* pathL) {
    const uint32_t n = q.counts[kCntDone];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t path = q.done[i];
        const V3 L = xyz(ps.L[path]);
        const float2 fp = ps.film[path];
        if (film64) splat(film64, rc.width, rc.height, fp.x, fp.y, L);
        else if (film32) splat(film32, rc.width, rc.height, fp.x, fp.y, L);
        if (pathL) {
            const uint32_t g = ps.meta[path].w;
            pathL[3 * static_cast<size_t>(g)] = L.x;
            pathL[3 * static_cast<size_t>(g) + 1] = L.y;
            pathL[3 * static_cast<size_t>(g) + 2] = L.z;
        }
    }
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
            const HitRec h = closestHit<W>(s, ray, r[6], r[7], lstack + lane, gs);
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
    if (s.width == 8)
        hipLaunchKernelGGL(kExtend<8>, grid, dim3(64), 0, st, s, ps, q.ext[curSlot], q.counts + curSlot, spill,
                           spillPerLane, tc);
    else
        hipLaunchKernelGGL(kExtend<4>, grid, dim3(64), 0, st, s, ps, q.ext[curSlot], q.counts + curSlot, spill,
                           spillPerLane, tc);
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
void launchSobol(hipStream_t st, const DevScene& s, uint32_t m, uint32_t dims, const uint32_t* pix, uint32_t n,
                 float* out, uint64_t* indexOut) {
    hipLaunchKernelGGL(kSobol, dim3(blocksFor(n, 256, 1u << 30)), dim3(256), 0, st, s, m, dims, pix, n, out, indexOut);
}

}  // namespace pg

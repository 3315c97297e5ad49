// Device-side building blocks of the gfx950 path tracer: vector math, Sobol, camera, watertight
// triangle test, light quads, wide-BVH traversal, BxDFs and lights.
//
// The whole translation unit is compiled with -ffp-contract=off: the camera, the triangle test and
// the Sobol conversion must produce the same float bits as the CPU restatement of the reference.
// Where speed matters and exactness does not (the ray/box slab test), fused multiply-adds are
// written out explicitly.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "pg_shared.hpp"

namespace pg {

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V3 mul(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float length(V3 a) { return sqrtf(dot(a, a)); }
// Normal's normalising constructor (Kernel/Transform.hpp:244-248): divide each component.
__device__ __forceinline__ V3 normalize(V3 a) {
    const float l = length(a);
    return {a.x / l, a.y / l, a.z / l};
}
// Shading arithmetic (hit normals and frames, BSDFs, light sampling, MIS weights) on the hardware
// reciprocal, square root and reciprocal square root (v_rcp_f32, v_sqrt_f32, v_rsq_f32, 1 ulp)
// instead of the correctly rounded sequences of about ten instructions each: the shade kernel is
// bound by instruction issue.
// Results move by an ulp against IEEE division, inside the per-path tolerance against the oracle.
// Camera rays, traversal and hit distances keep IEEE division and square root (compared bit for
// bit), and so do the light-quad tests that repeat the closest-hit decision.
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ V3 fnormalize(V3 a) { return a * __builtin_amdgcn_rsqf(dot(a, a)); }
__device__ __forceinline__ float maxc(V3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
// Spectrum::valid (Kernel/Protocol.hpp:92-94)
__device__ __forceinline__ bool valid(V3 a) { return a.x > 0.0f || a.y > 0.0f || a.z > 0.0f; }
__device__ __forceinline__ float comp(V3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
__device__ __forceinline__ V3 ld3(const float* p) { return {p[0], p[1], p[2]}; }
__device__ __forceinline__ V3 xyz(float4 v) { return {v.x, v.y, v.z}; }

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;

// ------------------------------------------------------------------------------------------
// Sobol: sobolSampleImpl (ScrambledSobolSampler/Kernel.cpp:23-29) walking only the set bits.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ float sobolSample(const uint32_t* __restrict__ m, uint64_t index, uint32_t dim) {
    if (dim >= kSobolDims) {
        // Past the table the reference switches to a per-path pseudo-random engine
        // (Tracer/Embree/Kernel.cpp:37-45); here a splitmix64 hash of (index, dim).
        uint64_t z = index ^ (static_cast<uint64_t>(dim) * 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return static_cast<float>(static_cast<uint32_t>(z >> 40)) * 0x1p-24f;
    }
    const uint32_t* col = m + dim * kSobolCols;
    uint32_t v = 0;
    while (index) {
        v ^= col[__builtin_ctzll(index)];
        index &= index - 1;
    }
    return static_cast<float>(v) * 0x1p-32f;
}

// Dimensions d0 .. d0+7 of one index at once, from the bit-major copy of the matrices
// (sobolT[bit * kSobolDims + dim]): each set index bit costs 8 independent loads from one 32-byte
// row segment instead of 8 separate dependent chains. Bit-identical to sobolSample.
__device__ __forceinline__ void sobolWindow(const DevScene& s, uint64_t index, uint32_t d0, float out[8]) {
    if (d0 + 8 <= kByteDims) {
        // Byte-sliced table: one row of 8 dimensions per index byte, all loads independent. The
        // low four bytes are read whatever their value (a zero byte selects an all-zero row), so
        // their loads go out together; higher bytes, set from 2^32 on, loop.
        uint32_t v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const uint32_t* __restrict__ t = s.sobolB + d0;
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            const uint32_t* row = t + (k * 256 + static_cast<uint32_t>((index >> (8 * k)) & 255u)) * kByteDims;
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] ^= row[j];
        }
        index >>= 32;
        for (uint32_t k = 4; index; ++k, index >>= 8) {
            const uint32_t* row = t + (k * 256 + static_cast<uint32_t>(index & 255u)) * kByteDims;
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] ^= row[j];
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) out[j] = static_cast<float>(v[j]) * 0x1p-32f;
        return;
    }
    if (d0 + 8 > kSobolDims) {
#pragma unroll
        for (int k = 0; k < 8; ++k) out[k] = sobolSample(s.sobol, index, d0 + k);
        return;
    }
    uint32_t v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const uint32_t* __restrict__ t = s.sobolT + d0;
    while (index) {
        const uint32_t* row = t + __builtin_ctzll(index) * kSobolDims;
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] ^= row[k];
        index &= index - 1;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) out[k] = static_cast<float>(v[k]) * 0x1p-32f;
}

// sobolIntervalToIndex (ScrambledSobolSampler/Kernel.cpp:31-50) with tables from a GF(2) inverse.
// sampleBits bounds the bit length of every sample number the caller passes.
__device__ __forceinline__ uint64_t intervalToIndex(const DevScene& s, uint32_t m, uint32_t sample, uint32_t px,
                                                    uint32_t py, uint32_t sampleBits) {
    if (m == 0) return 0;
    const uint32_t n = 2 * m;
    uint64_t index = static_cast<uint64_t>(sample) << n;
    uint64_t b = (static_cast<uint64_t>(px) << m) | py;
    // The tables are read through the constant address space: with m and sampleBits the same on
    // all lanes (one render's constants) every address is wave-uniform, so the entries come in as
    // scalar loads, eight a batch, and each lane keeps those its bits select. A batch reads entries
    // up to the next multiple of 8 past a bound; they exist (rows of 64 and 32 entries, sample <
    // 2^32, n <= 2 kMaxLog2Res = 32) and are masked off by zero bits.
    using CU64 = const __attribute__((address_space(4))) uint64_t*;
    const CU64 delta = (CU64)(s.sobolDelta + m * 64);
    const CU64 inv = (CU64)(s.sobolInv + m * 32);
#pragma unroll
    for (uint32_t j0 = 0; j0 < 32; j0 += 8) {
        if (j0 >= sampleBits) break;
        uint64_t t[8];
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) t[k] = delta[j0 + k];
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) b ^= t[k] & (0ull - ((sample >> (j0 + k)) & 1u));
    }
#pragma unroll
    for (uint32_t j0 = 0; j0 < 32; j0 += 8) {
        if (j0 >= n) break;
        uint64_t t[8];
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) t[k] = inv[j0 + k];
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) index ^= t[k] & (0ull - ((b >> (j0 + k)) & 1u));
    }
    return index;
}

struct Ray {
    V3 o, d;
};

// rayGen (Sensor/PerspectiveCamera/Kernel.cpp:22-48), same operation order as the CPU oracle.
__device__ __forceinline__ Ray cameraRay(const DevCamera& c, float ndcX, float ndcY, float u1, float u2) {
    const V3 base = ld3(c.base), lookAt = ld3(c.lookAt), upRef = ld3(c.up);
    const V3 forward = normalize(lookAt - base);
    const V3 right = cross(forward, upRef);
    const V3 up = cross(right, forward);
    const V3 filmHit = base + right * (c.sizeX * (0.5f - ndcX)) + up * (c.sizeY * (ndcY - 0.5f));
    const float focalDistance = dot(lookAt - base, forward);
    const float filmDistance = 1.0f / (1.0f / c.focal - 1.0f / focalDistance);
    const V3 lensCenter = base + forward * filmDistance;
    const float ang = u1 * (2.0f * kPi);  // sampleUniformDisk, Kernel/Sampling.hpp:23-29
    const float rad = sqrtf(fmaxf(0.0f, u2));
    const float lx = rad * cosf(ang), ly = rad * sinf(ang);
    const V3 lensHit = lensCenter + right * (c.aperture * lx) + up * (c.aperture * ly);
    const V3 dir = lensCenter - filmHit;
    const V3 focusHit = lensCenter + dir * (focalDistance / dot(dir, forward));
    return Ray{lensHit, normalize(focusHit - lensHit)};
}

// ------------------------------------------------------------------------------------------
// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013), replacing Embree's triangle
// intersector behind rtcIntersect1 (Tracer/Embree/Embree.cpp:300).
// ------------------------------------------------------------------------------------------
struct RayPre {
    int kx, ky, kz;
    float sx, sy, sz;
    float ix, iy, iz;     // 1/d with zero components replaced (box test only)
    float oix, oiy, oiz;  // o * (1/d)
};
__device__ __forceinline__ RayPre precompute(const Ray& r) {
    RayPre p;
    const float ax = fabsf(r.d.x), ay = fabsf(r.d.y), az = fabsf(r.d.z);
    p.kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    p.kx = p.kz == 2 ? 0 : p.kz + 1;
    p.ky = p.kx == 2 ? 0 : p.kx + 1;
    const float dz = comp(r.d, p.kz);
    if (dz < 0.0f) {
        const int t = p.kx;
        p.kx = p.ky;
        p.ky = t;
    }
    p.sx = comp(r.d, p.kx) / dz;
    p.sy = comp(r.d, p.ky) / dz;
    p.sz = 1.0f / dz;
    p.ix = 1.0f / (r.d.x == 0.0f ? 1e-20f : r.d.x);
    p.iy = 1.0f / (r.d.y == 0.0f ? 1e-20f : r.d.y);
    p.iz = 1.0f / (r.d.z == 0.0f ? 1e-20f : r.d.z);
    p.oix = r.o.x * p.ix;
    p.oiy = r.o.y * p.iy;
    p.oiz = r.o.z * p.iz;
    return p;
}

__device__ __forceinline__ bool intersectTri(const Ray& r, const RayPre& p, V3 v0, V3 v1, V3 v2, float& t, float& bu,
                                             float& bv) {
    const V3 a = v0 - r.o, b = v1 - r.o, c = v2 - r.o;
    const float az = comp(a, p.kz), bz = comp(b, p.kz), cz = comp(c, p.kz);
    const float ax = comp(a, p.kx) - p.sx * az, ay = comp(a, p.ky) - p.sy * az;
    const float bx = comp(b, p.kx) - p.sx * bz, by = comp(b, p.ky) - p.sy * bz;
    const float cx = comp(c, p.kx) - p.sx * cz, cy = comp(c, p.ky) - p.sy * cz;
    float U = cx * by - cy * bx;
    float V = ax * cy - ay * cx;
    float W = bx * ay - by * ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {  // exact edge: redo the 2D edge tests in double
        U = static_cast<float>(static_cast<double>(cx) * by - static_cast<double>(cy) * bx);
        V = static_cast<float>(static_cast<double>(ax) * cy - static_cast<double>(ay) * cx);
        W = static_cast<float>(static_cast<double>(bx) * ay - static_cast<double>(by) * ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float T = U * (p.sz * az) + V * (p.sz * bz) + W * (p.sz * cz);
    t = T / det;
    bu = V / det;
    bv = W / det;
    return true;
}

// planeIntersect + calcTexCoord (Geometry/BasicShapes/Kernel.cpp:30-74)
__device__ __forceinline__ bool intersectQuad(const Ray& r, const DevQuad& q, float tNear, float tFar, float& tOut) {
    const V3 qo = ld3(q.origin), qu = ld3(q.u), qv = ld3(q.v), qn = ld3(q.normal);
    const V3 delta = qo - r.o;
    const float b = dot(delta, qn);
    const float k = dot(r.d, qn);
    // Division-free rejection that never rejects a hit of the exact test below (tNear >= 0):
    // t > 0 needs b and k of one sign; |b| >= |k| tFar (1 + 1e-6) guarantees fl(b / k) >= tFar.
    if (!((b > 0.0f && k > 0.0f) || (b < 0.0f && k < 0.0f))) return false;
    if (fabsf(b) >= fabsf(k) * (tFar * 1.000001f)) return false;
    const float t = b / k;
    if (!(tNear < t && t < tFar)) return false;
    const V3 pos = r.d * t - delta;
    float invDet, x, y;
    if (q.maxDetComp == 0) {
        invDet = 1.0f / (qu.y * qv.z - qv.y * qu.z);
        x = pos.y * qv.z - qv.y * pos.z;
        y = qu.y * pos.z - pos.y * qu.z;
    } else if (q.maxDetComp == 1) {
        invDet = 1.0f / (qu.z * qv.x - qv.z * qu.x);
        x = pos.z * qv.x - qv.z * pos.x;
        y = qu.z * pos.x - pos.z * qu.x;
    } else {
        invDet = 1.0f / (qu.x * qv.y - qv.x * qu.y);
        x = pos.x * qv.y - qv.x * pos.y;
        y = qu.x * pos.y - pos.x * qu.y;
    }
    const float u = x * invDet, v = y * invDet;
    if (!(0.0f <= u && u <= 1.0f && 0.0f <= v && v <= 1.0f)) return false;
    tOut = t;
    return true;
}

// ------------------------------------------------------------------------------------------
// Wide-BVH traversal, one ray per lane. Each lane keeps a stack of (child reference, entry
// distance) pairs: the first entries (a launch setting, kLdsStack by default) live in the wave's
// LDS slice (lane-minor, so the 64 lanes of a wave hit distinct banks), deeper entries spill to a
// per-lane HBM region.
// Closest hit: nearest child first, the rest pushed and culled on pop against the best t.
// ------------------------------------------------------------------------------------------


struct HitRec {
    float t;
    uint32_t pos;   // leaf-order triangle position, kHitQuad | quad, or kHitMiss
    uint32_t prim;  // original triangle index (tie rule), or as pos for quads/miss
    float u, v;
};

// A lane's traversal stack: `depth` entries in the wave's LDS slice (entry k at ls[64 k]), the
// deeper ones in the lane's spill region gs.
struct TravStack {
    uint2* ls;
    uint2* gs;
    int depth;
};
__device__ __forceinline__ void stackPush(const TravStack& st, int& sp, uint32_t ref, float t) {
    const uint2 e = make_uint2(ref, __float_as_uint(t));
    if (sp < st.depth) st.ls[sp * 64] = e;
    else st.gs[sp - st.depth] = e;
    ++sp;
}
__device__ __forceinline__ uint2 stackPop(const TravStack& st, int& sp) {
    --sp;
    return sp < st.depth ? st.ls[sp * 64] : st.gs[sp - st.depth];
}

// ANY = false: closest hit of triangles in (tmin, h.t] (tie: smaller original prim).
// ANY = true : returns true at the first triangle with tmin < t < tmax.
struct TravCount {
    uint32_t nodes = 0, tris = 0;
};

// COUNT = true also counts wide nodes visited and triangles tested (roofline bookkeeping).
template <int W, bool ANY, bool COUNT = false>
__device__ __forceinline__ bool traverse(const DevScene& s, const Ray& r, const RayPre& p, float tmin, float tmax,
                                         HitRec& h, const TravStack& stk, TravCount& cnt, uint32_t root = 0) {
    const WideNode<W>* __restrict__ nodes = static_cast<const WideNode<W>*>(s.nodes);
    const float4* __restrict__ tris = s.tris;
    int sp = 0;
    uint32_t ref = root;  // node 0: the static triangles' root; an instance's BLAS root otherwise
    while (true) {
        if (!(ref & kLeafBit)) {
            if (COUNT) ++cnt.nodes;
            const WideNode<W> n = nodes[ref];
            const float tcull = ANY ? tmax : h.t * 1.00001f;
            float tn[W];
            uint32_t mask = 0;
#pragma unroll
            for (int c = 0; c < W; ++c) {
                const float ax = __builtin_fmaf(n.lox[c], p.ix, -p.oix), bx = __builtin_fmaf(n.hix[c], p.ix, -p.oix);
                const float ay = __builtin_fmaf(n.loy[c], p.iy, -p.oiy), by = __builtin_fmaf(n.hiy[c], p.iy, -p.oiy);
                const float az = __builtin_fmaf(n.loz[c], p.iz, -p.oiz), bz = __builtin_fmaf(n.hiz[c], p.iz, -p.oiz);
                const float t0 = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), tmin));
                const float t1 = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), tcull)) * 1.0000004f;
                tn[c] = t0;
                if (n.child[c] != kInvalidChild && t0 <= t1) mask |= 1u << c;
            }
            if (mask) {
                // Every index into the node is a compile-time constant (no dynamic indexing), so
                // the node stays in registers instead of being spilled to scratch memory.
                int best = 0;
                float bt = INFINITY;
                uint32_t bestRef = 0;
#pragma unroll
                for (int c = 0; c < W; ++c)
                    if (((mask >> c) & 1) && tn[c] < bt) bt = tn[c], best = c, bestRef = n.child[c];
#pragma unroll
                for (int c = 0; c < W; ++c)
                    if (((mask >> c) & 1) && c != best) stackPush(stk, sp, n.child[c], tn[c]);
                ref = bestRef;
                continue;
            }
        } else {
            const uint32_t first = ref & kLeafFirstMask;
            const uint32_t count = ((ref >> kLeafCountShift) & 0xFu) + 1u;
            if constexpr (ANY) {
                // Occlusion: the leaf's triangles two at a time, both loads issued before either
                // test (a lone last triangle is loaded twice, in bounds), one memory wait per pair.
                // The shadow stage takes 15.7 instead of 16.9 ms per 1080p64 frame this way; the
                // closest-hit kernel is slower with it (1.656 against 1.622 ms per launch).
                const uint32_t end = first + count;
                for (uint32_t k = first; k < end; k += 2) {
                    const bool two = k + 1 < end;
                    const uint32_t k1 = two ? k + 1 : k;
                    const float4 a0 = tris[3 * k], b0 = tris[3 * k + 1], c0 = tris[3 * k + 2];
                    const float4 a1 = tris[3 * k1], b1 = tris[3 * k1 + 1], c1 = tris[3 * k1 + 2];
                    __builtin_amdgcn_sched_barrier(0);  // keep the six loads ahead of their first use
                    if (COUNT) cnt.tris += two ? 2u : 1u;
                    float t, u, v;
                    if (intersectTri(r, p, xyz(a0), xyz(b0), xyz(c0), t, u, v) && t > tmin && t < tmax) return true;
                    if (two && intersectTri(r, p, xyz(a1), xyz(b1), xyz(c1), t, u, v) && t > tmin && t < tmax)
                        return true;
                }
            } else {
                for (uint32_t k = first; k < first + count; ++k) {
                    if (COUNT) ++cnt.tris;
                    const float4 a = tris[3 * k], b = tris[3 * k + 1], c = tris[3 * k + 2];
                    float t, u, v;
                    if (!intersectTri(r, p, xyz(a), xyz(b), xyz(c), t, u, v)) continue;
                    const uint32_t prim = __float_as_uint(a.w);
                    if (t > tmin && (t < h.t || (t == h.t && prim < h.prim))) h = HitRec{t, k, prim, u, v};
                }
            }
        }
        // pop the next entry that can still hold a closer hit
        bool found = false;
        while (sp > 0) {
            const uint2 e = stackPop(stk, sp);
            if (ANY || __uint_as_float(e.y) <= h.t * 1.00001f) {
                ref = e.x;
                found = true;
                break;
            }
        }
        if (!found) break;
    }
    return ANY ? false : h.pos != kHitMiss;
}

// ------------------------------------------------------------------------------------------
// Moving instances. Local-to-world transform at a time as Embree 3.12 forms it from the keys the
// reference sets with rtcSetGeometryTransformQuaternion (Embree.cpp:699-703, 756-772): scale,
// skew, shift and translation interpolate linearly between the two keys around the time, the
// rotation by spherical linear interpolation along the shorter arc (Shoemake 1985), normalised
// linear interpolation once the two quaternions' cosine exceeds 0.9995; the matrix is
// T R S with S = [[sx, kxy, kxz, hx], [0, sy, kyz, hy], [0, 0, sz, hz]] (Embree's quaternion
// decomposition), then the parents' static transform.
// ------------------------------------------------------------------------------------------
struct Affine {
    float m[12];  // row-major 3x4
};
__device__ __forceinline__ V3 affPoint(const Affine& a, V3 p) {
    return {a.m[0] * p.x + a.m[1] * p.y + a.m[2] * p.z + a.m[3], a.m[4] * p.x + a.m[5] * p.y + a.m[6] * p.z + a.m[7],
            a.m[8] * p.x + a.m[9] * p.y + a.m[10] * p.z + a.m[11]};
}
__device__ __forceinline__ V3 affVector(const Affine& a, V3 v) {
    return {a.m[0] * v.x + a.m[1] * v.y + a.m[2] * v.z, a.m[4] * v.x + a.m[5] * v.y + a.m[6] * v.z,
            a.m[8] * v.x + a.m[9] * v.y + a.m[10] * v.z};
}
__device__ __forceinline__ Affine affCompose(const float* a, const Affine& b) {  // a * b
    Affine r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            r.m[4 * i + j] = a[4 * i] * b.m[j] + a[4 * i + 1] * b.m[4 + j] + a[4 * i + 2] * b.m[8 + j] + (j == 3 ? a[4 * i + 3] : 0.0f);
    return r;
}
__device__ __forceinline__ Affine affInverse(const Affine& a) {
    const float* m = a.m;
    const float c00 = m[5] * m[10] - m[6] * m[9], c01 = m[6] * m[8] - m[4] * m[10], c02 = m[4] * m[9] - m[5] * m[8];
    const float inv = 1.0f / (m[0] * c00 + m[1] * c01 + m[2] * c02);
    Affine r;
    r.m[0] = c00 * inv, r.m[1] = (m[2] * m[9] - m[1] * m[10]) * inv, r.m[2] = (m[1] * m[6] - m[2] * m[5]) * inv;
    r.m[4] = c01 * inv, r.m[5] = (m[0] * m[10] - m[2] * m[8]) * inv, r.m[6] = (m[2] * m[4] - m[0] * m[6]) * inv;
    r.m[8] = c02 * inv, r.m[9] = (m[1] * m[8] - m[0] * m[9]) * inv, r.m[10] = (m[0] * m[5] - m[1] * m[4]) * inv;
#pragma unroll
    for (int i = 0; i < 3; ++i) r.m[4 * i + 3] = -(r.m[4 * i] * m[3] + r.m[4 * i + 1] * m[7] + r.m[4 * i + 2] * m[11]);
    return r;
}
// Unit quaternions (w, x, y, z) q0, q1 at fraction f.
__device__ __forceinline__ void quatSlerp(const float* q0, const float* q1in, float f, float q[4]) {
    float q1[4] = {q1in[0], q1in[1], q1in[2], q1in[3]};
    float c = q0[0] * q1[0] + q0[1] * q1[1] + q0[2] * q1[2] + q0[3] * q1[3];
    if (c < 0.0f) {
        c = -c;
#pragma unroll
        for (int i = 0; i < 4; ++i) q1[i] = -q1[i];
    }
    if (c > 0.9995f) {
        float n = 0.0f;
#pragma unroll
        for (int i = 0; i < 4; ++i) q[i] = q0[i] * (1.0f - f) + q1[i] * f, n += q[i] * q[i];
        const float inv = 1.0f / sqrtf(n);
#pragma unroll
        for (int i = 0; i < 4; ++i) q[i] *= inv;
        return;
    }
    float p[4], n = 0.0f;  // the part of q1 orthogonal to q0
#pragma unroll
    for (int i = 0; i < 4; ++i) p[i] = q1[i] - c * q0[i], n += p[i] * p[i];
    const float inv = 1.0f / sqrtf(n), phi = f * acosf(c);
    const float cs = cosf(phi), sn = sinf(phi) * inv;
#pragma unroll
    for (int i = 0; i < 4; ++i) q[i] = cs * q0[i] + sn * p[i];
}
__device__ __forceinline__ Affine instanceAt(const DevInstance& I, float time) {
    uint32_t seg = 0;
    float f = 0.0f;
    if (I.keys > 1) {
        const float x = (time - I.t0) * I.rate;
        const float fl = fminf(fmaxf(floorf(x), 0.0f), static_cast<float>(I.keys - 2));
        seg = static_cast<uint32_t>(fl);
        f = x - fl;
    }
    const float* a = I.srt[seg];
    const float* b = I.srt[I.keys > 1 ? seg + 1 : seg];
    float k[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) k[i] = a[i] * (1.0f - f) + b[i] * f;
    float q[4];
    quatSlerp(a + 9, b + 9, f, q);
    const float w = q[0], x = q[1], y = q[2], z = q[3];
    const float r[9] = {1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y - w * z), 2.0f * (x * z + w * y),
                        2.0f * (x * y + w * z), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z - w * x),
                        2.0f * (x * z - w * y), 2.0f * (y * z + w * x), 1.0f - 2.0f * (x * x + y * y)};
    Affine m;  // R S, then the translation
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float r0 = r[3 * i], r1 = r[3 * i + 1], r2 = r[3 * i + 2];
        m.m[4 * i] = r0 * k[0];
        m.m[4 * i + 1] = r0 * k[3] + r1 * k[1];
        m.m[4 * i + 2] = r0 * k[4] + r1 * k[5] + r2 * k[2];
        m.m[4 * i + 3] = r0 * k[6] + r1 * k[7] + r2 * k[8] + k[13 + i];
    }
    return affCompose(I.parent, m);
}
__device__ __forceinline__ bool instanceVisible(const DevInstance& I, float time) {
    return I.visible && time >= I.t0 && time <= I.t1;
}
// Slab test of a world box (the instance's swept bounds) against (tmin, tmax].
__device__ __forceinline__ bool boxHit(const float* lo, const float* hi, const RayPre& p, float tmin, float tmax) {
    const float ax = lo[0] * p.ix - p.oix, bx = hi[0] * p.ix - p.oix;
    const float ay = lo[1] * p.iy - p.oiy, by = hi[1] * p.iy - p.oiy;
    const float az = lo[2] * p.iz - p.oiz, bz = hi[2] * p.iz - p.oiz;
    const float t0 = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), tmin));
    const float t1 = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), tmax)) * 1.0000004f;
    return t0 <= t1;
}
// The instances after the static BVH (an Embree instance node over each moving mesh): the ray goes
// into the instance's local space at its time (same parameter t, the direction is not
// normalised) and through the instance's BLAS.
template <int W, bool ANY, bool COUNT = false>
__device__ __forceinline__ bool traverseInstances(const DevScene& s, const Ray& r, const RayPre& p, float time,
                                                  float tmin, float tmax, HitRec& h, const TravStack& stk,
                                                  TravCount& cnt) {
    for (uint32_t k = 0; k < s.numInst; ++k) {
        const DevInstance& I = s.inst[k];
        if (!instanceVisible(I, time) || !boxHit(I.lo, I.hi, p, tmin, ANY ? tmax : h.t * 1.00001f)) continue;
        const Affine inv = affInverse(instanceAt(I, time));
        const Ray lr{affPoint(inv, r.o), affVector(inv, r.d)};
        const RayPre lp = precompute(lr);
        if (traverse<W, ANY, COUNT>(s, lr, lp, tmin, tmax, h, stk, cnt, I.root) && ANY) return true;
    }
    return false;
}

// Closest hit of triangles and light quads (rtcIntersect1 with mask gsm|areaLight, Embree.cpp:292)
// at ray time `time` (the sample's dimension 2, Embree/Kernel.cpp:77; only moving instances use it).
// MOTION = false: the kernel variant for scenes without moving instances, which carries none of
// their code (and so none of its registers).
template <int W, bool COUNT = false, bool MOTION = false>
__device__ __forceinline__ HitRec closestHit(const DevScene& s, const Ray& r, float tmin, float tmax,
                                             const TravStack& stk, TravCount& cnt, float time = 0.0f) {
    HitRec h{tmax, kHitMiss, kHitMiss, 0.0f, 0.0f};
    const RayPre p = precompute(r);
    if (s.hasBvh) traverse<W, false, COUNT>(s, r, p, tmin, tmax, h, stk, cnt);
    if (MOTION && s.numInst) traverseInstances<W, false, COUNT>(s, r, p, time, tmin, tmax, h, stk, cnt);
    for (uint32_t q = 0; q < s.numQuads; ++q) {
        float t;
        if (intersectQuad(r, s.quads[q], tmin, h.t, t)) h = HitRec{t, kHitQuad | q, kHitQuad | q, 0.0f, 0.0f};
    }
    if (h.pos == kHitMiss) h.t = tmax;
    return h;
}

// rtcOccluded1 with mask gsm (Embree.cpp:357-391): light quads never occlude.
template <int W, bool MOTION = false>
__device__ __forceinline__ bool occluded(const DevScene& s, const Ray& r, float tmin, float tmax, const TravStack& stk,
                                         float time = 0.0f) {
    if (!s.hasBvh && !(MOTION && s.numInst)) return false;
    HitRec h{tmax, kHitMiss, kHitMiss, 0.0f, 0.0f};
    const RayPre p = precompute(r);
    TravCount cnt;
    if (s.hasBvh && traverse<W, true>(s, r, p, tmin, tmax, h, stk, cnt)) return true;
    return MOTION && s.numInst && traverseInstances<W, true>(s, r, p, time, tmin, tmax, h, stk, cnt);
}

// ------------------------------------------------------------------------------------------
// Shading frame and hit normal (calcTriangleMeshSurface, Tracer/Embree/Kernel.cpp:170-192;
// planeSurface, BasicShapes/Kernel.cpp:99-111; SurfaceIntersectionInfo, Protocol.hpp:47-66)
// ------------------------------------------------------------------------------------------
struct Frame {
    V3 t, b, n;
    __device__ __forceinline__ V3 toLocal(V3 v) const { return {dot(t, v), dot(b, v), dot(n, v)}; }
    __device__ __forceinline__ V3 toWorld(V3 v) const {
        return {t.x * v.x + b.x * v.y + n.x * v.z, t.y * v.x + b.y * v.y + n.y * v.z, t.z * v.x + b.z * v.y + n.z * v.z};
    }
};
__device__ __forceinline__ Frame shadingFrame(V3 n, bool d3) {
    Frame f;
    f.n = n;
    const V3 ex{1.0f, 0.0f, 0.0f}, ey{0.0f, 1.0f, 0.0f};
    f.t = fabsf(dot(n, ex)) < fabsf(dot(n, ey)) ? cross(n, ex) : cross(n, ey);
    if (!d3) f.t = fnormalize(f.t);
    f.b = cross(n, f.t);
    return f;
}
// A moving instance's triangle in world space at `time` (static triangles are stored there).
template <bool MOTION>
__device__ __forceinline__ void worldTriangle(const DevInstance* inst, float4 a, float4 b, float4 c, float time, V3& v0,
                                              V3& v1, V3& v2) {
    v0 = xyz(a), v1 = xyz(b), v2 = xyz(c);
    const uint32_t k = __float_as_uint(b.w);
    if (MOTION && k) {
        const Affine m = instanceAt(inst[k - 1], time);
        v0 = affPoint(m, v0), v1 = affPoint(m, v1), v2 = affPoint(m, v2);
    }
}
template <bool MOTION = false>
__device__ __forceinline__ V3 triNormal(const DevScene& s, uint32_t pos, V3 dir, bool* front = nullptr,
                                        float time = 0.0f) {
    V3 v0, v1, v2;
    worldTriangle<MOTION>(s.inst, s.tris[3 * pos], s.tris[3 * pos + 1], s.tris[3 * pos + 2], time, v0, v1, v2);
    const V3 ng = fnormalize(cross(v1 - v0, v2 - v0));
    const bool f = dot(dir, ng) < 0.0f;  // Embree.cpp:336: Front when dot(dir, Ng) < 0
    if (front) *front = f;
    return f ? ng : -ng;
}
__device__ __forceinline__ V3 quadNormal(const DevQuad& q, V3 dir) {
    const V3 n = ld3(q.normal);
    return dot(dir, n) > 0.0f ? -n : n;
}

// ------------------------------------------------------------------------------------------
// Sampling helpers (Kernel/Sampling.hpp:31-50) and BxDFs (Surface/Classic/Kernel.cpp)
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ V3 cosineHemisphere(float u1, float u2) {
    u1 = 2.0f * u1 - 1.0f;
    u2 = 2.0f * u2 - 1.0f;
    const float au1 = fabsf(u1), au2 = fabsf(u2);
    float x = 0.0f, y = 0.0f;
    if (!(fminf(au1, au2) < 1e-4f)) {
        const float rad = au1 > au2 ? u1 : u2;
        const float theta = au1 > au2 ? (kPi / 4.0f) * fdiv(u2, u1) : (kPi / 2.0f) - (kPi / 4.0f) * fdiv(u1, u2);
        x = cosf(theta) * rad;
        y = sinf(theta) * rad;
    }
    return {x, y, fsqrt(fmaxf(0.0f, 1.0f - (x * x + y * y)))};
}

__device__ __forceinline__ float ggxLambda(V3 w, float alpha) {  // Classic/Kernel.cpp:326-341
    const float cos2 = w.z * w.z;
    const float sin2 = fmaxf(0.0f, 1.0f - cos2);
    const float tan2 = fdiv(sin2, cos2);
    if (isinf(tan2)) return 0.0f;
    return (-1.0f + fsqrt(1.0f + alpha * alpha * tan2)) * 0.5f;
}
__device__ __forceinline__ float ggxD(V3 wh, float alpha) {  // Classic/Kernel.cpp:342-357
    const float cos2 = wh.z * wh.z;
    const float sin2 = fmaxf(0.0f, 1.0f - cos2);
    const float tan2 = fdiv(sin2, cos2);
    if (isinf(tan2)) return 0.0f;
    const float e = cos2 * (1.0f + fdiv(tan2, alpha * alpha));
    return fdiv(kInvPi, alpha * alpha * e * e);
}
// Visible-normal sampling (Heitz, JCGT 2018).
__device__ __forceinline__ V3 ggxSampleWh(V3 wo, float alpha, float u1, float u2) {
    const V3 vh = fnormalize(V3{alpha * wo.x, alpha * wo.y, wo.z});
    const float lensq = vh.x * vh.x + vh.y * vh.y;
    const V3 t1 = lensq > 0.0f ? V3{-vh.y, vh.x, 0.0f} * __builtin_amdgcn_rsqf(lensq) : V3{1.0f, 0.0f, 0.0f};
    const V3 t2 = cross(vh, t1);
    const float r = fsqrt(u1);
    const float phi = 2.0f * kPi * u2;
    const float p1 = r * cosf(phi);
    float p2 = r * sinf(phi);
    const float sm = 0.5f * (1.0f + vh.z);
    p2 = (1.0f - sm) * fsqrt(fmaxf(0.0f, 1.0f - p1 * p1)) + sm * p2;
    const V3 nh = t1 * p1 + t2 * p2 + vh * fsqrt(fmaxf(0.0f, 1.0f - p1 * p1 - p2 * p2));
    return fnormalize(V3{alpha * nh.x, alpha * nh.y, fmaxf(0.0f, nh.z)});
}
// Anisotropic Trowbridge-Reitz for glass and substrate, whose RoughnessX and RoughnessY give the
// alphas along the shading frame's x and y axes (MicrofacetDistribution, Classic/Kernel.cpp:289-357,
// in its form: the azimuth's cosine and sine from w over sin(theta), clamped to [-1, 1]).
struct Ggx {
    float ax, ay;
};
__device__ __forceinline__ float ggxLambda(V3 w, Ggx g) {  // Classic/Kernel.cpp:326-341
    const float cos2 = w.z * w.z;
    const float sin2 = fmaxf(0.0f, 1.0f - cos2);
    const float tan2 = fdiv(sin2, cos2);
    if (isinf(tan2)) return 0.0f;
    const float sinT = fsqrt(sin2);
    const float cosPhi = sinT == 0.0f ? 1.0f : fminf(1.0f, fmaxf(-1.0f, fdiv(w.x, sinT)));
    const float sinPhi = sinT == 0.0f ? 0.0f : fminf(1.0f, fmaxf(-1.0f, fdiv(w.y, sinT)));
    const float p1 = cosPhi * g.ax, p2 = sinPhi * g.ay;
    return (-1.0f + fsqrt(1.0f + (p1 * p1 + p2 * p2) * tan2)) * 0.5f;
}
__device__ __forceinline__ float ggxD(V3 wh, Ggx g) {  // Classic/Kernel.cpp:342-357
    const float cos2 = wh.z * wh.z;
    const float sin2 = fmaxf(0.0f, 1.0f - cos2);
    const float tan2 = fdiv(sin2, cos2);
    if (isinf(tan2)) return 0.0f;
    const float sinT = fsqrt(sin2);
    const float cosPhi = sinT == 0.0f ? 1.0f : fminf(1.0f, fmaxf(-1.0f, fdiv(wh.x, sinT)));
    const float sinPhi = sinT == 0.0f ? 0.0f : fminf(1.0f, fmaxf(-1.0f, fdiv(wh.y, sinT)));
    const float p1 = fdiv(cosPhi, g.ax), p2 = fdiv(sinPhi, g.ay);
    const float e = cos2 * (1.0f + (p1 * p1 + p2 * p2) * tan2);
    return fdiv(kInvPi, g.ax * g.ay * e * e);
}
// Visible-normal sampling (Heitz, JCGT 2018) with the view direction stretched by (ax, ay).
__device__ __forceinline__ V3 ggxSampleWh(V3 wo, Ggx g, float u1, float u2) {
    const V3 vh = fnormalize(V3{g.ax * wo.x, g.ay * wo.y, wo.z});
    const float lensq = vh.x * vh.x + vh.y * vh.y;
    const V3 t1 = lensq > 0.0f ? V3{-vh.y, vh.x, 0.0f} * __builtin_amdgcn_rsqf(lensq) : V3{1.0f, 0.0f, 0.0f};
    const V3 t2 = cross(vh, t1);
    const float r = fsqrt(u1);
    const float phi = 2.0f * kPi * u2;
    const float p1 = r * cosf(phi);
    float p2 = r * sinf(phi);
    const float sm = 0.5f * (1.0f + vh.z);
    p2 = (1.0f - sm) * fsqrt(fmaxf(0.0f, 1.0f - p1 * p1)) + sm * p2;
    const V3 nh = t1 * p1 + t2 * p2 + vh * fsqrt(fmaxf(0.0f, 1.0f - p1 * p1 - p2 * p2));
    return fnormalize(V3{g.ax * nh.x, g.ay * nh.y, fmaxf(0.0f, nh.z)});
}
// Density of a reflection drawn from the visible normals: D(wh) G1(wo) / (4 cos theta_o).
__device__ __forceinline__ float ggxReflectPdf(Ggx g, V3 wo, V3 wi) {
    if (wi.z <= 0.0f || wo.z <= 0.0f) return 0.0f;
    const V3 wh = fnormalize(wo + wi);
    return fdiv(ggxD(wh, g), 4.0f * wo.z * (1.0f + ggxLambda(wo, g)));
}
__device__ __forceinline__ float remapRoughness(float roughness) {  // Classic/Kernel.cpp:378-381
    const float x = logf(fmaxf(1e-3f, roughness));
    return fmaxf(1e-3f, 1.62142f + (0.819955f + (0.1734f + (0.0171201f + 0.000640711f * x) * x) * x) * x);
}

__device__ __forceinline__ V3 schlick(V3 f0, float c) {
    const float m = 1.0f - fminf(1.0f, fmaxf(0.0f, c));
    const float m5 = (m * m) * (m * m) * m;
    return V3{f0.x + (1.0f - f0.x) * m5, f0.y + (1.0f - f0.y) * m5, f0.z + (1.0f - f0.z) * m5};
}
__device__ __forceinline__ V3 conductorF(V3 f0, float alpha, V3 wo, V3 wi) {  // MicrofacetReflection::evaluate
    if (wo.z <= 0.0f || wi.z <= 0.0f) return {0.0f, 0.0f, 0.0f};
    const V3 wh = fnormalize(wo + wi);
    const float g = 1.0f + ggxLambda(wo, alpha) + ggxLambda(wi, alpha);
    const float k = fdiv(ggxD(wh, alpha), 4.0f * g * fabsf(wo.z * wi.z));
    return schlick(f0, dot(wi, wh)) * k;
}
__device__ __forceinline__ float conductorPdf(float alpha, V3 wo, V3 wi) {
    if (wi.z <= 0.0f || wo.z <= 0.0f) return 0.0f;
    const V3 wh = fnormalize(wo + wi);
    return fdiv(ggxD(wh, alpha), 4.0f * wo.z * (1.0f + ggxLambda(wo, alpha)));
}

// FresnelDielectric (Classic/Kernel.cpp:271-286).
__device__ __forceinline__ float fresnelDielectric(float cosI, float etaI, float etaT) {
    cosI = fmaxf(0.0f, fminf(1.0f, fabsf(cosI)));
    const float sinI = fsqrt(fmaxf(0.0f, 1.0f - cosI * cosI));
    const float sinT = fdiv(etaI, etaT) * sinI;
    if (sinT >= 1.0f) return 1.0f;
    const float cosT = fsqrt(fmaxf(0.0f, 1.0f - sinT * sinT));
    const float p1 = etaT * cosI, p2 = etaI * cosT;
    const float par = fdiv(p1 - p2, p1 + p2);
    const float p3 = etaI * cosI, p4 = etaT * cosT;
    const float perp = fdiv(p3 - p4, p3 + p4);
    return 0.5f * (par * par + perp * perp);
}

// Textures (Texture/ProceduralTexture/Kernel.cpp:28-38, Texture/SampledTexture/Kernel.cpp:22-82).
__device__ __forceinline__ float checkerMod(float x, float scale) {
    const float rem = remainderf(x * scale, 2.0f);
    return rem < 0.0f ? rem + 2.0f : rem;
}
__device__ __forceinline__ uint32_t texLocate(float p, uint32_t size, uint32_t wrap) {
    const float fs = static_cast<float>(size);
    const float rem = remainderf(p, fs);
    const float q = wrap == kWrapRepeat ? (rem < 0.0f ? rem + fs : rem) : fabsf(rem);
    return min(static_cast<uint32_t>(q), size - 1u);
}
__device__ __forceinline__ uint32_t texNext(uint32_t p, uint32_t size, uint32_t wrap) {
    const uint32_t q = p + 1u;
    return q == size ? (wrap == kWrapRepeat ? 0u : p) : q;
}
__device__ __noinline__ float4 sampleTexture(const DevTexture* texTab, const uint8_t* texels, uint32_t t, float2 uv) {
    const DevTexture& tx = texTab[t];
    float c[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (tx.type == kTexChecker) {
        const bool white = (checkerMod(uv.x, tx.scale) > 1.0f) == (checkerMod(uv.y, tx.scale) > 1.0f);
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i) c[i] = i < tx.channels ? (white ? tx.white[i] : tx.black[i]) : 0.0f;
    } else {  // bilinear over 8-bit texels, left tap at floor(p - 0.5)
        const float pu = uv.x * static_cast<float>(tx.width), pv = uv.y * static_cast<float>(tx.height);
        const uint32_t lu = texLocate(pu - 0.5f, tx.width, tx.wrap), ru = texNext(lu, tx.width, tx.wrap);
        const uint32_t lv = texLocate(pv - 0.5f, tx.height, tx.wrap), rv = texNext(lv, tx.height, tx.wrap);
        const float lwx = 0.5f + floorf(pu + 0.5f) - pu, lwy = 0.5f + floorf(pv + 0.5f) - pv;
        const float rwx = 1.0f - lwx, rwy = 1.0f - lwy;
        const uint8_t* base = texels + tx.offset;
        const uint32_t stride = tx.width * tx.channels;
        const uint32_t taps[4][2] = {{lu, lv}, {ru, lv}, {lu, rv}, {ru, rv}};
        const float wts[4] = {lwx * lwy, rwx * lwy, lwx * rwy, rwx * rwy};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint8_t* px = base + taps[k][1] * stride + taps[k][0] * tx.channels;
#pragma unroll
            for (uint32_t i = 0; i < 4; ++i)
                if (i < tx.channels) c[i] += wts[k] * static_cast<float>(px[i]);
        }
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i) c[i] *= 0.003921568f;
    }
    return make_float4(c[0], c[1], c[2], c[3]);
}
// calcTriangleMeshSurface (Tracer/Embree/Kernel.cpp:175-188): corners 1 and 2 weighted by the
// hit's barycentrics u and v, corner 0 by 1 - u - v.
__device__ __forceinline__ float2 hitTexcoord(const DevScene& s, uint32_t pos, float bu, float bv) {
    if (!s.triUV) return make_float2(0.0f, 0.0f);
    const uint32_t prim = __float_as_uint(s.tris[3 * pos].w);
    const float* t = s.triUV + 6 * static_cast<size_t>(prim);
    const float w = 1.0f - bu - bv;
    return make_float2(t[2] * bu + t[4] * bv + t[0] * w, t[3] * bu + t[5] * bv + t[1] * w);
}

// A hit's material. The plastic and Oren-Nayar extras (spec) stay in the scene's table (or its
// texture) and are read where a lobe needs them, so they hold no registers across the vertex.
constexpr uint32_t kRoughFromTexture = 0x200u;  // Mat::bits: alpha (and Oren-Nayar's B) come from a texture
struct Mat {
    uint32_t bits;  // type (0 Lambert, 1 conductor, 2 plastic, 3 mirror, 4 glass, 5 substrate,
                    // 16 Oren-Nayar, 17 rough glass) | front face << 8 | kRoughFromTexture | kMatTextured
    V3 albedo;      // diffuse colour, conductor F0, mirror / glass reflection (texture applied)
    float alpha;    // GGX alpha (glass, substrate: along x); Oren-Nayar: A
    float aux;      // glass, substrate: GGX alpha along y; Oren-Nayar: B when the roughness comes from a texture
    uint32_t id;
    float2 uv;      // the hit's texture coordinate (textured materials)
    const DevMaterial* tab;
    const DevTexture* texTab;
    const uint8_t* texels;
    const uint4* matTex;
    __device__ __forceinline__ uint32_t type() const { return bits & 0xFFu; }
    // plastic / glass: the hit is on the triangle's front face (etaI, etaT of their init)
    __device__ __forceinline__ bool front() const { return ((bits >> 8) & 1u) != 0u; }
    // glass, substrate: the anisotropic distribution of their RoughnessX / RoughnessY
    __device__ __forceinline__ Ggx ggx() const { return Ggx{alpha, aux}; }
    // plastic / substrate specular, glass transmission; Oren-Nayar: B in x
    __device__ __forceinline__ V3 spec() const {
        if (bits & kMatTextured) {
            const uint32_t t = matTex[id].y;
            if (t != kTexNone) {
                const float4 c = sampleTexture(texTab, texels, t, uv);
                return {c.x, c.y, c.z};
            }
        }
        return ld3(tab[id].spec);
    }
};
// TEX false: the scene has no textured material, so the flag is cleared and every texture lookup
// (here and in Mat::spec) compiles away.
template <bool TEX = true>
__device__ __forceinline__ Mat loadMat(const DevScene& s, uint32_t id, bool front = true, uint32_t pos = 0,
                                       float bu = 0.0f, float bv = 0.0f) {
    const DevMaterial& m = s.mats[id];
    const uint32_t type = TEX ? m.type : (m.type & 0xFFu);  // without textures: the plain type byte
    Mat r{type | (front ? 0x100u : 0u), ld3(m.albedo), m.alpha, m.alphaY, id, make_float2(0.0f, 0.0f), s.mats,
          s.textures, s.texels, s.matTex};
    if (type & kMatTextured) {
        r.uv = hitTexcoord(s, pos, bu, bv);
        const uint4 b = s.matTex[id];
        if (b.x != kTexNone) {
            const float4 c = sampleTexture(s.textures, s.texels, b.x, r.uv);
            r.albedo = {c.x, c.y, c.z};
        }
        const uint32_t texY = m.texY;
        if (b.z != kTexNone || texY != kTexNone) {  // the roughness at this hit picks the lobe as each surface's init does
            const float rough = b.z != kTexNone ? sampleTexture(s.textures, s.texels, b.z, r.uv).x : m.roughX;
            uint32_t t = b.w;
            if (b.w == 0u && rough > 0.0f) {  // matteInit: Oren-Nayar, sigma = roughness x pi
                const float sigma = rough * kPi, s2 = sigma * sigma;
                t = kDevOrenNayar;
                r.alpha = 1.0f - fdiv(s2, 2.0f * (s2 + 0.33f));
                r.aux = fdiv(0.45f * s2, s2 + 0.09f);
            } else if (b.w == 1u) {
                r.alpha = fmaxf(rough * rough, 1e-7f);
            } else if (b.w == 2u) {
                r.alpha = remapRoughness(rough);
            } else if (b.w == 4u || b.w == 5u) {
                // glassInit (Classic/Kernel.cpp:630-655): smooth only when both roughnesses are zero;
                // substrateInit (:769-773): always the remapped pair
                const float roughY = texY != kTexNone ? sampleTexture(s.textures, s.texels, texY, r.uv).x : m.roughY;
                if (b.w == 5u || rough > 0.0f || roughY > 0.0f) {
                    if (b.w == 4u) t = kDevRoughGlass;
                    r.alpha = remapRoughness(rough);
                    r.aux = remapRoughness(roughY);
                } else {
                    r.alpha = 0.0f;
                }
            }
            r.bits = (r.bits & ~0xFFu) | t | kRoughFromTexture;
        }
    }
    return r;
}

// Plastic's glossy lobe: MicrofacetReflection<FresnelDielectric> (Classic/Kernel.cpp:384-433,
// plasticInit :441-471, eta 1.46), with the exact visible-normal pdf as the conductor.
__device__ __forceinline__ V3 plasticGlossyF(const Mat& m, V3 wo, V3 wi) {
    if (wo.z <= 0.0f || wi.z <= 0.0f) return {0.0f, 0.0f, 0.0f};
    const V3 wh = fnormalize(wo + wi);
    const float g = 1.0f + ggxLambda(wo, m.alpha) + ggxLambda(wi, m.alpha);
    const float k = fdiv(ggxD(wh, m.alpha), 4.0f * g * fabsf(wo.z * wi.z));
    const float etaI = m.front() ? 1.0f : 1.46f, etaT = m.front() ? 1.46f : 1.0f;
    return m.spec() * (fresnelDielectric(dot(wi, wh), etaI, etaT) * k);
}
__device__ __forceinline__ float lambertPdf(V3 wi) { return wi.z > 0.0f ? fabsf(wi.z) * kInvPi : 0.0f; }

// Rough glass (glassInit with roughness > 0, Classic/Kernel.cpp:641-655): reflection =
// MicrofacetReflection with the dielectric Fresnel (eta 1.5, face-swapped) scaled by the reflection
// colour; transmission = microfacet refraction (Walter et al. 2007, in pbrt-v3's form with radiance
// scaled by (eta_i / eta_t)^2, which cancels its eta^2) scaled by the transmission colour. The
// reference's MicrofacetTransmission::sample refracts with eta_t / eta_i, so its samples do not
// follow its own pdf; this is the consistent form.
__device__ __forceinline__ float glassEtaI(const Mat& m) { return m.front() ? 1.0f : 1.5f; }
__device__ __forceinline__ float glassEtaT(const Mat& m) { return m.front() ? 1.5f : 1.0f; }
__device__ __forceinline__ V3 roughGlassR(const Mat& m, V3 wo, V3 wi) {
    if (wo.z <= 0.0f || wi.z <= 0.0f) return {0.0f, 0.0f, 0.0f};
    const V3 wh = fnormalize(wo + wi);
    const float g = 1.0f + ggxLambda(wo, m.ggx()) + ggxLambda(wi, m.ggx());
    const float k = fdiv(ggxD(wh, m.ggx()), 4.0f * g * fabsf(wo.z * wi.z));
    return m.albedo * (fresnelDielectric(dot(wi, wh), glassEtaI(m), glassEtaT(m)) * k);
}
__device__ __forceinline__ V3 roughGlassT(const Mat& m, V3 wo, V3 wi) {
    if (wo.z <= 0.0f || wi.z >= 0.0f) return {0.0f, 0.0f, 0.0f};
    const float eta = fdiv(glassEtaT(m), glassEtaI(m));
    V3 wh = fnormalize(wo + wi * eta);
    if (wh.z < 0.0f) wh = -wh;
    const float ow = dot(wo, wh), iw = dot(wi, wh);
    if (!(ow > 0.0f && iw < 0.0f)) return {0.0f, 0.0f, 0.0f};  // Walter's chi+: a microfacet wo sees
    const float fr = fresnelDielectric(ow, glassEtaI(m), glassEtaT(m));
    const float den = ow + eta * iw;
    const float g = 1.0f + ggxLambda(wo, m.ggx()) + ggxLambda(wi, m.ggx());
    const float v = fdiv((1.0f - fr) * ggxD(wh, m.ggx()) * fabsf(iw) * fabsf(ow), g * fabsf(wo.z * wi.z) * den * den);
    return m.spec() * v;
}
__device__ __forceinline__ float roughGlassTPdf(const Mat& m, V3 wo, V3 wi) {
    if (wo.z <= 0.0f || wi.z >= 0.0f) return 0.0f;
    const float eta = fdiv(glassEtaT(m), glassEtaI(m));
    V3 wh = fnormalize(wo + wi * eta);
    if (wh.z < 0.0f) wh = -wh;
    const float ow = dot(wo, wh), iw = dot(wi, wh);
    if (!(ow > 0.0f && iw < 0.0f)) return 0.0f;  // only normals the visible-normal sampler draws
    const float den = ow + eta * iw;
    const float dwh = fabsf(fdiv(eta * eta * iw, den * den));  // d omega_h / d omega_i
    return fdiv(ggxD(wh, m.ggx()) * fabsf(ow), wo.z * (1.0f + ggxLambda(wo, m.ggx()))) * dwh;
}
__device__ __forceinline__ uint32_t glassLobes(const Mat& m) { return (valid(m.albedo) ? 1u : 0u) | (valid(m.spec()) ? 2u : 0u); }
__device__ __forceinline__ V3 roughGlassEval(const Mat& m, V3 wo, V3 wi, V3 ng) {
    const uint32_t lobes = glassLobes(m);
    if (dot(wo, ng) * dot(wi, ng) > 0.0f) return (lobes & 1u) ? roughGlassR(m, wo, wi) : V3{0.0f, 0.0f, 0.0f};
    return (lobes & 2u) ? roughGlassT(m, wo, wi) : V3{0.0f, 0.0f, 0.0f};
}
__device__ __forceinline__ float roughGlassPdf(const Mat& m, V3 wo, V3 wi) {
    const uint32_t lobes = glassLobes(m);
    float p = 0.0f;
    if (lobes & 1u) p += ggxReflectPdf(m.ggx(), wo, wi);
    if (lobes & 2u) p += roughGlassTPdf(m, wo, wi);
    return lobes == 3u ? p / 2.0f : p;
}

// OrenNayar::evaluate (Classic/Kernel.cpp:108-131): A and B are precomputed on upload; the value is
// zero when either direction is within sin(theta) < 1e-4 of the normal, as the reference has it.
__device__ __forceinline__ V3 orenNayarF(const Mat& m, V3 wo, V3 wi) {
    const float sinI = fsqrt(fmaxf(0.0f, 1.0f - wi.z * wi.z));
    const float sinO = fsqrt(fmaxf(0.0f, 1.0f - wo.z * wo.z));
    if (sinI < 1e-4f || sinO < 1e-4f) return {0.0f, 0.0f, 0.0f};
    const float cosPhiI = fminf(1.0f, fmaxf(-1.0f, fdiv(wi.x, sinI))), sinPhiI = fminf(1.0f, fmaxf(-1.0f, fdiv(wi.y, sinI)));
    const float cosPhiO = fminf(1.0f, fmaxf(-1.0f, fdiv(wo.x, sinO))), sinPhiO = fminf(1.0f, fmaxf(-1.0f, fdiv(wo.y, sinO)));
    const float maxCos = fmaxf(0.0f, cosPhiI * cosPhiO + sinPhiI * sinPhiO);
    const bool iSteeper = fabsf(wi.z) > fabsf(wo.z);
    const float sinAlpha = iSteeper ? sinO : sinI;
    const float tanBeta = iSteeper ? fdiv(sinI, fabsf(wi.z)) : fdiv(sinO, fabsf(wo.z));
    const float b = (m.bits & kRoughFromTexture) ? m.aux : m.spec().x;
    return m.albedo * (kInvPi * (m.alpha + b * maxCos * sinAlpha * tanBeta));
}

// FresnelBlend::evaluate (Classic/Kernel.cpp:705-722): Ashikhmin-Shirley diffuse + Schlick-weighted
// GGX specular (substrate: diffuse = albedo, specular = spec).
__device__ __forceinline__ float pow5(float v) { return (v * v) * (v * v) * v; }
__device__ __forceinline__ V3 fresnelBlendF(const Mat& m, V3 wo, V3 wi) {
    const V3 rs = m.spec();
    const V3 negRem = rs - V3{1.0f, 1.0f, 1.0f};
    const float kd = (-28.0f / (23.0f * kPi)) * (1.0f - pow5(1.0f - 0.5f * fabsf(wi.z))) *
                     (1.0f - pow5(1.0f - 0.5f * fabsf(wo.z)));
    const V3 wh = fnormalize(wo + wi);
    const float c = dot(wi, wh);
    const V3 schlick = rs + negRem * pow5(c - 1.0f);
    const float ks = fdiv(ggxD(wh, m.ggx()), 4.0f * fmaxf(fabsf(wi.z), fabsf(wo.z)) * fabsf(c));
    return mul(m.albedo, negRem) * kd + schlick * ks;
}
// FresnelBlend::pdf with the visible-normal density of its GGX half (the reference's pdf divides by
// 1 + Lambda(wo) + Lambda(wh), which does not match the sampler it pairs with).
__device__ __forceinline__ float fresnelBlendPdf(const Mat& m, V3 wo, V3 wi) {
    if (wi.z <= 0.0f) return 0.0f;
    const V3 wh = fnormalize(wo + wi);
    return 0.5f * (fabsf(wi.z) * kInvPi + fdiv(ggxD(wh, m.ggx()), 4.0f * fabsf(wo.z) * (1.0f + ggxLambda(wo, m.ggx()))));
}

// Mirror and smooth glass are specular: their init clears noSpecular, so the vertex does no light
// sampling and the next hit's emission counts (PathTracing/Kernel.cpp:119, 158-163, 189).
__device__ __forceinline__ bool isSpecular(const Mat& m) { return m.type() == 3u || m.type() == 4u; }

// The generic BxDF mixture (Classic/Kernel.cpp:144-237) over plastic's valid lobes (diffuse,
// glossy): evaluate sums them, pdf averages them, sample picks one by u1.
__device__ __forceinline__ uint32_t plasticLobes(const Mat& m) { return (valid(m.albedo) ? 1u : 0u) | (valid(m.spec()) ? 2u : 0u); }

__device__ __forceinline__ V3 bsdfEval(const Mat& m, V3 wo, V3 wi, V3 ng) {
    if (m.type() == kDevRoughGlass) return roughGlassEval(m, wo, wi, ng);
    if (!(dot(wo, ng) * dot(wi, ng) > 0.0f)) return {0.0f, 0.0f, 0.0f};  // Classic/Kernel.cpp:220
    if (m.type() == 0) return m.albedo * kInvPi;
    if (m.type() == 1) return conductorF(m.albedo, m.alpha, wo, wi);
    if (m.type() == kDevOrenNayar) return orenNayarF(m, wo, wi);
    if (m.type() == 5u) return fresnelBlendF(m, wo, wi);
    if (m.type() != 2u) return {0.0f, 0.0f, 0.0f};  // specular: blackBodyEvaluate / no lobe to evaluate
    const uint32_t lobes = plasticLobes(m);
    V3 f{0.0f, 0.0f, 0.0f};
    if (lobes & 1u) f = f + m.albedo * kInvPi;
    if (lobes & 2u) f = f + plasticGlossyF(m, wo, wi);
    return f;
}
__device__ __forceinline__ float bsdfPdf(const Mat& m, V3 wo, V3 wi) {
    if (m.type() == 0 || m.type() == kDevOrenNayar) return lambertPdf(wi);
    if (m.type() == 1) return conductorPdf(m.alpha, wo, wi);
    if (m.type() == 5u) return fresnelBlendPdf(m, wo, wi);
    if (m.type() == kDevRoughGlass) return roughGlassPdf(m, wo, wi);
    if (m.type() != 2u) return 0.0f;
    const uint32_t lobes = plasticLobes(m);
    float p = 0.0f;
    if (lobes & 1u) p = p + lambertPdf(wi);
    if (lobes & 2u) p = p + conductorPdf(m.alpha, wo, wi);
    return lobes == 3u ? p / 2.0f : p;
}
struct BsdfSample {
    V3 wi, f;
    float pdf;
};
template <typename A>  // float: isotropic alpha; Ggx: the anisotropic pair
__device__ __forceinline__ BsdfSample ggxReflectSample(V3 wo, A alpha, float u1, float u2) {
    BsdfSample s{{0, 0, 0}, {0, 0, 0}, 0.0f};
    if (wo.z <= 0.0f) return s;
    const V3 wh = ggxSampleWh(wo, alpha, u1, u2);
    const float c = dot(wo, wh);
    if (c < 0.0f) return s;
    s.wi = wh * (2.0f * c) - wo;
    if (s.wi.z <= 0.0f) return s;
    if constexpr (std::is_same<A, Ggx>::value) s.pdf = ggxReflectPdf(alpha, wo, s.wi);
    else s.pdf = conductorPdf(alpha, wo, s.wi);
    return s;
}
__device__ __forceinline__ BsdfSample bsdfSample(const Mat& m, V3 wo, V3 ng, float u1, float u2) {
    BsdfSample s{{0, 0, 0}, {0, 0, 0}, 0.0f};
    if (m.type() == 0 || m.type() == kDevOrenNayar) {  // DefaultBxDFHelper::sample (Classic/Kernel.cpp:67-74)
        s.wi = cosineHemisphere(u1, u2);
        s.f = m.type() == 0 ? m.albedo * kInvPi : orenNayarF(m, wo, s.wi);
        s.pdf = lambertPdf(s.wi);
        return s;
    }
    if (m.type() == 1) {
        s = ggxReflectSample(wo, m.alpha, u1, u2);
        if (s.pdf > 0.0f) s.f = conductorF(m.albedo, m.alpha, wo, s.wi);
        return s;
    }
    if (m.type() == 3u) {  // mirrorSample (Classic/Kernel.cpp:678-692)
        s.wi = V3{-wo.x, -wo.y, wo.z};
        s.f = m.albedo * __builtin_amdgcn_rcpf(fabsf(s.wi.z));
        s.pdf = 1.0f;
        return s;
    }
    if (m.type() == 4u) {  // FresnelSpecular::sample (Classic/Kernel.cpp:522-540), glassInit eta 1.5 (:608-640)
        const V3 rt = m.spec();
        if (!valid(m.albedo) && !valid(rt)) return s;  // glassInit: mask 0
        const float etaI = m.front() ? 1.0f : 1.5f, etaT = m.front() ? 1.5f : 1.0f;
        const float fr = fresnelDielectric(wo.z, etaI, etaT);
        if (u1 < fr) {
            s.wi = V3{-wo.x, -wo.y, wo.z};
            s.pdf = fr;
            s.f = m.albedo * fdiv(fr, fabsf(s.wi.z));
        } else {  // refract(wo, +z, etaI / etaT), Kernel/Transform.hpp:410-421
            const float eta = fdiv(etaI, etaT);
            const float cosI = wo.z;
            const float sin2T = eta * eta * fmaxf(0.0f, 1.0f - cosI * cosI);
            if (sin2T >= 1.0f) return s;
            const float cosT = fsqrt(fmaxf(0.0f, 1.0f - sin2T));
            s.wi = fnormalize(V3{0.0f, 0.0f, eta * cosI - cosT} - wo * eta);
            s.pdf = 1.0f - fr;
            s.f = rt * fdiv(eta * eta * s.pdf, fabsf(s.wi.z));
        }
        return s;
    }
    if (m.type() == kDevRoughGlass) {  // the generic mixture over the reflection and transmission lobes
        const uint32_t lobes = glassLobes(m);
        const uint32_t count = __popc(lobes);
        if (count == 0) return s;
        const uint32_t select = min(static_cast<uint32_t>(floorf(u1 * static_cast<float>(count))), count - 1u);
        u1 = u1 * static_cast<float>(count) - static_cast<float>(select);
        const bool reflect = (lobes & 1u) && select == 0;
        if (reflect) {
            s = ggxReflectSample(wo, m.ggx(), u1, u2);
            if (s.pdf > 0.0f) s.f = roughGlassR(m, wo, s.wi);
        } else {  // a visible microfacet normal, then Snell's law about it with eta_i / eta_t
            if (wo.z <= 0.0f) return s;
            const V3 wh = ggxSampleWh(wo, m.ggx(), u1, u2);
            const float c = dot(wo, wh);
            if (c < 0.0f) return s;
            const float eta = fdiv(glassEtaI(m), glassEtaT(m));
            const float sin2T = eta * eta * fmaxf(0.0f, 1.0f - c * c);
            if (sin2T >= 1.0f) return s;
            const float cosT = fsqrt(fmaxf(0.0f, 1.0f - sin2T));
            s.wi = fnormalize(wh * (eta * c - cosT) - wo * eta);
            s.pdf = roughGlassTPdf(m, wo, s.wi);
            if (s.pdf > 0.0f) s.f = roughGlassT(m, wo, s.wi);
        }
        if (!(s.pdf > 0.0f)) return s;
        if (count > 1) {  // the other lobe: its value where its side matches, its pdf always
            const bool same = dot(wo, ng) * dot(s.wi, ng) > 0.0f;
            if (reflect && !same) s.f = s.f + roughGlassT(m, wo, s.wi);
            if (!reflect && same) s.f = s.f + roughGlassR(m, wo, s.wi);
            s.pdf = (s.pdf + (reflect ? roughGlassTPdf(m, wo, s.wi) : ggxReflectPdf(m.ggx(), wo, s.wi))) / 2.0f;
        }
        return s;
    }
    if (m.type() == 5u) {  // FresnelBlend::sample (Classic/Kernel.cpp:734-750)
        if (u1 < 0.5f) {
            s.wi = cosineHemisphere(fminf(2.0f * u1, 1.0f), u2);
        } else {
            const V3 wh = ggxSampleWh(wo, m.ggx(), fminf(2.0f * (u1 - 0.5f), 1.0f), u2);
            s.wi = wh * (2.0f * dot(wo, wh)) - wo;
            if (s.wi.z <= 0.0f) return s;
        }
        s.pdf = fresnelBlendPdf(m, wo, s.wi);
        s.f = fresnelBlendF(m, wo, s.wi);
        return s;
    }
    const uint32_t lobes = plasticLobes(m);
    const uint32_t count = __popc(lobes);
    if (count == 0) return s;
    const uint32_t select = min(static_cast<uint32_t>(floorf(u1 * static_cast<float>(count))), count - 1u);
    u1 = u1 * static_cast<float>(count) - static_cast<float>(select);
    const bool diffuse = (lobes & 1u) && select == 0;  // lobes in order: diffuse, glossy
    if (diffuse) {
        s.wi = cosineHemisphere(u1, u2);
        s.f = m.albedo * kInvPi;
        s.pdf = lambertPdf(s.wi);
    } else {
        s = ggxReflectSample(wo, m.alpha, u1, u2);
        if (s.pdf > 0.0f) s.f = plasticGlossyF(m, wo, s.wi);
    }
    if (!(s.pdf > 0.0f)) return s;
    if (count > 1) {  // the other lobe: its value on the same side, its pdf always
        if (dot(wo, ng) * dot(s.wi, ng) > 0.0f) s.f = s.f + (diffuse ? plasticGlossyF(m, wo, s.wi) : m.albedo * kInvPi);
        s.pdf = s.pdf + (diffuse ? conductorPdf(m.alpha, wo, s.wi) : lambertPdf(s.wi));
        s.pdf = fdiv(s.pdf, static_cast<float>(count));
    }
    return s;
}

// ------------------------------------------------------------------------------------------
// Lights: uniformSelect (BasicLight/Kernel.cpp:91-98), diffuseSample over planeSample
// (AreaLight/Kernel.cpp:24-41, BasicShapes/Kernel.cpp:113-134), SESample (BasicLight :63-76)
// ------------------------------------------------------------------------------------------
struct LightSample {
    V3 dir, rad;
    float dist, pdf;
};
// select (Kernel/Sampling.hpp:52-63) over an area light's primitive CDF: the first primitive whose
// CDF entry exceeds u (same bisection), then u rescaled into that primitive: by the CDF's lower
// bound in spec mode, by its upper bound cdf[r] with D1 as the reference does. u = 1 (the Sobol
// conversion has no clamp) picks the last primitive where the reference reads past its CDF.
__device__ __forceinline__ uint32_t selectPrimitive(const DevScene& s, const DevAreaLight L, float& u, bool d1) {
    uint32_t l = 0, r = L.count;
    while (l < r) {
        const uint32_t mid = (l + r) >> 1;
        if (u >= s.quads[L.first + mid].cdf) l = mid + 1;
        else r = mid;
    }
    r = min(r, L.count - 1u);
    const float hiCdf = s.quads[L.first + r].cdf;
    const float lower = d1 ? hiCdf : (r ? s.quads[L.first + r - 1u].cdf : 0.0f);
    u = fdiv(u - lower, s.quads[L.first + r].pdf);
    return L.first + r;
}

// The primitive of area light `light` (0-based) that the ray meets first, by closestHit's quad
// rule (smaller t, ties to the smaller index), or false.
__device__ __forceinline__ bool firstLightPrimitive(const DevScene& s, uint32_t light, const Ray& r, float& tq,
                                                    uint32_t& qi) {
    const DevAreaLight L = s.lights[light];
    bool hit = false;
    tq = 1e5f;
    for (uint32_t k = L.first; k < L.first + L.count; ++k) {
        float t;
        if (intersectQuad(r, s.quads[k], 1e-3f, tq, t)) tq = t, qi = k, hit = true;
    }
    return hit;
}
// No quad outside area light `light` meets the ray before quad qi at tq (closestHit's quad order).
__device__ __forceinline__ bool noOtherLightFirst(const DevScene& s, uint32_t light, const Ray& r, float tq,
                                                  uint32_t qi) {
    const DevAreaLight L = s.lights[light];
    for (uint32_t k = 0; k < s.numQuads; ++k) {
        if (k == L.first) {
            k += L.count - 1u;
            continue;
        }
        float t2;
        if (intersectQuad(r, s.quads[k], 1e-3f, 1e5f, t2) && (t2 < tq || (t2 == tq && k < qi))) return false;
    }
    return true;
}

__device__ __forceinline__ LightSample sampleLight(const DevScene& s, uint32_t light, V3 hit, float u1, float u2,
                                                   bool d1) {
    LightSample ls{{0, 0, 0}, {0, 0, 0}, 0.0f, 0.0f};
    if (light == 0) {
        const float theta = u1 * kPi, phi = u2 * kPi;
        const float st = sinf(theta);
        ls.dir = {st * cosf(phi), st * sinf(phi), cosf(theta)};
        ls.rad = ld3(s.env);
        ls.pdf = st == 0.0f ? 0.0f : __builtin_amdgcn_rcpf(2.0f * kPi * kPi * st);
        ls.dist = 1e5f;
        return ls;
    }
    if (light > s.numLights) {  // pointSample (Light/BasicLight/Kernel.cpp:35-45): a delta light, pdf 1
        const float* p = s.points + 6 * (light - 1u - s.numLights);
        const V3 delta = V3{p[0], p[1], p[2]} - hit;
        const float dis2 = dot(delta, delta);
        const float inv2 = __builtin_amdgcn_rcpf(dis2);
        ls.rad = {p[3] * inv2, p[4] * inv2, p[5] * inv2};
        ls.dist = fsqrt(fmaxf(dis2, 0.0f));
        const float invD = __builtin_amdgcn_rcpf(ls.dist);
        ls.dir = {delta.x * invD, delta.y * invD, delta.z * invD};
        ls.pdf = 1.0f;
        return ls;
    }
    float uu = u1;
    const DevQuad& q = s.quads[selectPrimitive(s, s.lights[light - 1u], uu, d1)];
    const V3 src = ld3(q.origin) + ld3(q.u) * uu + ld3(q.v) * u2;
    V3 n = ld3(q.normal);
    if (dot(hit - src, n) <= 0.0f) n = -n;
    const float pdfA = __builtin_amdgcn_rcpf(q.area);
    const V3 delta = src - hit;
    if (dot(delta, n) >= 0.0f) return ls;
    const float dis2 = dot(delta, delta);
    ls.rad = ld3(q.radiance);
    ls.dist = fsqrt(dis2);
    const float invD = __builtin_amdgcn_rcpf(ls.dist);
    ls.dir = {delta.x * invD, delta.y * invD, delta.z * invD};
    ls.pdf = fdiv(pdfA * dis2, fabsf(dot(n, ls.dir)));
    return ls;
}
__device__ __forceinline__ float envPdf(V3 dir) {  // SEPdf, BasicLight/Kernel.cpp:82-89
    const float theta = acosf(fminf(1.0f, fmaxf(-1.0f, dir.z)));
    const float st = sinf(theta);
    return st == 0.0f ? 0.0f : __builtin_amdgcn_rcpf(2.0f * kPi * kPi * st);
}
__device__ __forceinline__ float powerHeuristic(float a, float b) {  // PathTracing/Kernel.cpp:27-31
    const float sa = a * a, sb = b * b;
    return fdiv(sa, sa + sb);
}

// Wave-level compaction: every lane of the wave calls this together; lanes with pred get
// consecutive slots from one atomic per wave.
__device__ __forceinline__ uint32_t wavePush(uint32_t* counter, bool pred) {
    const uint64_t mask = __ballot(pred);
    if (mask == 0) return 0;
    const uint32_t lane = __lane_id();
    const uint32_t leader = static_cast<uint32_t>(__builtin_ctzll(mask));
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, static_cast<uint32_t>(__popcll(mask)));
    base = __shfl(base, static_cast<int>(leader));
    return base + static_cast<uint32_t>(__popcll(mask & ((1ull << lane) - 1ull)));
}

}  // namespace pg

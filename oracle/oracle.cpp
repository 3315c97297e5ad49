// CPU oracle: a plain C++ restatement of Piper2's CPU path-tracing path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker that the HIP product path is compared
// against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the
// library built from it. The product library never links or calls it.
//
// Each function cites the reference file:line it restates (paths relative to the Piper2 tree).
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off, so no FMA contraction: the camera, the
// watertight triangle test and the Sobol float conversion must match the GPU bit for bit).
//
// Reference defects (spec mode corrects them; the compat bits reproduce the reference):
//   D1  light CDF remap uses the upper bound (Kernel/Sampling.hpp:61)          -> OR_COMPAT_D1
//   D3  shading T/B not normalised (Tracer/Embree/Kernel.cpp:180-186)           -> OR_COMPAT_D3
//   D9  u2 drawn before u1 under the MS ABI (Tracer/Embree/Kernel.cpp:117,145)  -> OR_COMPAT_D9
//   D13 light pass-through decrements depth, wrapping at 0
//       (Integrator/PathTracing/Kernel.cpp:143)                                 -> OR_COMPAT_D13
// GGX sampling uses exact visible-normal sampling with its exact pdf in both modes (the reference's
// pdf omits D(wh), Surface/Classic/Kernel.cpp:405-412, so its plain estimator is biased).
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------------------------
// Sobol (Sampler/ScrambledSobolSampler/Kernel.cpp:23-67, ScrambledSobolSampler.cpp:66-72)
// ---------------------------------------------------------------------------------------------
constexpr uint32_t kSobolDims = 1024;  // Shared.hpp:22
constexpr uint32_t kSobolCols = 52;    // Shared.hpp:23

struct JoeKuo {
    uint32_t s, a;
    uint32_t m[18];
};
const JoeKuo kJoeKuo[kSobolDims] = {
#include "../piper2_amd/csrc/joe_kuo_1024.inc"
};

uint32_t g_sobol[kSobolDims * kSobolCols];
// Interval-to-index tables for resolution 2^m, m = 1..16: column c of the inverse of the
// (2m x 2m) GF(2) map from the low 2m index bits to (top m bits of dim 0, top m bits of dim 1),
// and the contribution of sample-index bit c (index bit 2m+c) to those 2m output bits.
uint64_t g_inv[17][32];
uint64_t g_delta[17][64];

// Bratley-Fox / Joe-Kuo recurrence at 64-bit precision; column j of dimension d keeps the high
// 32 bits (v_j = m_j / 2^j).
void buildSobolMatrices() {
    for (uint32_t d = 0; d < kSobolDims; ++d) {
        uint64_t v[kSobolCols];
        if (d == 0) {
            for (uint32_t j = 0; j < kSobolCols; ++j) v[j] = 1ULL << (63 - j);
        } else {
            const uint32_t s = kJoeKuo[d].s, a = kJoeKuo[d].a;
            for (uint32_t j = 0; j < kSobolCols; ++j) {
                if (j < s) {
                    v[j] = static_cast<uint64_t>(kJoeKuo[d].m[j]) << (63 - j);
                } else {
                    uint64_t x = v[j - s] ^ (v[j - s] >> s);
                    for (uint32_t k = 1; k < s; ++k)
                        if ((a >> (s - 1 - k)) & 1) x ^= v[j - k];
                    v[j] = x;
                }
            }
        }
        for (uint32_t j = 0; j < kSobolCols; ++j) g_sobol[d * kSobolCols + j] = static_cast<uint32_t>(v[j] >> 32);
    }
}

// Output bits (top m of dim 0 << m | top m of dim 1) produced by index bit j.
uint64_t topBits(uint32_t m, uint32_t j) {
    const uint32_t x = g_sobol[0 * kSobolCols + j] >> (32 - m);
    const uint32_t y = g_sobol[1 * kSobolCols + j] >> (32 - m);
    return (static_cast<uint64_t>(x) << m) | y;
}

void buildIntervalTables() {
    for (uint32_t m = 1; m <= 16; ++m) {
        const uint32_t n = 2 * m;
        // Gauss-Jordan over GF(2): columns A[j] = topBits(j), solve A x = e_i for each unit e_i.
        uint64_t rows[32];  // augmented: row r has bits of A's row r (n bits) | identity (n bits) << n
        for (uint32_t r = 0; r < n; ++r) {
            uint64_t row = 0;
            for (uint32_t j = 0; j < n; ++j)
                if ((topBits(m, j) >> r) & 1) row |= 1ULL << j;
            rows[r] = row | (1ULL << (n + r));
        }
        for (uint32_t c = 0; c < n; ++c) {
            uint32_t p = c;
            while (p < n && !((rows[p] >> c) & 1)) ++p;
            if (p == n) std::abort();  // singular: not a (0,2)-sequence
            std::swap(rows[p], rows[c]);
            for (uint32_t r = 0; r < n; ++r)
                if (r != c && ((rows[r] >> c) & 1)) rows[r] ^= rows[c];
        }
        // rows[j] = (e_j | inverse row j); inverse column i = bits i of rows[*] >> n
        for (uint32_t i = 0; i < n; ++i) {
            uint64_t col = 0;
            for (uint32_t j = 0; j < n; ++j)
                if ((rows[j] >> (n + i)) & 1) col |= 1ULL << j;
            g_inv[m][i] = col;
        }
        for (uint32_t c = 0; c + n < kSobolCols && c < 64; ++c) g_delta[m][c] = topBits(m, n + c);
    }
}

struct SobolInit {
    SobolInit() {
        buildSobolMatrices();
        buildIntervalTables();
    }
} g_sobolInit;

// sobolSampleImpl, Kernel.cpp:23-29 (scramble is always 0, ScrambledSobolSampler.cpp:52-53)
float sobolSample(uint64_t index, uint32_t dim) {
    if (dim >= kSobolDims) {
        // Beyond the table the reference draws from mt19937_64 (Tracer/Embree/Kernel.cpp:37-45);
        // both sides of this build use a splitmix64 hash of (index, dim) instead.
        uint64_t z = index ^ (static_cast<uint64_t>(dim) * 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return static_cast<float>(static_cast<uint32_t>(z >> 40)) * 0x1p-24f;
    }
    uint32_t v = 0;
    for (uint32_t i = dim * kSobolCols; index; index >>= 1, ++i)
        if (index & 1) v ^= g_sobol[i];
    return static_cast<float>(v) * 0x1p-32f;
}

// sobolIntervalToIndex, Kernel.cpp:31-50: the unique index whose top bits are `sample` and whose
// first two coordinates fall in pixel (px, py) at resolution 2^m.
uint64_t intervalToIndex(uint32_t m, uint32_t sample, uint32_t px, uint32_t py) {
    if (m == 0) return 0;
    const uint32_t n = 2 * m;
    uint64_t index = static_cast<uint64_t>(sample) << n;
    uint64_t b = (static_cast<uint64_t>(px) << m) | py;
    for (uint32_t c = 0; sample; sample >>= 1, ++c)
        if (sample & 1) b ^= g_delta[m][c];
    for (uint32_t c = 0; b; b >>= 1, ++c)
        if (b & 1) index ^= g_inv[m][c];
    return index;
}

// ---------------------------------------------------------------------------------------------
// Small vector math (Kernel/Transform.hpp)
// ---------------------------------------------------------------------------------------------
struct V3 {
    float x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 mul(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
// Normal's normalising constructor, Transform.hpp:244-248
inline V3 normalize(V3 a) {
    const float l = length(a);
    return {a.x / l, a.y / l, a.z / l};
}
inline float maxc(V3 a) { return std::fmax(a.x, std::fmax(a.y, a.z)); }
inline bool valid(V3 a) { return a.x > 0.0f || a.y > 0.0f || a.z > 0.0f; }  // Protocol.hpp Spectrum::valid
inline float comp(V3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;

// ---------------------------------------------------------------------------------------------
// Scene
// ---------------------------------------------------------------------------------------------
enum : uint32_t { kMatLambert = 0, kMatConductor = 1, kMatPlastic = 2, kMatMirror = 3, kMatGlass = 4, kMatSubstrate = 5 };
struct Material {
    uint32_t type;
    V3 albedo;          // Lambert rho, conductor F0, plastic diffuse
    float alpha;        // GGX alpha (plastic: remapRoughness of the scene's roughness; glass, substrate:
                        // of RoughnessX)
    float alphaY = 0.0f;  // GGX alpha along the shading frame's y axis (glass, substrate: of
                          // RoughnessY; every other surface: alpha)
    float roughX = 0.0f, roughY = 0.0f;  // glass, substrate: RoughnessX / RoughnessY as configured
    uint32_t texY = ~0u;  // glass, substrate: texture of RoughnessY (~0u: the constant)
    bool ownY = false;    // RoughnessY given on its own (or_scene_set_roughness_y), not RoughnessX's
    V3 spec{0, 0, 0};   // plastic / substrate specular, glass transmission
    bool front = true;  // plastic: the hit is on the triangle's front face
    float onA = 1.0f, onB = 0.0f;  // matte with roughness > 0: Oren-Nayar A and B
    bool orenNayar = false;
};
struct Quad {  // BasicShapes/Shared.hpp:22-28, BasicShapes.cpp:70-90
    V3 origin, u, v, normal, radiance;
    float area;           // area of the owning light: every primitive of its Plane (BasicShapes.cpp:86)
    uint32_t light = 0;   // owning area light
    float cdf = 1.0f, pdf = 1.0f;  // area CDF and probability within the light (BasicShapes.cpp:87-94)
};
// Area light: DiffuseAreaLight over a Plane with `count` primitives, quads [first, first + count).
struct AreaLight {
    uint32_t first, count;
};
// Textures: CheckBoard (ProceduralTexture/Kernel.cpp:28-38) and 8-bit sampled images
// (SampledTexture/Kernel.cpp:22-82).
struct Texture {
    uint32_t type;  // 1 checkerboard, 2 image
    uint32_t channels;
    float black[4], white[4], scale;
    uint32_t width, height, wrap;  // wrap 0 repeat, 1 mirror
    std::vector<uint8_t> texels;
};
struct PointLight {  // PointLightData, BasicLight/Shared.hpp:26-29 (position from the node)
    V3 pos, intensity;
};
struct Camera {  // Sensor/PerspectiveCamera/Shared.hpp:22-29
    V3 base, lookAt, up;
    float sizeX, sizeY, focal, aperture;
};
struct Tri {
    V3 v0, v1, v2;
};
struct BNode {
    float lo[3], hi[3];
    uint32_t left, count;  // count > 0: leaf over prims [left, left+count)
};

struct Scene {
    std::vector<Tri> tris;
    std::vector<uint32_t> triObj, objMat;
    std::vector<Material> mats;
    std::vector<Quad> quads;
    std::vector<AreaLight> lights;
    std::vector<PointLight> points;
    std::vector<float> triUV;  // 6 per triangle: texture coordinates of corners 0, 1, 2
    std::vector<Texture> textures;
    std::vector<std::array<uint32_t, 3>> matTex;  // per material: albedo, spec, roughness texture (~0u: none)
    Camera cam;
    V3 env;
    std::vector<BNode> nodes;
    std::vector<uint32_t> order;  // BVH leaf order -> triangle index
    // Optional wide BVH as the GPU library builds and lays it out (pg_scene_bvh_download): nodes of
    // `wideWidth` children, each six planes of W floats (lo x, y, z, hi x, y, z) then W child
    // references (leaf: top bit, (count - 1) << 27, first triangle); triangles in leaf order as
    // 12 floats (v0, original index bits, v1, -, v2, -). When set, rays traverse it instead.
    uint32_t wideWidth = 0;
    std::vector<uint32_t> wideNodes;  // raw 32-bit words
    std::vector<float> wideTris;
    // Moving instances: triangles [numStatic, tris.size()) are instance meshes in their own space.
    struct Instance {
        uint32_t first = 0, count = 0;
        uint32_t keys = 1;
        bool visible = true;
        double t0 = -INFINITY, t1 = INFINITY;
        double srt[8][16] = {};  // TransformSRT keys, unit quaternions
        double parent[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    };
    uint32_t numStatic = 0;
    std::vector<Instance> inst;
};

// Binned-SAH BVH2 over triangle boxes (the oracle's own acceleration structure; only the hit
// result matters, which is defined independently of the tree by the closest-(t, prim) rule).
struct Box {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const Box& b) {
        for (int k = 0; k < 3; ++k) lo[k] = std::fmin(lo[k], b.lo[k]), hi[k] = std::fmax(hi[k], b.hi[k]);
    }
    void grow(V3 p) {
        const float c[3] = {p.x, p.y, p.z};
        for (int k = 0; k < 3; ++k) lo[k] = std::fmin(lo[k], c[k]), hi[k] = std::fmax(hi[k], c[k]);
    }
    float area() const {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx < 0) return 0.0f;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

void buildBVH(Scene& s) {
    const uint32_t n = s.numStatic;  // moving instances are tested apart (instanceHits)
    s.nodes.clear();
    s.order.resize(n);
    if (n == 0) return;
    std::vector<Box> boxes(n);
    std::vector<float> cen(3 * n);
    float maxAbs = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
        boxes[i].grow(s.tris[i].v0), boxes[i].grow(s.tris[i].v1), boxes[i].grow(s.tris[i].v2);
        for (int k = 0; k < 3; ++k) {
            cen[3 * i + k] = 0.5f * (boxes[i].lo[k] + boxes[i].hi[k]);
            maxAbs = std::fmax(maxAbs, std::fmax(std::fabs(boxes[i].lo[k]), std::fabs(boxes[i].hi[k])));
        }
        s.order[i] = i;
    }
    const float pad = 1e-5f * std::fmax(maxAbs, 1.0f);
    struct Item {
        uint32_t node, begin, end;
    };
    s.nodes.push_back({});
    std::vector<Item> work{{0, 0, n}};
    while (!work.empty()) {
        const Item it = work.back();
        work.pop_back();
        Box b, cb;
        for (uint32_t i = it.begin; i < it.end; ++i) {
            const uint32_t t = s.order[i];
            b.grow(boxes[t]);
            cb.grow(V3{cen[3 * t], cen[3 * t + 1], cen[3 * t + 2]});
        }
        BNode& node = s.nodes[it.node];
        for (int k = 0; k < 3; ++k) node.lo[k] = b.lo[k] - pad, node.hi[k] = b.hi[k] + pad;
        const uint32_t cnt = it.end - it.begin;
        int bestAxis = -1;
        uint32_t bestSplit = 0;
        float bestCost = static_cast<float>(cnt) * b.area();
        constexpr int kBins = 16;
        if (cnt > 2) {
            for (int k = 0; k < 3; ++k) {
                const float ext = cb.hi[k] - cb.lo[k];
                if (!(ext > 0.0f)) continue;
                Box bb[kBins];
                uint32_t bc[kBins] = {};
                for (uint32_t i = it.begin; i < it.end; ++i) {
                    const uint32_t t = s.order[i];
                    int bi = static_cast<int>((cen[3 * t + k] - cb.lo[k]) / ext * kBins);
                    bi = std::min(std::max(bi, 0), kBins - 1);
                    bb[bi].grow(boxes[t]);
                    ++bc[bi];
                }
                float rightArea[kBins];
                uint32_t rightCnt[kBins];
                Box acc;
                uint32_t ac = 0;
                for (int i = kBins - 1; i > 0; --i) {
                    acc.grow(bb[i]), ac += bc[i];
                    rightArea[i] = acc.area(), rightCnt[i] = ac;
                }
                Box lacc;
                uint32_t lc = 0;
                for (int i = 0; i < kBins - 1; ++i) {
                    lacc.grow(bb[i]), lc += bc[i];
                    if (lc == 0 || rightCnt[i + 1] == 0) continue;
                    const float cost = 0.125f * b.area() + lacc.area() * lc + rightArea[i + 1] * rightCnt[i + 1];
                    if (cost < bestCost) bestCost = cost, bestAxis = k, bestSplit = static_cast<uint32_t>(i + 1);
                }
            }
        }
        uint32_t mid;
        if (bestAxis >= 0) {
            const float ext = cb.hi[bestAxis] - cb.lo[bestAxis];
            const auto it2 = std::partition(s.order.begin() + it.begin, s.order.begin() + it.end, [&](uint32_t t) {
                int bi = static_cast<int>((cen[3 * t + bestAxis] - cb.lo[bestAxis]) / ext * kBins);
                bi = std::min(std::max(bi, 0), kBins - 1);
                return static_cast<uint32_t>(bi) < bestSplit;
            });
            mid = static_cast<uint32_t>(it2 - s.order.begin());
        } else if (cnt > 4) {
            mid = it.begin + cnt / 2;  // SAH prefers a leaf but it is too large: median split
        } else {
            node.left = it.begin, node.count = cnt;
            continue;
        }
        const uint32_t l = static_cast<uint32_t>(s.nodes.size());
        s.nodes.push_back({});
        s.nodes.push_back({});
        s.nodes[it.node].left = l;
        s.nodes[it.node].count = 0;
        work.push_back({l, it.begin, mid});
        work.push_back({l + 1, mid, it.end});
    }
}

// ---------------------------------------------------------------------------------------------
// Ray casting (Tracer/Embree/Embree.cpp:267-391)
// ---------------------------------------------------------------------------------------------
struct Ray {
    V3 o, d;
};
constexpr uint32_t kMiss = 0xFFFFFFFFu;
constexpr uint32_t kQuadBit = 0x80000000u;
struct Hit {
    float t;
    uint32_t prim;  // triangle index, kQuadBit|quad, or kMiss
    float u, v;     // barycentrics of v1 and v2 (Embree convention, Embree/Kernel.cpp:176-178)
};

struct RayPre {  // per-ray constants of the watertight test (Woop, Benthin, Wald, JCGT 2013)
    int kx, ky, kz;
    float sx, sy, sz;
    float inv[3];
};
RayPre precompute(const Ray& r) {
    RayPre p;
    const float ad[3] = {std::fabs(r.d.x), std::fabs(r.d.y), std::fabs(r.d.z)};
    p.kz = (ad[0] >= ad[1] && ad[0] >= ad[2]) ? 0 : (ad[1] >= ad[2] ? 1 : 2);
    p.kx = p.kz == 2 ? 0 : p.kz + 1;
    p.ky = p.kx == 2 ? 0 : p.kx + 1;
    if (comp(r.d, p.kz) < 0.0f) std::swap(p.kx, p.ky);
    p.sx = comp(r.d, p.kx) / comp(r.d, p.kz);
    p.sy = comp(r.d, p.ky) / comp(r.d, p.kz);
    p.sz = 1.0f / comp(r.d, p.kz);
    for (int k = 0; k < 3; ++k) {
        float dk = comp(r.d, k);
        if (dk == 0.0f) dk = 1e-20f;  // box test only: avoids 0 * inf
        p.inv[k] = 1.0f / dk;
    }
    return p;
}

// Watertight ray/triangle test. Returns t (or NaN on miss) and barycentrics (u: v1, v: v2).
bool intersectTri(const Ray& r, const RayPre& p, const Tri& tri, float& t, float& bu, float& bv) {
    const V3 a = tri.v0 - r.o, b = tri.v1 - r.o, c = tri.v2 - r.o;
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

// planeIntersect, BasicShapes/Kernel.cpp:30-74 (calcTexCoord + parallelogram bounds)
bool intersectQuad(const Ray& r, const Quad& q, float tNear, float tFar, float& tOut) {
    const V3 delta = q.origin - r.o;
    const float b = dot(delta, q.normal);
    const float k = dot(r.d, q.normal);
    const float t = b / k;
    if (!(tNear < t && t < tFar)) return false;
    const V3 pos = r.d * t - delta;
    const V3 det3 = cross(q.u, q.v);
    const float a0 = std::fabs(det3.x), a1 = std::fabs(det3.y), a2 = std::fabs(det3.z);
    const int sel = (a0 >= a1 && a0 >= a2) ? 0 : ((a1 >= a0 && a1 >= a2) ? 1 : 2);
    float invDet, x, y;
    if (sel == 0) {
        invDet = 1.0f / (q.u.y * q.v.z - q.v.y * q.u.z);
        x = pos.y * q.v.z - q.v.y * pos.z;
        y = q.u.y * pos.z - pos.y * q.u.z;
    } else if (sel == 1) {
        invDet = 1.0f / (q.u.z * q.v.x - q.v.z * q.u.x);
        x = pos.z * q.v.x - q.v.z * pos.x;
        y = q.u.z * pos.x - pos.z * q.u.x;
    } else {
        invDet = 1.0f / (q.u.x * q.v.y - q.v.x * q.u.y);
        x = pos.x * q.v.y - q.v.x * pos.y;
        y = q.u.x * pos.y - pos.x * q.u.y;
    }
    const float u = x * invDet, v = y * invDet;
    if (!(0.0f <= u && u <= 1.0f && 0.0f <= v && v <= 1.0f)) return false;
    tOut = t;
    return true;
}

// Closest-hit rule: accept t in (tmin, best.t], ties broken by the smaller primitive index.
inline void consider(Hit& best, float t, uint32_t prim, float u, float v, float tmin) {
    if (t > tmin && (t < best.t || (t == best.t && prim < best.prim))) best = {t, prim, u, v};
}

bool boxHit(const BNode& n, const Ray& r, const RayPre& p, float tmin, float tmax) {
    float t0 = tmin, t1 = tmax;
    const float o[3] = {r.o.x, r.o.y, r.o.z};
    for (int k = 0; k < 3; ++k) {
        float a = (n.lo[k] - o[k]) * p.inv[k], b = (n.hi[k] - o[k]) * p.inv[k];
        if (a > b) std::swap(a, b);
        t0 = std::fmax(t0, a);
        t1 = std::fmin(t1, b * 1.0000004f);
    }
    return t0 <= t1;
}

bool traverseWide(const Scene& s, const Ray& r, float tmin, float tmax, bool any, Hit& h, uint64_t* counts);

// ---------------------------------------------------------------------------------------------
// Moving instances: a Group child with a keyframed TransformSRT (EmbreeBranchNode, Embree.cpp:
// 777-835) is an Embree instance whose transform updateTransformSRT (:709-773) sets as time steps
// with rtcSetGeometryTimeRange / TimeStepCount / TransformQuaternion. Embree 3.12 (the version
// Tracer/Embree/CMakeLists.txt:3 asks for) then draws it at the ray's time (RTCRay::time,
// Embree.cpp:290, the sample's dimension 2, Embree/Kernel.cpp:77) as its documentation states for
// quaternion motion: over n = keys - 1 equal segments of [t0, t1], segment i = clamp(floor(x), 0,
// n - 1) of x = n (time - t0) / (t1 - t0) at fraction f = x - i; scale, skew, shift and
// translation linear in f; the rotation spherical-linear on the shorter arc (Shoemake 1985,
// q(f) = (sin((1 - f) W) q0 + sin(f W) q1) / sin W with cos W = q0 . q1 >= 0; normalised
// linear interpolation when cos W > 0.9995); matrix T R S with
// S = [[sx, kxy, kxz, hx], [0, sy, kyz, hy], [0, 0, sz, hz]]; the geometry is present only for
// t0 <= time <= t1 (a time range inside the shutter lets it appear and disappear).
// Computed here in double.
// ---------------------------------------------------------------------------------------------
void instanceMatrix(const Scene::Instance& I, double time, double m[12]) {
    uint32_t i = 0;
    double f = 0.0;
    if (I.keys > 1) {
        const double n = I.keys - 1.0;
        const double x = n * (time - I.t0) / (I.t1 - I.t0);
        i = static_cast<uint32_t>(std::min(std::max(std::floor(x), 0.0), n - 1.0));
        f = x - i;
    }
    const double* a = I.srt[i];
    const double* b = I.srt[I.keys > 1 ? i + 1 : i];
    double k[16];
    for (int j = 0; j < 16; ++j) k[j] = (1.0 - f) * a[j] + f * b[j];
    double q1[4] = {b[9], b[10], b[11], b[12]};
    double cw = a[9] * q1[0] + a[10] * q1[1] + a[11] * q1[2] + a[12] * q1[3];
    if (cw < 0.0) {
        cw = -cw;
        for (double& c : q1) c = -c;
    }
    double q[4];
    if (cw > 0.9995) {
        double nn = 0.0;
        for (int j = 0; j < 4; ++j) q[j] = (1.0 - f) * a[9 + j] + f * q1[j], nn += q[j] * q[j];
        for (double& c : q) c /= std::sqrt(nn);
    } else {
        const double W = std::acos(cw), sw = std::sin(W);
        const double ca = std::sin((1.0 - f) * W) / sw, cb = std::sin(f * W) / sw;
        for (int j = 0; j < 4; ++j) q[j] = ca * a[9 + j] + cb * q1[j];
    }
    const double w = q[0], x = q[1], y = q[2], z = q[3];
    // rotation matrix of the unit quaternion, rows
    const double R[3][3] = {{w * w + x * x - y * y - z * z, 2 * (x * y - w * z), 2 * (x * z + w * y)},
                            {2 * (x * y + w * z), w * w - x * x + y * y - z * z, 2 * (y * z - w * x)},
                            {2 * (x * z - w * y), 2 * (y * z + w * x), w * w - x * x - y * y + z * z}};
    const double S[3][4] = {{k[0], k[3], k[4], k[6]}, {0, k[1], k[5], k[7]}, {0, 0, k[2], k[8]}};
    double L[3][4];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 4; ++c) {
            L[r][c] = R[r][0] * S[0][c] + R[r][1] * S[1][c] + R[r][2] * S[2][c];
            if (c == 3) L[r][c] += k[13 + r];
        }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 4; ++c) {
            m[4 * r + c] = I.parent[4 * r] * L[0][c] + I.parent[4 * r + 1] * L[1][c] + I.parent[4 * r + 2] * L[2][c];
            if (c == 3) m[4 * r + c] += I.parent[4 * r + 3];
        }
}

bool instancePresent(const Scene::Instance& I, double time) { return I.visible && time >= I.t0 && time <= I.t1; }

// The ray in the instance's space at `time`: inverse of the affine map (cofactors / determinant).
Ray instanceRay(const Scene::Instance& I, const Ray& r, double time) {
    double m[12];
    instanceMatrix(I, time, m);
    const double det = m[0] * (m[5] * m[10] - m[6] * m[9]) - m[1] * (m[4] * m[10] - m[6] * m[8]) +
                       m[2] * (m[4] * m[9] - m[5] * m[8]);
    const double inv[3][3] = {{(m[5] * m[10] - m[6] * m[9]) / det, (m[2] * m[9] - m[1] * m[10]) / det,
                               (m[1] * m[6] - m[2] * m[5]) / det},
                              {(m[6] * m[8] - m[4] * m[10]) / det, (m[0] * m[10] - m[2] * m[8]) / det,
                               (m[2] * m[4] - m[0] * m[6]) / det},
                              {(m[4] * m[9] - m[5] * m[8]) / det, (m[1] * m[8] - m[0] * m[9]) / det,
                               (m[0] * m[5] - m[1] * m[4]) / det}};
    const double o[3] = {r.o.x - m[3], r.o.y - m[7], r.o.z - m[11]}, d[3] = {r.d.x, r.d.y, r.d.z};
    float lo[3], ld[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = static_cast<float>(inv[k][0] * o[0] + inv[k][1] * o[1] + inv[k][2] * o[2]);
        ld[k] = static_cast<float>(inv[k][0] * d[0] + inv[k][1] * d[1] + inv[k][2] * d[2]);
    }
    return Ray{{lo[0], lo[1], lo[2]}, {ld[0], ld[1], ld[2]}};
}

// Triangle `prim` in world space at `time` (static triangles are stored there).
Tri worldTri(const Scene& s, uint32_t prim, double time) {
    const Tri& t = s.tris[prim];
    if (prim < s.numStatic) return t;
    for (const Scene::Instance& I : s.inst)
        if (prim >= I.first && prim < I.first + I.count) {
            double m[12];
            instanceMatrix(I, time, m);
            auto ap = [&](V3 v) {
                return V3{static_cast<float>(m[0] * v.x + m[1] * v.y + m[2] * v.z + m[3]),
                          static_cast<float>(m[4] * v.x + m[5] * v.y + m[6] * v.z + m[7]),
                          static_cast<float>(m[8] * v.x + m[9] * v.y + m[10] * v.z + m[11])};
            };
            return Tri{ap(t.v0), ap(t.v1), ap(t.v2)};
        }
    return t;
}

// Every moving instance present at `time`: its triangles against the ray in its own space (the
// ray parameter t is the world one: the direction is mapped, not normalised). any: stop at the
// first t in (tmin, tmax).
bool instanceHits(const Scene& s, const Ray& r, double time, float tmin, float tmax, bool any, Hit& best) {
    for (const Scene::Instance& I : s.inst) {
        if (!instancePresent(I, time)) continue;
        const Ray lr = instanceRay(I, r, time);
        const RayPre lp = precompute(lr);
        float t, u, v;
        for (uint32_t i = I.first; i < I.first + I.count; ++i) {
            if (!intersectTri(lr, lp, s.tris[i], t, u, v)) continue;
            if (any) {
                if (t > tmin && t < tmax) return true;
            } else {
                consider(best, t, i, u, v, tmin);
            }
        }
    }
    return false;
}

// rtcIntersect1 with mask gsm|areaLight (Embree.cpp:267-355): closest triangle or light quad.
Hit traceClosest(const Scene& s, const Ray& r, float tmin, float tmax, bool brute, uint64_t* counts,
                 double time = 0.0) {
    Hit best{tmax, kMiss, 0.0f, 0.0f};
    const RayPre p = precompute(r);
    float t, u, v;
    if (brute) {
        for (uint32_t i = 0; i < s.numStatic; ++i)
            if (intersectTri(r, p, s.tris[i], t, u, v)) consider(best, t, i, u, v, tmin);
    } else if (s.wideWidth) {
        if (s.numStatic && !s.wideTris.empty()) traverseWide(s, r, tmin, tmax, false, best, counts);
    } else if (!s.nodes.empty()) {
        uint32_t stack[128];
        int sp = 0;
        stack[sp++] = 0;
        while (sp) {
            const BNode& n = s.nodes[stack[--sp]];
            if (counts) ++counts[0];
            if (!boxHit(n, r, p, tmin, best.t * 1.00001f)) continue;
            if (n.count) {
                for (uint32_t i = n.left; i < n.left + n.count; ++i) {
                    const uint32_t ti = s.order[i];
                    if (counts) ++counts[1];
                    if (intersectTri(r, p, s.tris[ti], t, u, v)) consider(best, t, ti, u, v, tmin);
                }
            } else {
                stack[sp++] = n.left + 1;
                stack[sp++] = n.left;
            }
        }
    }
    instanceHits(s, r, time, tmin, tmax, false, best);
    // light quads after triangles; a quad must be strictly closer to win
    for (uint32_t q = 0; q < s.quads.size(); ++q)
        if (intersectQuad(r, s.quads[q], tmin, best.t, t)) best = {t, kQuadBit | q, 0.0f, 0.0f};
    if (best.prim == kMiss) best.t = tmax;
    return best;
}

// Traversal of the wide BVH in the order the GPU kernel uses (pg_device.hpp traverse()): slab
// distances as fma(plane, 1/d, -o/d), the nearest hit child entered first, the other hit
// children pushed in child order, popped last-in first-out and skipped when their entry distance
// exceeds the best hit by more than 1e-5 relative; leaves test their triangles in order. So the
// nodes visited and triangles tested are the GPU's own counts. any = true: return at the first
// triangle hit in (tmin, tmax).
constexpr uint32_t kWideLeaf = 0x80000000u, kWideCountShift = 27, kWideFirstMask = (1u << 27) - 1u;
constexpr uint32_t kWideInvalid = 0xFFFFFFFFu;
bool traverseWide(const Scene& s, const Ray& r, float tmin, float tmax, bool any, Hit& h, uint64_t* counts) {
    const uint32_t W = s.wideWidth;
    const RayPre p = precompute(r);
    const float ix = p.inv[0], iy = p.inv[1], iz = p.inv[2];
    const float oix = r.o.x * ix, oiy = r.o.y * iy, oiz = r.o.z * iz;
    uint32_t hprim = kMiss;  // original index of the current best (tie rule)
    std::vector<std::pair<uint32_t, float>> stack;
    uint32_t ref = 0;
    while (true) {
        if (!(ref & kWideLeaf)) {
            if (counts) ++counts[0];
            const uint32_t* n = &s.wideNodes[static_cast<size_t>(ref) * 7 * W];
            auto plane = [&](int k, uint32_t c) {
                float f;
                std::memcpy(&f, &n[k * W + c], 4);
                return f;
            };
            const float tcull = any ? tmax : h.t * 1.00001f;
            float tn[8];
            uint32_t mask = 0;
            for (uint32_t c = 0; c < W; ++c) {
                const float ax = std::fma(plane(0, c), ix, -oix), bx = std::fma(plane(3, c), ix, -oix);
                const float ay = std::fma(plane(1, c), iy, -oiy), by = std::fma(plane(4, c), iy, -oiy);
                const float az = std::fma(plane(2, c), iz, -oiz), bz = std::fma(plane(5, c), iz, -oiz);
                const float t0 = std::fmax(std::fmax(std::fmin(ax, bx), std::fmin(ay, by)), std::fmax(std::fmin(az, bz), tmin));
                const float t1 = std::fmin(std::fmin(std::fmax(ax, bx), std::fmax(ay, by)),
                                           std::fmin(std::fmax(az, bz), tcull)) * 1.0000004f;
                tn[c] = t0;
                if (n[6 * W + c] != kWideInvalid && t0 <= t1) mask |= 1u << c;
            }
            if (mask) {
                uint32_t best = 0;
                float bt = INFINITY;
                for (uint32_t c = 0; c < W; ++c)
                    if (((mask >> c) & 1) && tn[c] < bt) bt = tn[c], best = c;
                for (uint32_t c = 0; c < W; ++c)
                    if (((mask >> c) & 1) && c != best) stack.push_back({n[6 * W + c], tn[c]});
                ref = n[6 * W + best];
                continue;
            }
        } else {
            const uint32_t first = ref & kWideFirstMask, count = ((ref >> kWideCountShift) & 0xFu) + 1u;
            for (uint32_t k = first; k < first + count; ++k) {
                if (counts) ++counts[1];
                const float* q = &s.wideTris[static_cast<size_t>(k) * 12];
                const Tri tri{{q[0], q[1], q[2]}, {q[4], q[5], q[6]}, {q[8], q[9], q[10]}};
                float t, u, v;
                if (!intersectTri(r, p, tri, t, u, v)) continue;
                if (any) {
                    if (t > tmin && t < tmax) return true;
                } else {
                    uint32_t prim;
                    std::memcpy(&prim, &q[3], 4);
                    if (t > tmin && (t < h.t || (t == h.t && prim < hprim))) h = {t, prim, u, v}, hprim = prim;
                }
            }
        }
        bool found = false;
        while (!stack.empty()) {
            const auto e = stack.back();
            stack.pop_back();
            if (any || e.second <= h.t * 1.00001f) {
                ref = e.first;
                found = true;
                break;
            }
        }
        if (!found) break;
    }
    return !any && h.prim != kMiss;
}

// rtcOccluded1 with mask gsm only (Embree.cpp:357-391): lights never occlude.
bool occluded(const Scene& s, const Ray& r, float tmin, float tmax, double time = 0.0) {
    Hit ih{tmax, kMiss, 0.0f, 0.0f};
    if (instanceHits(s, r, time, tmin, tmax, true, ih)) return true;
    if (s.wideWidth) {
        Hit h{tmax, kMiss, 0.0f, 0.0f};
        return s.numStatic && !s.wideTris.empty() && traverseWide(s, r, tmin, tmax, true, h, nullptr);
    }
    if (s.nodes.empty()) return false;
    const RayPre p = precompute(r);
    float t, u, v;
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const BNode& n = s.nodes[stack[--sp]];
        if (!boxHit(n, r, p, tmin, tmax * 1.00001f)) continue;
        if (n.count) {
            for (uint32_t i = n.left; i < n.left + n.count; ++i)
                if (intersectTri(r, p, s.tris[s.order[i]], t, u, v) && t > tmin && t < tmax) return true;
        } else {
            stack[sp++] = n.left + 1;
            stack[sp++] = n.left;
        }
    }
    return false;
}

// ---------------------------------------------------------------------------------------------
// Sampling helpers (Kernel/Sampling.hpp:23-68)
// ---------------------------------------------------------------------------------------------
void concentricDisk(float u1, float u2, float& x, float& y) {  // Sampling.hpp:31-42
    u1 = 2.0f * u1 - 1.0f;
    u2 = 2.0f * u2 - 1.0f;
    const float au1 = std::fabs(u1), au2 = std::fabs(u2);
    if (std::fmin(au1, au2) < 1e-4f) {
        x = y = 0.0f;
        return;
    }
    const float rad = au1 > au2 ? u1 : u2;
    const float theta = au1 > au2 ? (kPi / 4.0f) * (u2 / u1) : (kPi / 2.0f) - (kPi / 4.0f) * (u1 / u2);
    x = std::cos(theta) * rad;
    y = std::sin(theta) * rad;
}
V3 cosineHemisphere(float u1, float u2) {  // Sampling.hpp:44-50
    float x, y;
    concentricDisk(u1, u2, x, y);
    const float z = std::sqrt(std::fmax(0.0f, 1.0f - (x * x + y * y)));
    return {x, y, z};
}

// ---------------------------------------------------------------------------------------------
// BxDFs (Surface/Classic/Kernel.cpp). Shading space: z = shading normal.
// ---------------------------------------------------------------------------------------------
enum : uint32_t { kPartReflection = 1, kPartTransmission = 2, kPartSpecular = 4 };  // BxDFPart subset
struct BsdfSample {
    V3 wi, f;
    float pdf;
    uint32_t part = kPartReflection;
};

// GGX (Trowbridge-Reitz) distribution, Classic/Kernel.cpp:289-382 (MicrofacetDistribution with
// alphaX, alphaY; the conductor and plastic pass one alpha twice). The azimuth's cosine and sine come
// from w over sin(theta), clamped to [-1, 1], as the reference computes them.
struct Dist {
    float ax, ay;
};
Dist dist(const Material& m) { return {m.alpha, m.alphaY}; }
void azimuth(V3 w, float sin2, float& cosPhi, float& sinPhi) {
    const float st = std::sqrt(sin2);
    cosPhi = st == 0.0f ? 1.0f : std::fmin(1.0f, std::fmax(-1.0f, w.x / st));
    sinPhi = st == 0.0f ? 0.0f : std::fmin(1.0f, std::fmax(-1.0f, w.y / st));
}
float ggxLambda(V3 w, Dist d) {  // Kernel.cpp:326-341
    const float cos2 = w.z * w.z;
    const float sin2 = std::fmax(0.0f, 1.0f - cos2);
    const float tan2 = sin2 / cos2;
    if (std::isinf(tan2)) return 0.0f;
    float cp, sp;
    azimuth(w, sin2, cp, sp);
    const float p1 = cp * d.ax, p2 = sp * d.ay;
    return (-1.0f + std::sqrt(1.0f + (p1 * p1 + p2 * p2) * tan2)) * 0.5f;
}
float ggxD(V3 wh, Dist d) {  // Kernel.cpp:342-357
    const float cos2 = wh.z * wh.z;
    const float sin2 = std::fmax(0.0f, 1.0f - cos2);
    const float tan2 = sin2 / cos2;
    if (std::isinf(tan2)) return 0.0f;
    float cp, sp;
    azimuth(wh, sin2, cp, sp);
    const float p1 = cp / d.ax, p2 = sp / d.ay;
    const float e = cos2 * (1.0f + (p1 * p1 + p2 * p2) * tan2);
    return kInvPi / (d.ax * d.ay * e * e);
}
// Visible-normal sampling (Heitz, "Sampling the GGX Distribution of Visible Normals", JCGT 2018):
// the view direction stretched by (ax, ay), a point of the projected hemisphere, unstretched.
V3 ggxSampleWh(V3 wo, Dist d, float u1, float u2) {
    const V3 vh = normalize(V3{d.ax * wo.x, d.ay * wo.y, wo.z});
    const float lensq = vh.x * vh.x + vh.y * vh.y;
    V3 t1 = lensq > 0.0f ? V3{-vh.y, vh.x, 0.0f} * (1.0f / std::sqrt(lensq)) : V3{1.0f, 0.0f, 0.0f};
    const V3 t2 = cross(vh, t1);
    const float r = std::sqrt(u1);
    const float phi = 2.0f * kPi * u2;
    const float p1 = r * std::cos(phi);
    float p2 = r * std::sin(phi);
    const float sm = 0.5f * (1.0f + vh.z);
    p2 = (1.0f - sm) * std::sqrt(std::fmax(0.0f, 1.0f - p1 * p1)) + sm * p2;
    const V3 nh = t1 * p1 + t2 * p2 + vh * std::sqrt(std::fmax(0.0f, 1.0f - p1 * p1 - p2 * p2));
    return normalize(V3{d.ax * nh.x, d.ay * nh.y, std::fmax(0.0f, nh.z)});
}
V3 schlick(V3 f0, float c) {
    const float m = 1.0f - std::fmin(1.0f, std::fmax(0.0f, c));
    const float m5 = (m * m) * (m * m) * m;
    return V3{f0.x + (1.0f - f0.x) * m5, f0.y + (1.0f - f0.y) * m5, f0.z + (1.0f - f0.z) * m5};
}
// MicrofacetReflection::evaluate, Kernel.cpp:398-404, with a Schlick conductor Fresnel.
V3 conductorF(const Material& m, V3 wo, V3 wi) {
    if (wo.z <= 0.0f || wi.z <= 0.0f) return {0.0f, 0.0f, 0.0f};
    const V3 wh = normalize(wo + wi);
    const float g = 1.0f + ggxLambda(wo, dist(m)) + ggxLambda(wi, dist(m));
    const float k = ggxD(wh, dist(m)) / (4.0f * g * std::fabs(wo.z * wi.z));
    return schlick(m.albedo, dot(wi, wh)) * k;
}
float conductorPdf(const Material& m, V3 wo, V3 wi) {  // D(wh) G1(wo) / (4 cos(theta_o))
    if (wi.z <= 0.0f || wo.z <= 0.0f) return 0.0f;
    const V3 wh = normalize(wo + wi);
    return ggxD(wh, dist(m)) / (4.0f * wo.z * (1.0f + ggxLambda(wo, dist(m))));
}

// MicrofacetDistribution::remapRoughness, Classic/Kernel.cpp:378-381 (a quartic in log roughness).
float remapRoughness(float roughness) {
    const float x = std::log(std::fmax(1e-3f, roughness));
    const float c[5] = {1.62142f, 0.819955f, 0.1734f, 0.0171201f, 0.000640711f};
    float r = c[4];
    for (int k = 3; k >= 0; --k) r = c[k] + r * x;
    return std::fmax(1e-3f, r);
}
// FresnelDielectric::operator(), Classic/Kernel.cpp:271-286 (unpolarised, Snell's law; total
// internal reflection gives 1).
float fresnelDielectric(float cosI, float etaI, float etaT) {
    cosI = std::fmin(1.0f, std::fmax(0.0f, std::fabs(cosI)));
    const float sinT = etaI / etaT * std::sqrt(std::fmax(0.0f, 1.0f - cosI * cosI));
    if (sinT >= 1.0f) return 1.0f;
    const float cosT = std::sqrt(std::fmax(0.0f, 1.0f - sinT * sinT));
    const float rPar = (etaT * cosI - etaI * cosT) / (etaT * cosI + etaI * cosT);
    const float rPerp = (etaI * cosI - etaT * cosT) / (etaI * cosI + etaT * cosT);
    return 0.5f * (rPar * rPar + rPerp * rPerp);
}

// One BxDF of a material. Kinds: the Lambert lobe (LambertianReflection, Kernel.cpp:81-92), the
// GGX lobe with a Schlick conductor Fresnel, and the GGX lobe with the dielectric Fresnel of
// plasticInit (Kernel.cpp:441-471: eta 1.46 outside-in on the front face, swapped on the back).
enum LobeKind : uint32_t { kLobeLambert, kLobeConductor, kLobeDielectric, kLobeOrenNayar, kLobeMirror, kLobeGlass, kLobeBlend,
                           kLobeGlassR, kLobeGlassT };
struct Lobe {
    LobeKind kind;
    V3 scale;
};
struct Lobes {
    Lobe l[2];
    uint32_t n = 0;
};
// The valid lobes in the order of the BSDF's members (PlasticBSDF: diffuse, then glossy);
// plasticInit keeps only lobes whose colour is valid().
Lobes lobesOf(const Material& m) {
    Lobes r;
    if (m.type == kMatLambert) {  // matteInit (Kernel.cpp:249-269): one lobe either way
        r.l[r.n++] = {m.orenNayar ? kLobeOrenNayar : kLobeLambert, m.albedo};
    } else if (m.type == kMatConductor) {
        r.l[r.n++] = {kLobeConductor, m.albedo};
    } else if (m.type == kMatMirror) {  // mirrorInit, Kernel.cpp:667-676
        r.l[r.n++] = {kLobeMirror, m.albedo};
    } else if (m.type == kMatGlass && m.alpha == 0.0f) {  // glassInit, zero roughness: FresnelSpecular, Kernel.cpp:608-640
        if (valid(m.albedo) || valid(m.spec)) r.l[r.n++] = {kLobeGlass, m.albedo};
    } else if (m.type == kMatGlass) {  // rough: microfacet reflection and transmission, Kernel.cpp:641-655
        if (valid(m.albedo)) r.l[r.n++] = {kLobeGlassR, m.albedo};
        if (valid(m.spec)) r.l[r.n++] = {kLobeGlassT, m.spec};
    } else if (m.type == kMatSubstrate) {  // substrateInit, Kernel.cpp:757-777
        if (valid(m.albedo) || valid(m.spec)) r.l[r.n++] = {kLobeBlend, m.albedo};
    } else {
        if (valid(m.albedo)) r.l[r.n++] = {kLobeLambert, m.albedo};
        if (valid(m.spec)) r.l[r.n++] = {kLobeDielectric, m.spec};
    }
    return r;
}
// OrenNayar::evaluate (Kernel.cpp:108-131), in terms of the polar and azimuth angles of wo and wi:
// (A + B max(0, cos(phi_i - phi_o)) sin(alpha) tan(beta)) / pi with alpha the larger polar angle and
// beta the smaller; zero within sin(theta) < 1e-4 of the normal, as the reference has it.
float orenNayarScale(const Material& m, V3 wo, V3 wi) {
    const float sI = std::sqrt(std::fmax(0.0f, 1.0f - wi.z * wi.z));
    const float sO = std::sqrt(std::fmax(0.0f, 1.0f - wo.z * wo.z));
    if (sI < 1e-4f || sO < 1e-4f) return 0.0f;
    auto clamp1 = [](float x) { return std::fmin(1.0f, std::fmax(-1.0f, x)); };
    const float cosDphi = clamp1(wi.x / sI) * clamp1(wo.x / sO) + clamp1(wi.y / sI) * clamp1(wo.y / sO);
    const float cI = std::fabs(wi.z), cO = std::fabs(wo.z);
    const float sinAlpha = cI > cO ? sO : sI;       // alpha = max(theta_i, theta_o)
    const float tanBeta = cI > cO ? sI / cI : sO / cO;  // beta = min(theta_i, theta_o)
    return kInvPi * (m.onA + m.onB * std::fmax(0.0f, cosDphi) * sinAlpha * tanBeta);
}
// FresnelBlend (Kernel.cpp:694-751): Ashikhmin-Shirley's diffuse term and a Schlick-weighted GGX
// term, diffuse = scale, specular = m.spec.
float fifth(float v) { return v * v * v * v * v; }
V3 blendF(const Material& m, const V3& rd, V3 wo, V3 wi) {
    const V3 rs = m.spec;
    const V3 oneMinus{1.0f - rs.x, 1.0f - rs.y, 1.0f - rs.z};
    const float shape = (28.0f / (23.0f * kPi)) * (1.0f - fifth(1.0f - 0.5f * std::fabs(wi.z))) *
                        (1.0f - fifth(1.0f - 0.5f * std::fabs(wo.z)));
    const V3 wh = normalize(wo + wi);
    const float c = dot(wi, wh);
    const float tail = fifth(1.0f - c);  // Schlick: rs + (1 - rs) (1 - cos)^5
    const V3 fres{rs.x + oneMinus.x * tail, rs.y + oneMinus.y * tail, rs.z + oneMinus.z * tail};
    const float spec = ggxD(wh, dist(m)) / (4.0f * std::fmax(std::fabs(wi.z), std::fabs(wo.z)) * std::fabs(c));
    return mul(rd, oneMinus) * shape + fres * spec;
}
// Half cosine, half visible-normal GGX (the reference's pdf divides by 1 + Lambda(wo) + Lambda(wh),
// which is not the density its sampler draws from).
float blendPdf(const Material& m, V3 wo, V3 wi) {
    if (wi.z <= 0.0f) return 0.0f;
    const V3 wh = normalize(wo + wi);
    return 0.5f * (std::fabs(wi.z) * kInvPi + ggxD(wh, dist(m)) / (4.0f * std::fabs(wo.z) * (1.0f + ggxLambda(wo, dist(m)))));
}
// Rough glass. Reflection: the GGX lobe with the dielectric Fresnel of the face (eta 1.5).
// Transmission: Walter et al. 2007's microfacet BTDF with the generalized half vector
// h = -(eta_i wo + eta_t wi) (made to face +z; zero unless wo.h > 0 > wi.h, their chi+ terms),
// carrying radiance, so scaled by (eta_i / eta_t)^2:
// f = |wi.h| |wo.h| (1 - F(wo.h)) D(h) G / (|wi.z| |wo.z| (wo.h + (eta_t/eta_i) wi.h)^2), and a pdf
// of the visible normal h times the Jacobian (eta_t/eta_i)^2 |wi.h| / (wo.h + (eta_t/eta_i) wi.h)^2.
float glassEta(const Material& m) { return m.front ? 1.5f : 1.0f / 1.5f; }  // eta_t / eta_i
V3 glassHalf(const Material& m, V3 wo, V3 wi) {
    V3 h = normalize(wo + wi * glassEta(m));
    return h.z < 0.0f ? -h : h;
}
float glassTPdf(const Material& m, V3 wo, V3 wi) {
    if (wo.z <= 0.0f || wi.z >= 0.0f) return 0.0f;
    const V3 h = glassHalf(m, wo, wi);
    const float ow = dot(wo, h), iw = dot(wi, h);
    if (!(ow > 0.0f && iw < 0.0f)) return 0.0f;
    const float e = glassEta(m), den = ow + e * iw;
    const float visible = ggxD(h, dist(m)) * std::fabs(ow) / (wo.z * (1.0f + ggxLambda(wo, dist(m))));
    return visible * std::fabs(e * e * iw / (den * den));
}
V3 lobeF(const Material& m, const Lobe& b, V3 wo, V3 wi) {
    if (b.kind == kLobeLambert) return b.scale * kInvPi;
    if (b.kind == kLobeGlassT) {
        if (wo.z <= 0.0f || wi.z >= 0.0f) return {0.0f, 0.0f, 0.0f};
        const V3 h = glassHalf(m, wo, wi);
        const float ow = dot(wo, h), iw = dot(wi, h);
        if (!(ow > 0.0f && iw < 0.0f)) return {0.0f, 0.0f, 0.0f};  // Walter's chi+(i.h / i.n)
        const float fr = m.front ? fresnelDielectric(ow, 1.0f, 1.5f) : fresnelDielectric(ow, 1.5f, 1.0f);
        const float den = ow + glassEta(m) * iw;
        const float g = 1.0f / (1.0f + ggxLambda(wo, dist(m)) + ggxLambda(wi, dist(m)));
        return b.scale * ((1.0f - fr) * ggxD(h, dist(m)) * g * std::fabs(iw) * std::fabs(ow) /
                          (std::fabs(wo.z * wi.z) * den * den));
    }
    if (b.kind == kLobeGlassR) {
        if (wo.z <= 0.0f || wi.z <= 0.0f) return {0.0f, 0.0f, 0.0f};
        const V3 wh = normalize(wo + wi);
        const float g = 1.0f + ggxLambda(wo, dist(m)) + ggxLambda(wi, dist(m));
        const float k = ggxD(wh, dist(m)) / (4.0f * g * std::fabs(wo.z * wi.z));
        const float fr = m.front ? fresnelDielectric(dot(wi, wh), 1.0f, 1.5f) : fresnelDielectric(dot(wi, wh), 1.5f, 1.0f);
        return b.scale * (fr * k);
    }
    if (b.kind == kLobeMirror || b.kind == kLobeGlass) return {0.0f, 0.0f, 0.0f};  // delta lobes evaluate to 0
    if (b.kind == kLobeBlend) return blendF(m, b.scale, wo, wi);
    if (b.kind == kLobeOrenNayar) return b.scale * orenNayarScale(m, wo, wi);
    if (wo.z <= 0.0f || wi.z <= 0.0f) return {0.0f, 0.0f, 0.0f};
    if (b.kind == kLobeConductor) return conductorF(m, wo, wi);
    const V3 wh = normalize(wo + wi);  // MicrofacetReflection::evaluate, Kernel.cpp:398-404
    const float g = 1.0f + ggxLambda(wo, dist(m)) + ggxLambda(wi, dist(m));
    const float k = ggxD(wh, dist(m)) / (4.0f * g * std::fabs(wo.z * wi.z));
    const float fr = m.front ? fresnelDielectric(dot(wi, wh), 1.0f, 1.46f) : fresnelDielectric(dot(wi, wh), 1.46f, 1.0f);
    return b.scale * (fr * k);
}
bool lobeTransmits(const Lobe& b) { return b.kind == kLobeGlassT; }
float lobePdf(const Material& m, const Lobe& b, V3 wo, V3 wi) {
    if (b.kind == kLobeMirror || b.kind == kLobeGlass) return 0.0f;
    if (b.kind == kLobeGlassT) return glassTPdf(m, wo, wi);
    if (b.kind == kLobeBlend) return blendPdf(m, wo, wi);
    if (b.kind == kLobeLambert || b.kind == kLobeOrenNayar) return wi.z > 0.0f ? std::fabs(wi.z) * kInvPi : 0.0f;  // Kernel.cpp:76-78
    return conductorPdf(m, wo, wi);  // the same visible-normal pdf for either Fresnel
}
BsdfSample lobeSample(const Material& m, const Lobe& b, V3 wo, float u1, float u2) {
    BsdfSample s{{0, 0, 0}, {0, 0, 0}, 0.0f};
    if (b.kind == kLobeMirror) {  // mirrorSample, Kernel.cpp:678-692: pdf 1, f = R / |cos|
        s.wi = V3{-wo.x, -wo.y, wo.z};
        s.f = b.scale * (1.0f / std::fabs(s.wi.z));
        s.pdf = 1.0f;
        s.part = kPartReflection | kPartSpecular;
        return s;
    }
    if (b.kind == kLobeGlass) {  // FresnelSpecular::sample, Kernel.cpp:522-540 (eta 1.5, face-swapped)
        const float eI = m.front ? 1.0f : 1.5f, eT = m.front ? 1.5f : 1.0f;
        const float fr = fresnelDielectric(wo.z, eI, eT);
        if (u1 < fr) {
            s.wi = V3{-wo.x, -wo.y, wo.z};
            s.pdf = fr;
            s.f = b.scale * (fr / std::fabs(s.wi.z));
            s.part = kPartReflection | kPartSpecular;
            return s;
        }
        // Snell's law about +z with eta = eI / eT (refract, Transform.hpp:410-421)
        const float eta = eI / eT;
        const float sin2T = eta * eta * std::fmax(0.0f, 1.0f - wo.z * wo.z);
        if (sin2T >= 1.0f) return s;
        const float cosT = std::sqrt(std::fmax(0.0f, 1.0f - sin2T));
        s.wi = normalize(V3{-eta * wo.x, -eta * wo.y, (eta * wo.z - cosT) - eta * wo.z});
        s.pdf = 1.0f - fr;
        s.f = m.spec * (eta * eta * s.pdf / std::fabs(s.wi.z));  // radiance scaled by (eI / eT)^2
        s.part = kPartTransmission | kPartSpecular;
        return s;
    }
    if (b.kind == kLobeGlassT) {  // a visible microfacet normal h, then Snell's law about h
        if (wo.z <= 0.0f) return s;
        const V3 h = ggxSampleWh(wo, dist(m), u1, u2);
        const float c = dot(wo, h);
        if (c < 0.0f) return s;
        const float eta = 1.0f / glassEta(m);  // eta_i / eta_t
        const float sin2T = eta * eta * std::fmax(0.0f, 1.0f - c * c);
        if (sin2T >= 1.0f) return s;
        const float cosT = std::sqrt(std::fmax(0.0f, 1.0f - sin2T));
        s.wi = normalize(h * (eta * c - cosT) - wo * eta);
        s.pdf = glassTPdf(m, wo, s.wi);
        if (s.pdf > 0.0f) s.f = lobeF(m, b, wo, s.wi);
        s.part = kPartTransmission;
        return s;
    }
    if (b.kind == kLobeBlend) {  // FresnelBlend::sample, Kernel.cpp:734-750
        if (u1 < 0.5f) {
            s.wi = cosineHemisphere(std::fmin(2.0f * u1, 1.0f), u2);
        } else {
            const V3 wh = ggxSampleWh(wo, dist(m), std::fmin(2.0f * (u1 - 0.5f), 1.0f), u2);
            s.wi = wh * (2.0f * dot(wo, wh)) - wo;
            if (s.wi.z <= 0.0f) return s;
        }
        s.pdf = blendPdf(m, wo, s.wi);
        s.f = blendF(m, b.scale, wo, s.wi);
        return s;
    }
    if (b.kind == kLobeLambert || b.kind == kLobeOrenNayar) {  // DefaultBxDFHelper::sample, Kernel.cpp:67-74
        s.wi = cosineHemisphere(u1, u2);
        s.f = lobeF(m, b, wo, s.wi);
        s.pdf = s.wi.z > 0.0f ? std::fabs(s.wi.z) * kInvPi : 0.0f;
        return s;
    }
    if (wo.z <= 0.0f) return s;  // MicrofacetReflection::sample, Kernel.cpp:413-433
    const V3 wh = ggxSampleWh(wo, dist(m), u1, u2);
    const float c = dot(wo, wh);
    if (c < 0.0f) return s;
    s.wi = wh * (2.0f * c) - wo;  // reflect, Transform.hpp:405-408
    if (s.wi.z <= 0.0f) return s;
    s.f = lobeF(m, b, wo, s.wi);
    s.pdf = lobePdf(m, b, wo, s.wi);
    return s;
}

// Reflection/transmission side test of the evaluate path, Classic/Kernel.cpp:220 and :211-215;
// every lobe here is a reflection lobe, so a pair on opposite sides of Ng evaluates to zero.
inline bool sameSide(V3 wo, V3 wi, V3 ng) { return dot(wo, ng) * dot(wi, ng) > 0.0f; }

// The mixture over a material's lobes (Classic/Kernel.cpp:144-237): evaluate sums the lobes,
// pdf is their mean, and sample draws lobe floor(u1 count), rescales u1, then adds the other
// lobes' values (same side only) and pdfs and divides the pdf by the count.
V3 bsdfEval(const Material& m, V3 wo, V3 wi, V3 ng) {
    const bool same = sameSide(wo, wi, ng);  // addition: Reflection on the same side, else Transmission
    const Lobes ls = lobesOf(m);
    V3 f{0.0f, 0.0f, 0.0f};
    for (uint32_t k = 0; k < ls.n; ++k)
        if (lobeTransmits(ls.l[k]) != same) f = f + lobeF(m, ls.l[k], wo, wi);
    return f;
}
float bsdfPdf(const Material& m, V3 wo, V3 wi) {
    const Lobes ls = lobesOf(m);
    float p = 0.0f;
    for (uint32_t k = 0; k < ls.n; ++k) p += lobePdf(m, ls.l[k], wo, wi);
    return ls.n > 1 ? p / static_cast<float>(ls.n) : p;
}
BsdfSample bsdfSample(const Material& m, V3 wo, V3 ng, float u1, float u2) {
    const Lobes ls = lobesOf(m);
    if (ls.n == 0) return BsdfSample{{0, 0, 0}, {0, 0, 0}, 0.0f};
    const float cnt = static_cast<float>(ls.n);
    const uint32_t pick = std::min(static_cast<uint32_t>(std::floor(u1 * cnt)), ls.n - 1);
    BsdfSample s = lobeSample(m, ls.l[pick], wo, u1 * cnt - static_cast<float>(pick), u2);
    if (s.pdf <= 0.0f) return s;
    const bool same = sameSide(wo, s.wi, ng);
    for (uint32_t k = 0; k < ls.n && !(s.part & kPartSpecular); ++k) {
        if (k == pick) continue;
        if (lobeTransmits(ls.l[k]) != same) s.f = s.f + lobeF(m, ls.l[k], wo, s.wi);
        s.pdf += lobePdf(m, ls.l[k], wo, s.wi);
    }
    if (ls.n > 1) s.pdf = s.pdf / cnt;
    return s;
}

// Texture lookups. The checkerboard folds x * scale into [0, 2) with an IEEE remainder; the image
// takes the texels left and right of p - 1/2 with wrap, weighted by the distance of p to the texel
// centre, and scales the 8-bit values by 1/255.
std::array<float, 4> sampleTexture(const Texture& t, float u, float v) {
    std::array<float, 4> c{0.0f, 0.0f, 0.0f, 0.0f};
    if (t.type == 1) {
        auto fold = [&](float x) {
            const float r = std::remainder(x * t.scale, 2.0f);
            return r < 0.0f ? r + 2.0f : r;
        };
        const float* src = (fold(u) > 1.0f) == (fold(v) > 1.0f) ? t.white : t.black;
        for (uint32_t i = 0; i < t.channels; ++i) c[i] = src[i];
        return c;
    }
    auto wrapIndex = [&](float p, uint32_t size) {
        const float r = std::remainder(p, static_cast<float>(size));
        const float q = t.wrap == 0 ? (r < 0.0f ? r + static_cast<float>(size) : r) : std::fabs(r);
        return std::min(static_cast<uint32_t>(q), size - 1);
    };
    auto step = [&](uint32_t i, uint32_t size) { return i + 1 == size ? (t.wrap == 0 ? 0u : i) : i + 1; };
    const float pu = u * static_cast<float>(t.width), pv = v * static_cast<float>(t.height);
    const uint32_t x0 = wrapIndex(pu - 0.5f, t.width), x1 = step(x0, t.width);
    const uint32_t y0 = wrapIndex(pv - 0.5f, t.height), y1 = step(y0, t.height);
    const float ax = 0.5f + std::floor(pu + 0.5f) - pu, ay = 0.5f + std::floor(pv + 0.5f) - pv;
    const float bx = 1.0f - ax, by = 1.0f - ay;
    const uint32_t xs[4] = {x0, x1, x0, x1}, ys[4] = {y0, y0, y1, y1};
    const float ws[4] = {ax * ay, bx * ay, ax * by, bx * by};
    for (int k = 0; k < 4; ++k) {
        const uint8_t* px = t.texels.data() + (static_cast<size_t>(ys[k]) * t.width + xs[k]) * t.channels;
        for (uint32_t i = 0; i < t.channels; ++i) c[i] += ws[k] * static_cast<float>(px[i]);
    }
    for (uint32_t i = 0; i < t.channels; ++i) c[i] *= 0.003921568f;
    return c;
}
// The material at a triangle hit with its textures applied at the hit's texture coordinate
// (corner 1 weighted by u, corner 2 by v, corner 0 by 1 - u - v).
void applyTextures(const Scene& s, uint32_t matId, const Hit& h, Material& m) {
    static const std::array<uint32_t, 3> kNone{~0u, ~0u, ~0u};
    const auto& b = matId < s.matTex.size() ? s.matTex[matId] : kNone;
    if (b[0] == ~0u && b[1] == ~0u && b[2] == ~0u && m.texY == ~0u) return;
    float tu = 0.0f, tv = 0.0f;
    if (!s.triUV.empty()) {
        const float* c = &s.triUV[6 * static_cast<size_t>(h.prim)];
        const float w = 1.0f - h.u - h.v;
        tu = c[2] * h.u + c[4] * h.v + c[0] * w;
        tv = c[3] * h.u + c[5] * h.v + c[1] * w;
    }
    if (b[0] != ~0u) {
        const auto c = sampleTexture(s.textures[b[0]], tu, tv);
        m.albedo = {c[0], c[1], c[2]};
    }
    if (b[1] != ~0u) {
        const auto c = sampleTexture(s.textures[b[1]], tu, tv);
        m.spec = {c[0], c[1], c[2]};
    }
    if (b[2] != ~0u || m.texY != ~0u) {  // the roughness at this hit, read as each surface's init reads its config
        const float r = b[2] != ~0u ? sampleTexture(s.textures[b[2]], tu, tv)[0] : m.roughX;
        if (m.type == kMatLambert) {
            m.orenNayar = r > 0.0f;
            if (m.orenNayar) {
                const float sigma = r * kPi, s2 = sigma * sigma;
                m.onA = 1.0f - s2 / (2.0f * (s2 + 0.33f));
                m.onB = 0.45f * s2 / (s2 + 0.09f);
            }
        } else if (m.type == kMatConductor) {
            m.alpha = m.alphaY = std::fmax(r * r, 1e-7f);
        } else if (m.type == kMatPlastic) {
            m.alpha = m.alphaY = remapRoughness(r);
        } else if (m.type == kMatGlass || m.type == kMatSubstrate) {
            // RoughnessY from its own texture or constant; glassInit (Kernel.cpp:630-655) is smooth only
            // when both are zero, substrateInit (:769-773) always takes the remapped pair
            const float ry = m.texY != ~0u ? sampleTexture(s.textures[m.texY], tu, tv)[0] : m.roughY;
            const bool rough = m.type == kMatSubstrate || r > 0.0f || ry > 0.0f;
            m.alpha = rough ? remapRoughness(r) : 0.0f;
            m.alphaY = rough ? remapRoughness(ry) : 0.0f;
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Integrator (Integrator/PathTracing/Kernel.cpp) and per-sample driver (Tracer/Embree/Kernel.cpp)
// ---------------------------------------------------------------------------------------------
enum : uint32_t { OR_COMPAT_D1 = 1, OR_COMPAT_D3 = 2, OR_COMPAT_D9 = 4, OR_COMPAT_D13 = 8 };

struct Params {
    uint32_t width, height, spp, maxDepth, minRR, flags;
    uint32_t rectX, rectY, rectW, rectH;
    uint32_t sppBegin, sppEnd;
    float ndc[4];  // SensorNDCAffineTransform {ox, oy, sx, sy}
    uint32_t filter;        // 0 box, 1 triangle, 2 gaussian (FixedSampler/Kernel.cpp:55-70)
    float negAlpha, sub;    // GaussianFilterData (FixedSampler.cpp:40-44): -Alpha, exp(-Alpha)
};

struct Sampler {  // PerSampleContext dimension counter, Embree/Kernel.cpp:37-45
    uint64_t index;
    uint32_t dim;
    float time = 0.0f;  // PerSampleContext::time: the sample's dimension 2 (Embree/Kernel.cpp:77)
    float next() { return sobolSample(index, dim++); }
    void next2(bool d9, float& u1, float& u2) {
        const float a = next(), b = next();
        if (d9) u2 = a, u1 = b;
        else u1 = a, u2 = b;
    }
};

struct Frame {  // SurfaceIntersectionInfo, Protocol.hpp:47-66
    V3 t, b, n;
    V3 toLocal(V3 v) const { return {dot(t, v), dot(b, v), dot(n, v)}; }
    V3 toWorld(V3 v) const {
        return {t.x * v.x + b.x * v.y + n.x * v.z, t.y * v.x + b.y * v.y + n.y * v.z, t.z * v.x + b.z * v.y + n.z * v.z};
    }
};

// Hit post-process: calcTriangleMeshSurface (Embree/Kernel.cpp:170-192) and planeSurface
// (BasicShapes/Kernel.cpp:99-111). Returns the world normal facing the incoming ray.
V3 triNg(const Tri& tri) { return normalize(cross(tri.v1 - tri.v0, tri.v2 - tri.v0)); }
V3 hitNormal(const Scene& s, const Hit& h, V3 dir, double time = 0.0) {
    if (h.prim & kQuadBit) {
        const Quad& q = s.quads[h.prim & ~kQuadBit];
        return dot(dir, q.normal) > 0.0f ? -q.normal : q.normal;  // face Front -> -normal
    }
    const V3 ng = triNg(worldTri(s, h.prim, time));
    return dot(dir, ng) < 0.0f ? ng : -ng;  // Embree.cpp:336: Front when dot(dir, Ng) < 0
}
Frame shadingFrame(V3 n, bool d3) {
    Frame f;
    f.n = n;
    const V3 ex{1.0f, 0.0f, 0.0f}, ey{0.0f, 1.0f, 0.0f};
    f.t = std::fabs(dot(n, ex)) < std::fabs(dot(n, ey)) ? cross(n, ex) : cross(n, ey);
    if (!d3) f.t = normalize(f.t);
    f.b = cross(n, f.t);
    return f;
}

// uniformSelect (Light/BasicLight/Kernel.cpp:91-98) over {environment, area lights, point lights}
struct LightSel {
    uint32_t light;  // 0 = environment, 1 + l = area light l, 1 + nl + p = point light p
    float pdf;
    bool delta;      // a point light (LightAttributes::Delta)
};
LightSel selectLight(const Scene& s, float u) {  // uniformSelect, BasicLight/Kernel.cpp:91-98
    const uint32_t nl = static_cast<uint32_t>(s.lights.size());
    const uint32_t count = 1 + nl + static_cast<uint32_t>(s.points.size());
    const uint32_t sel = std::min(static_cast<uint32_t>(std::floor(u * static_cast<float>(count))), count - 1);
    return {sel, 1.0f / static_cast<float>(count), sel > nl};
}

struct LightSample {
    V3 dir, rad;
    float dist, pdf;
};
// select (Kernel/Sampling.hpp:52-63): the first primitive whose CDF entry exceeds u, found by the
// same bisection. u is then rescaled into that primitive's own [0, 1): spec mode subtracts the CDF's
// lower bound (the previous entry), D1 subtracts its upper bound cdf[r] as the reference does.
// u = 1 (possible: the Sobol conversion has no clamp) would index one past the CDF in the
// reference; here it picks the last primitive.
uint32_t selectPrimitive(const Scene& s, const AreaLight& L, float& u, bool d1) {
    uint32_t l = 0, r = L.count;
    while (l < r) {
        const uint32_t mid = (l + r) >> 1;
        if (u >= s.quads[L.first + mid].cdf) l = mid + 1;
        else r = mid;
    }
    r = std::min(r, L.count - 1);
    const Quad& q = s.quads[L.first + r];
    const float lower = d1 ? q.cdf : (r ? s.quads[L.first + r - 1].cdf : 0.0f);
    u = (u - lower) / q.pdf;
    return L.first + r;
}

// diffuseSample (AreaLight/Kernel.cpp:24-41) over planeSample (BasicShapes/Kernel.cpp:113-134);
// SESample (BasicLight/Kernel.cpp:63-76) for the environment.
LightSample sampleLight(const Scene& s, uint32_t light, V3 hit, float u1, float u2, bool d1) {
    LightSample ls{{0, 0, 0}, {0, 0, 0}, 0.0f, 0.0f};
    if (light == 0) {
        const float theta = u1 * kPi, phi = u2 * kPi;
        const float st = std::sin(theta);
        ls.dir = {st * std::cos(phi), st * std::sin(phi), std::cos(theta)};
        ls.rad = s.env;
        ls.pdf = st == 0.0f ? 0.0f : 1.0f / (2.0f * kPi * kPi * st);
        ls.dist = 1e5f;
        return ls;
    }
    if (light > s.lights.size()) {  // pointSample, BasicLight/Kernel.cpp:35-45
        const PointLight& p = s.points[light - 1 - s.lights.size()];
        const V3 delta = p.pos - hit;
        const float dis2 = dot(delta, delta);
        ls.rad = V3{p.intensity.x / dis2, p.intensity.y / dis2, p.intensity.z / dis2};
        ls.dist = std::sqrt(std::fmax(dis2, 0.0f));
        ls.dir = V3{delta.x / ls.dist, delta.y / ls.dist, delta.z / ls.dist};
        ls.pdf = 1.0f;
        return ls;
    }
    float uu = u1;
    const Quad& q = s.quads[selectPrimitive(s, s.lights[light - 1], uu, d1)];
    const V3 src = q.origin + q.u * uu + q.v * u2;
    V3 n = q.normal;
    if (dot(hit - src, n) <= 0.0f) n = -n;
    const float pdfA = 1.0f / q.area;
    const V3 delta = src - hit;
    if (dot(delta, n) >= 0.0f) return ls;
    const float dis2 = dot(delta, delta);
    ls.rad = q.radiance;
    ls.dist = std::sqrt(dis2);
    ls.dir = {delta.x / ls.dist, delta.y / ls.dist, delta.z / ls.dist};
    ls.pdf = pdfA * dis2 / std::fabs(dot(n, ls.dir));
    return ls;
}

inline float powerHeuristic(float a, float b) {  // PathTracing/Kernel.cpp:27-31
    const float sa = a * a, sb = b * b;
    return sa / (sa + sb);
}

struct PathStats {
    uint64_t ext = 0, shadow = 0, mis = 0;
    uint64_t trav[2] = {0, 0};  // extension rays: nodes visited, triangles tested
};

// trace() + multipleImportanceSampling(), PathTracing/Kernel.cpp:23-193
V3 tracePath(const Scene& s, const Params& p, Sampler& smp, Ray ray, PathStats& st) {
    const bool d1 = p.flags & OR_COMPAT_D1, d3 = p.flags & OR_COMPAT_D3, d9 = p.flags & OR_COMPAT_D9,
               d13 = p.flags & OR_COMPAT_D13;
    V3 pf{1.0f, 1.0f, 1.0f}, L{0.0f, 0.0f, 0.0f};
    bool specular = true;
    uint32_t depth = 0;
    while (true) {
        const Hit h = traceClosest(s, ray, 1e-3f, 1e5f, false, st.trav, smp.time);
        ++st.ext;
        const bool isLight = h.prim != kMiss && (h.prim & kQuadBit);
        if (specular && (isLight || h.prim == kMiss)) {  // :119-140
            const V3 rad = isLight ? s.quads[h.prim & ~kQuadBit].radiance : s.env;  // two-sided emission
            L = L + mul(pf, rad);
        }
        if (isLight) {  // :142-146 pass through the light
            if (d13) --depth;
            ray.o = ray.o + ray.d * (h.t + 0.001f);
            continue;
        }
        if (h.prim == kMiss || depth >= p.maxDepth) break;  // :148

        const uint32_t matId = s.objMat[s.triObj[h.prim]];
        Material mat = s.mats[matId];
        applyTextures(s, matId, h, mat);
        const V3 n = hitNormal(s, h, ray.d, smp.time);
        mat.front = dot(ray.d, triNg(worldTri(s, h.prim, smp.time))) < 0.0f;  // Embree.cpp:336
        const Frame fr = shadingFrame(n, d3);
        const V3 wo = fr.toLocal(-ray.d);
        const V3 ng = fr.toLocal(n);
        const V3 hit = ray.o + ray.d * h.t;

        // multipleImportanceSampling, :23-107, only where init left noSpecular set (:158-163)
        if (mat.type != kMatMirror && !(mat.type == kMatGlass && mat.alpha == 0.0f)) {
            const LightSel sel = selectLight(s, smp.next());
            V3 partA{0, 0, 0}, partB{0, 0, 0};
            float u1, u2;
            smp.next2(d9, u1, u2);
            const LightSample ls = sampleLight(s, sel.light, hit, u1, u2, d1);
            if (ls.pdf > 0.0f && valid(ls.rad)) {
                const V3 wi = fr.toLocal(ls.dir);
                const V3 f = bsdfEval(mat, wo, wi, ng);
                if (valid(f)) {
                    ++st.shadow;
                    if (!occluded(s, Ray{hit, ls.dir}, std::fmin(1e-3f, ls.dist), ls.dist, smp.time)) {
                        float w = std::fabs(wi.z) / ls.pdf;
                        if (!sel.delta) w = w * powerHeuristic(ls.pdf, bsdfPdf(mat, wo, wi));  // :51-55
                        partA = mul(f, ls.rad) * w;
                    }
                }
            }
            if (!sel.delta) {  // sampleBSDF returns before drawing for a delta light (:61-62)
                smp.next2(d9, u1, u2);
                const BsdfSample bs = bsdfSample(mat, wo, ng, u1, u2);
                if (bs.pdf > 0.0f && valid(bs.f)) {
                    float weight = std::fabs(bs.wi.z) / bs.pdf;
                    const V3 dir = fr.toWorld(bs.wi);
                    const Hit lh = traceClosest(s, Ray{hit, dir}, 1e-3f, 1e5f, false, nullptr, smp.time);
                    ++st.mis;
                    V3 rad{0, 0, 0};
                    float lpdf = 0.0f;
                    if (lh.prim == kMiss) {  // environment: SEEvaluate / SEPdf (BasicLight/Kernel.cpp:77-89)
                        rad = s.env;
                        const float theta = std::acos(std::fmin(1.0f, std::fmax(-1.0f, dir.z)));
                        const float st2 = std::sin(theta);
                        lpdf = st2 == 0.0f ? 0.0f : 1.0f / (2.0f * kPi * kPi * st2);
                    } else if ((lh.prim & kQuadBit) && sel.light == 1 + s.quads[lh.prim & ~kQuadBit].light) {
                        const Quad& q = s.quads[lh.prim & ~kQuadBit];
                        const V3 qn = hitNormal(s, lh, dir);
                        rad = dot(qn, dir) < 0.0f ? q.radiance : V3{0, 0, 0};  // diffuseEvaluate
                        lpdf = lh.t * lh.t / (std::fabs(dot(qn, dir)) * q.area);  // calcGeometrySamplePdf
                    }
                    if (valid(rad) && lpdf > 0.0f) {
                        weight = weight * powerHeuristic(bs.pdf, lpdf);
                        partB = mul(rad, bs.f) * weight;
                    }
                }
            }
            L = L + mul(pf, (partA + partB) * (1.0f / sel.pdf));
        }

        float u1, u2;
        smp.next2(d9, u1, u2);
        const BsdfSample ss = bsdfSample(mat, wo, ng, u1, u2);  // :164-169
        if (!(ss.pdf > 0.0f && valid(ss.f))) break;
        pf = mul(pf, ss.f * (std::fabs(ss.wi.z) / ss.pdf));
        if (depth > p.minRR) {  // :177-182
            const float q = std::fmax(0.05f, 1.0f - maxc(pf));
            if (smp.next() < q) break;
            pf = pf * (1.0f / (1.0f - q));
        }
        ray.d = fr.toWorld(ss.wi);
        ray.o = hit + n * ((ss.part & kPartReflection) ? 0.001f : -0.001f);  // :184-187
        specular = (ss.part & kPartSpecular) != 0;  // :189
        ++depth;
    }
    return L;
}

// rayGen, Sensor/PerspectiveCamera/Kernel.cpp:22-48 (base = camera position)
Ray cameraRay(const Camera& c, float ndcX, float ndcY, float u1, float u2) {
    const V3 forward = normalize(c.lookAt - c.base);
    const V3 right = cross(forward, c.up);
    const V3 up = cross(right, forward);
    const V3 filmHit = c.base + right * (c.sizeX * (0.5f - ndcX)) + up * (c.sizeY * (ndcY - 0.5f));
    const float focalDistance = dot(c.lookAt - c.base, forward);
    const float filmDistance = 1.0f / (1.0f / c.focal - 1.0f / focalDistance);
    const V3 lensCenter = c.base + forward * filmDistance;
    const float ang = u1 * (2.0f * kPi);  // sampleUniformDisk, Sampling.hpp:23-29
    const float rad = std::sqrt(std::fmax(0.0f, u2));
    const float lx = rad * std::cos(ang), ly = rad * std::sin(ang);
    const V3 lensHit = lensCenter + right * (c.aperture * lx) + up * (c.aperture * ly);
    const V3 dir = lensCenter - filmHit;
    const V3 focusHit = lensCenter + dir * (focalDistance / dot(dir, forward));
    return Ray{lensHit, normalize(focusHit - lensHit)};
}

// One (pixel, sample): piperMain, Tracer/Embree/Kernel.cpp:47-100.
V3 renderSample(const Scene& s, const Params& p, uint32_t m, uint32_t px, uint32_t py, uint32_t si, float& fx,
                float& fy, PathStats& st) {
    Sampler smp{intervalToIndex(m, si, px, py), 5};
    smp.time = sobolSample(smp.index, 2);
    const float res = static_cast<float>(1u << m);
    // sobolStart, Sampler/ScrambledSobolSampler/Kernel.cpp:52-62
    fx = std::fmin(std::fmax(sobolSample(smp.index, 0) * res - static_cast<float>(px), 0.0f), 1.0f);
    fy = std::fmin(std::fmax(sobolSample(smp.index, 1) * res - static_cast<float>(py), 0.0f), 1.0f);
    fx += static_cast<float>(px);
    fy += static_cast<float>(py);
    const float ndcX = p.ndc[0] + p.ndc[2] * fx / static_cast<float>(p.width);
    const float ndcY = p.ndc[1] + p.ndc[3] * fy / static_cast<float>(p.height);
    float u1 = sobolSample(smp.index, 3), u2 = sobolSample(smp.index, 4);
    const Ray ray = cameraRay(s.cam, ndcX, ndcY, u1, u2);
    return tracePath(s, p, smp, ray, st);
}

// Filter weight of a sample at offset (dx, dy) from a pixel centre (FixedSampler/Kernel.cpp:55-70).
float filterWeight(const Params& p, float dx, float dy) {
    if (p.filter == 1) return std::fmax(0.0f, 1.0f - std::fabs(dx)) * std::fmax(0.0f, 1.0f - std::fabs(dy));
    if (p.filter == 2) {
        const auto w = [&](float d) { return std::fmax(std::exp(p.negAlpha * d * d) - p.sub, 0.0f); };
        return w(dx) * w(dy);
    }
    return 1.0f;
}

// accumulate (RenderDriver/FixedSampler/Kernel.cpp:30-52): the sample splats to the 2x2 pixels from
// (trunc(fx - 0.5), trunc(fy - 0.5)) with the filter's weight of its offset from each pixel centre.
template <typename F>
void splat(F* rgbw, const Params& p, float fx, float fy, V3 L, std::vector<std::atomic<uint32_t>>* locks) {
    const int32_t bx = static_cast<int32_t>(fx - 0.5f), by = static_cast<int32_t>(fy - 0.5f);
    for (int k = 0; k < 4; ++k) {
        const uint32_t x = static_cast<uint32_t>(bx + (k & 1)), y = static_cast<uint32_t>(by + (k >> 1));
        if (!(x < p.width && y < p.height)) continue;
        const float w = filterWeight(p, fx - (static_cast<float>(x) + 0.5f), fy - (static_cast<float>(y) + 0.5f));
        F* px = rgbw + 4 * (static_cast<size_t>(y) * p.width + x);
        std::atomic<uint32_t>& lk = (*locks)[(static_cast<size_t>(y) * p.width + x) & (locks->size() - 1)];
        while (lk.exchange(1, std::memory_order_acquire)) {
        }
        px[0] += L.x * w, px[1] += L.y * w, px[2] += L.z * w, px[3] += w;
        lk.store(0, std::memory_order_release);
    }
}

}  // namespace

// =============================================================================================
// C ABI for ctypes (tests / bench cpu_baseline only)
// =============================================================================================
// One material row of or_scene_create's mats array.
Material materialOf(const float* m) {
    Material r{static_cast<uint32_t>(m[0]), {m[1], m[2], m[3]}, m[4]};
    if (r.type == kMatPlastic || r.type == kMatSubstrate || (r.type == kMatGlass && m[4] > 0.0f))
        r.alpha = remapRoughness(m[4]);
    r.alphaY = r.alpha;  // isotropic until or_scene_set_roughness_y gives glass or substrate a RoughnessY
    r.roughX = r.roughY = m[4];
    if (r.type == kMatPlastic || r.type == kMatSubstrate || r.type == kMatGlass) r.spec = {m[5], m[6], m[7]};
    if (r.type == kMatLambert && m[4] > 0.0f) {  // OrenNayar(reflection, sigma), Kernel.cpp:99-106
        const float sigma = m[4] * kPi, s2 = sigma * sigma;
        r.orenNayar = true;
        r.onA = 1.0f - s2 / (2.0f * (s2 + 0.33f));
        r.onB = 0.45f * s2 / (s2 + 0.09f);
    }
    return r;
}

extern "C" {

void or_sobol_matrices(uint32_t* out) { std::memcpy(out, g_sobol, sizeof(g_sobol)); }
float or_sobol_sample(uint64_t index, uint32_t dim) { return sobolSample(index, dim); }
uint64_t or_interval_to_index(uint32_t m, uint32_t sample, uint32_t px, uint32_t py) {
    return intervalToIndex(m, sample, px, py);
}

// Scene arrays:
//   verts[3*nv]; idx[3*nt]; triObj[nt]; objMat[nobj];
//   mats[8*nmat] = {type, r, g, b, alpha, spec r, g, b} (matte, plastic and substrate: the alpha
//   slot holds the roughness; glass: spec is the transmission colour);
//   quads[12*nq] = {origin xyz, U xyz, V xyz, radiance rgb};
//   cam[14] = {base xyz, lookAt xyz, up xyz, sizeX, sizeY, focal, aperture, 0}; env[3]
void* or_scene_create(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt, const uint32_t* triObj,
                      const uint32_t* objMat, uint32_t nobj, const float* mats, uint32_t nmat, const float* quads,
                      uint32_t nq, const float* cam, const float* env) {
    (void)nv;
    (void)nmat;
    auto* s = new Scene;
    s->tris.resize(nt);
    for (uint32_t i = 0; i < nt; ++i) {
        auto vtx = [&](uint32_t k) {
            const uint32_t j = idx[3 * i + k];
            return V3{verts[3 * j], verts[3 * j + 1], verts[3 * j + 2]};
        };
        s->tris[i] = {vtx(0), vtx(1), vtx(2)};
    }
    s->triObj.assign(triObj, triObj + nt);
    s->objMat.assign(objMat, objMat + nobj);
    for (uint32_t i = 0; i < nmat; ++i) {
        const float* m = mats + 8 * i;
        s->mats.push_back(materialOf(m));
    }
    for (uint32_t i = 0; i < nq; ++i) {
        const float* q = quads + 12 * i;
        Quad Q;
        Q.origin = {q[0], q[1], q[2]};
        Q.u = {q[3], q[4], q[5]};
        Q.v = {q[6], q[7], q[8]};
        Q.radiance = {q[9], q[10], q[11]};
        const V3 det = cross(Q.u, Q.v);  // BasicShapes.cpp:76-90
        Q.area = length(det);
        Q.normal = {det.x / Q.area, det.y / Q.area, det.z / Q.area};
        Q.light = i;
        s->quads.push_back(Q);
        s->lights.push_back({i, 1});
    }
    s->cam = {{cam[0], cam[1], cam[2]}, {cam[3], cam[4], cam[5]}, {cam[6], cam[7], cam[8]},
              cam[9], cam[10], cam[11], cam[12]};
    s->env = {env[0], env[1], env[2]};
    s->numStatic = nt;
    buildBVH(*s);
    return s;
}
void or_scene_destroy(void* s) { delete static_cast<Scene*>(s); }

// Moving instances: triangles [first[k], first[k + 1]) (the last up to the scene's end) are
// instance k's mesh in its own space; triangles before first[0] are static. Each instance starts
// fixed at the identity until or_scene_set_motion.
void or_scene_set_instances(void* sp, const uint32_t* first, uint32_t n) {
    Scene& s = *static_cast<Scene*>(sp);
    const uint32_t nt = static_cast<uint32_t>(s.tris.size());
    s.inst.assign(n, Scene::Instance{});
    for (uint32_t k = 0; k < n; ++k) {
        Scene::Instance& I = s.inst[k];
        I.first = first[k];
        I.count = (k + 1 < n ? first[k + 1] : nt) - first[k];
        const double id[16] = {1, 1, 1, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0};
        std::memcpy(I.srt[0], id, sizeof(id));
    }
    s.numStatic = n ? first[0] : nt;
    buildBVH(s);
}
// keys[16 * numKeys]: TransformSRT (scale xyz, skew xy xz yz, shift xyz, quaternion w x y z,
// translation xyz); parent[12]: row-major 3x4. numKeys 1: fixed and always present.
void or_scene_set_motion(void* sp, uint32_t k, uint32_t numKeys, int visible, float t0, float t1, const float* keys,
                         const float* parent) {
    Scene::Instance& I = static_cast<Scene*>(sp)->inst[k];
    I.keys = numKeys;
    I.visible = visible != 0;
    I.t0 = numKeys > 1 ? t0 : -INFINITY;
    I.t1 = numKeys > 1 ? t1 : INFINITY;
    for (uint32_t j = 0; j < numKeys; ++j) {
        for (int c = 0; c < 16; ++c) I.srt[j][c] = keys[16 * j + c];
        double n = 0.0;
        for (int c = 9; c < 13; ++c) n += I.srt[j][c] * I.srt[j][c];
        for (int c = 9; c < 13; ++c) I.srt[j][c] /= std::sqrt(n);
    }
    for (int c = 0; c < 12; ++c) I.parent[c] = parent[c];
}

// Traverse a wide BVH laid out as the GPU library builds it (pg_scene_bvh_download: numNodes nodes
// of 28 * width bytes, numTris leaf-order triangles of 48 bytes) instead of the oracle's own BVH2.
void or_scene_set_wide_bvh(void* sp, uint32_t width, const void* nodes, uint64_t numNodes, const void* tris,
                           uint64_t numTris) {
    Scene& s = *static_cast<Scene*>(sp);
    s.wideWidth = width;
    s.wideNodes.resize(static_cast<size_t>(numNodes) * 7 * width);
    std::memcpy(s.wideNodes.data(), nodes, s.wideNodes.size() * 4);
    s.wideTris.resize(static_cast<size_t>(numTris) * 12);
    std::memcpy(s.wideTris.data(), tris, s.wideTris.size() * 4);
}

// rays[8*n] = {o xyz, d xyz, tmin, tmax}; outputs t, prim (kMiss / kQuadBit|q / tri), u, v.
// brute != 0 tests every triangle instead of using the BVH. counts (optional) gets
// {nodes visited, triangles tested} summed over the rays.
void or_trace(void* sp, const float* rays, uint32_t n, float* t, uint32_t* prim, float* uv, int brute,
              uint64_t* counts) {
    const Scene& s = *static_cast<Scene*>(sp);
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        const Hit h = traceClosest(s, Ray{{r[0], r[1], r[2]}, {r[3], r[4], r[5]}}, r[6], r[7], brute != 0, counts);
        t[i] = h.t, prim[i] = h.prim, uv[2 * i] = h.u, uv[2 * i + 1] = h.v;
    }
}
// or_trace at per-ray times (moving instances).
void or_trace_at(void* sp, const float* rays, const float* times, uint32_t n, float* t, uint32_t* prim, float* uv) {
    const Scene& s = *static_cast<Scene*>(sp);
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        const Hit h = traceClosest(s, Ray{{r[0], r[1], r[2]}, {r[3], r[4], r[5]}}, r[6], r[7], false, nullptr, times[i]);
        t[i] = h.t, prim[i] = h.prim, uv[2 * i] = h.u, uv[2 * i + 1] = h.v;
    }
}
void or_occluded(void* sp, const float* rays, uint32_t n, uint8_t* out) {
    const Scene& s = *static_cast<Scene*>(sp);
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        out[i] = occluded(s, Ray{{r[0], r[1], r[2]}, {r[3], r[4], r[5]}}, r[6], r[7]);
    }
}

// Camera rays exactly as the renderer makes them; rays[8*n] as in or_trace (tmin 1e-3, tmax 1e5).
// pix[3*n] = {px, py, sample}.
void or_camera_rays(void* sp, const uint32_t* iparams /*width,height,spp*/, const float* ndc, const uint32_t* pix,
                    uint32_t n, float* rays, float* film) {
    const Scene& s = *static_cast<Scene*>(sp);
    const uint32_t w = iparams[0], h = iparams[1];
    uint32_t m = 0;
    while ((1u << m) < std::max(w, h)) ++m;
    const float res = static_cast<float>(1u << m);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t px = pix[3 * i], py = pix[3 * i + 1], si = pix[3 * i + 2];
        const uint64_t idx = intervalToIndex(m, si, px, py);
        float fx = std::fmin(std::fmax(sobolSample(idx, 0) * res - static_cast<float>(px), 0.0f), 1.0f);
        float fy = std::fmin(std::fmax(sobolSample(idx, 1) * res - static_cast<float>(py), 0.0f), 1.0f);
        fx += static_cast<float>(px);
        fy += static_cast<float>(py);
        const float ndcX = ndc[0] + ndc[2] * fx / static_cast<float>(w);
        const float ndcY = ndc[1] + ndc[3] * fy / static_cast<float>(h);
        const Ray r = cameraRay(s.cam, ndcX, ndcY, sobolSample(idx, 3), sobolSample(idx, 4));
        float* o = rays + 8 * i;
        o[0] = r.o.x, o[1] = r.o.y, o[2] = r.o.z, o[3] = r.d.x, o[4] = r.d.y, o[5] = r.d.z, o[6] = 1e-3f, o[7] = 1e5f;
        if (film) film[2 * i] = fx, film[2 * i + 1] = fy;
    }
}

// Render a rect over samples [sppBegin, sppEnd).
//   iparams = {width, height, spp, maxDepth, minRR, flags, rectX, rectY, rectW, rectH, sppBegin, sppEnd,
//              filter (0 box, 1 triangle, 2 gaussian)}
//   ndc = {ox, oy, sx, sy, gaussian Alpha}; film = RGBW double [height*width*4] (accumulated into);
//   pathL (optional) = per-path radiance [rectH*rectW*(sppEnd-sppBegin)*3], path order
//   ((py-rectY)*rectW + (px-rectX))*(sppEnd-sppBegin) + (s-sppBegin);
//   rayCounts (optional) = {extension, shadow, mis} ray totals.
void or_render(void* sp, const uint32_t* iparams, const float* ndc, double* film, float* pathL, int threads,
               uint64_t* rayCounts) {
    const Scene& s = *static_cast<Scene*>(sp);
    Params p;
    p.width = iparams[0], p.height = iparams[1], p.spp = iparams[2], p.maxDepth = iparams[3], p.minRR = iparams[4];
    p.flags = iparams[5], p.rectX = iparams[6], p.rectY = iparams[7], p.rectW = iparams[8], p.rectH = iparams[9];
    p.sppBegin = iparams[10], p.sppEnd = iparams[11];
    std::memcpy(p.ndc, ndc, sizeof(p.ndc));
    p.filter = iparams[12];
    p.negAlpha = -ndc[4];
    p.sub = std::exp(-ndc[4]);
    uint32_t m = 0;
    while ((1u << m) < std::max(p.width, p.height)) ++m;  // ScrambledSobolSampler.cpp:66-72
    const uint32_t ns = p.sppEnd - p.sppBegin;
    // 8x8 tiles handed to worker threads, samples looped serially per pixel
    // (FixedSampler.cpp:73-94, Parallel.cpp:475-491)
    const uint32_t tx = (p.rectW + 7) / 8, ty = (p.rectH + 7) / 8;
    std::atomic<uint32_t> next{0};
    std::vector<std::atomic<uint32_t>> locks(1 << 16);
    for (auto& l : locks) l.store(0);
    std::vector<PathStats> stats(std::max(threads, 1));
    auto worker = [&](int tid) {
        PathStats& st = stats[tid];
        for (uint32_t tile; (tile = next.fetch_add(1)) < tx * ty;) {
            const uint32_t bx = tile % tx, by = tile / tx;
            for (uint32_t yy = by * 8; yy < std::min(p.rectH, by * 8 + 8); ++yy)
                for (uint32_t xx = bx * 8; xx < std::min(p.rectW, bx * 8 + 8); ++xx)
                    for (uint32_t si = p.sppBegin; si < p.sppEnd; ++si) {
                        float fx, fy;
                        const uint32_t px = p.rectX + xx, py = p.rectY + yy;
                        const V3 L = renderSample(s, p, m, px, py, si, fx, fy, st);
                        if (pathL) {
                            float* o = pathL + 3 * ((static_cast<size_t>(yy) * p.rectW + xx) * ns + (si - p.sppBegin));
                            o[0] = L.x, o[1] = L.y, o[2] = L.z;
                        }
                        if (film) splat(film, p, fx, fy, L, &locks);
                    }
        }
    };
    std::vector<std::thread> pool;
    for (int i = 1; i < threads; ++i) pool.emplace_back(worker, i);
    worker(0);
    for (auto& t : pool) t.join();
    if (rayCounts) {  // extension, shadow and MIS rays; nodes visited and triangles tested by extension rays
        for (int k = 0; k < 5; ++k) rayCounts[k] = 0;
        for (auto& st : stats)
            rayCounts[0] += st.ext, rayCounts[1] += st.shadow, rayCounts[2] += st.mis, rayCounts[3] += st.trav[0],
                rayCounts[4] += st.trav[1];
    }
}

// Furnace-style and analytic helpers used by the oracle's own pinning tests.
// mat: 8 floats as in or_scene_create; front: the face flag of a plastic hit. The shading frame is
// the geometric one (Ng = +z).
void or_bsdf_eval(const float* mat, int front, const float* wo, const float* wi, float* f, float* pdf) {
    Material m = materialOf(mat);
    m.front = front != 0;
    const V3 o{wo[0], wo[1], wo[2]}, i{wi[0], wi[1], wi[2]};
    const V3 r = bsdfEval(m, o, i, V3{0, 0, 1});
    f[0] = r.x, f[1] = r.y, f[2] = r.z;
    *pdf = bsdfPdf(m, o, i);
}
void or_bsdf_sample(const float* mat, int front, const float* wo, float u1, float u2, float* wi, float* f, float* pdf) {
    Material m = materialOf(mat);
    m.front = front != 0;
    const BsdfSample s = bsdfSample(m, V3{wo[0], wo[1], wo[2]}, V3{0, 0, 1}, u1, u2);
    wi[0] = s.wi.x, wi[1] = s.wi.y, wi[2] = s.wi.z;
    f[0] = s.f.x, f[1] = s.f.y, f[2] = s.f.z;
    *pdf = s.pdf;
}
// The same with RoughnessY = roughY for glass and substrate (their anisotropic distribution).
Material anisotropicOf(const float* mat, float roughY) {
    Material m = materialOf(mat);
    if (m.type != kMatGlass && m.type != kMatSubstrate) return m;
    m.roughY = roughY;
    const bool rough = m.type == kMatSubstrate || m.roughX > 0.0f || roughY > 0.0f;
    m.alpha = rough ? remapRoughness(m.roughX) : 0.0f;
    m.alphaY = rough ? remapRoughness(roughY) : 0.0f;
    return m;
}
void or_bsdf_eval_y(const float* mat, float roughY, int front, const float* wo, const float* wi, float* f, float* pdf) {
    Material m = anisotropicOf(mat, roughY);
    m.front = front != 0;
    const V3 o{wo[0], wo[1], wo[2]}, i{wi[0], wi[1], wi[2]};
    const V3 r = bsdfEval(m, o, i, V3{0, 0, 1});
    f[0] = r.x, f[1] = r.y, f[2] = r.z;
    *pdf = bsdfPdf(m, o, i);
}
void or_bsdf_sample_y(const float* mat, float roughY, int front, const float* wo, float u1, float u2, float* wi, float* f,
                      float* pdf) {
    Material m = anisotropicOf(mat, roughY);
    m.front = front != 0;
    const BsdfSample s = bsdfSample(m, V3{wo[0], wo[1], wo[2]}, V3{0, 0, 1}, u1, u2);
    wi[0] = s.wi.x, wi[1] = s.wi.y, wi[2] = s.wi.z;
    f[0] = s.f.x, f[1] = s.f.y, f[2] = s.f.z;
    *pdf = s.pdf;
}
// Area lights of several primitives: light l is the next counts[l] quads (sum = number of quads);
// without this call every quad is a light of its own. As the Plane constructor
// (BasicShapes.cpp:59-94): the light's area is the float sum of its primitives' |U x V|, the CDF
// entry of primitive i the running sum divided by that area, its probability its area divided by it.
int or_scene_set_area_lights(void* scene, const uint32_t* counts, uint32_t n) {
    auto* s = static_cast<Scene*>(scene);
    uint32_t first = 0;
    for (uint32_t l = 0; l < n; ++l) first += counts[l];
    if (first != s->quads.size()) return -1;
    s->lights.clear();
    first = 0;
    for (uint32_t l = 0; l < n; ++l) {
        float total = 0.0f;
        std::vector<float> cum, area;
        for (uint32_t k = 0; k < counts[l]; ++k) {
            const Quad& q = s->quads[first + k];
            const float a = length(cross(q.u, q.v));
            total = total + a;
            cum.push_back(total);
            area.push_back(a);
        }
        for (uint32_t k = 0; k < counts[l]; ++k) {
            Quad& q = s->quads[first + k];
            q.light = l;
            q.cdf = cum[k] / total;
            q.pdf = area[k] / total;
            q.area = total;
        }
        s->lights.push_back({first, counts[l]});
        first += counts[l];
    }
    return 0;
}
// Anisotropic roughness of glass and substrate (RoughnessX is the mats row's roughness): ry[nmat]
// RoughnessY per material (or NULL: equal to RoughnessX), texY[nmat] its texture (~0u: the constant;
// or NULL). Returns -1 if a surface other than glass or substrate is given a second roughness.
int or_scene_set_roughness_y(void* scene, const float* ry, const uint32_t* texY) {
    auto* s = static_cast<Scene*>(scene);
    for (size_t i = 0; i < s->mats.size(); ++i) {
        Material& m = s->mats[i];
        const float y = ry ? ry[i] : m.roughX;
        const uint32_t t = texY ? texY[i] : ~0u;
        const bool two = m.type == kMatGlass || m.type == kMatSubstrate;
        if (!two && (y != m.roughX || t != ~0u)) return -1;
        if (!two) continue;
        m.roughY = y, m.texY = t, m.ownY = true;
        // glassInit: smooth (alpha 0) only when both roughnesses are zero
        const bool rough = m.type == kMatSubstrate || m.roughX > 0.0f || y > 0.0f;
        m.alpha = rough ? remapRoughness(m.roughX) : 0.0f;
        m.alphaY = rough ? remapRoughness(y) : 0.0f;
    }
    return 0;
}
// points[6*n] = {position xyz, intensity rgb}; light order: environment, area lights, points.
void or_scene_set_points(void* scene, const float* points, uint32_t n) {
    auto* s = static_cast<Scene*>(scene);
    s->points.clear();
    for (uint32_t i = 0; i < n; ++i) {
        const float* p = points + 6 * i;
        s->points.push_back({{p[0], p[1], p[2]}, {p[3], p[4], p[5]}});
    }
}
// Textures: triuv[6*nt] texture coordinates of each triangle's corners (or NULL), one texture per
// or_scene_add_texture call, mtex[3*nmat] albedo / spec / roughness texture per material (~0u: none).
void or_scene_set_textures(void* scene, const float* triuv, const uint32_t* mtex) {
    auto* s = static_cast<Scene*>(scene);
    s->triUV.assign(triuv, triuv ? triuv + 6 * s->tris.size() : triuv);
    s->matTex.resize(s->mats.size());
    for (size_t i = 0; i < s->mats.size(); ++i) {
        s->matTex[i] = {mtex[3 * i], mtex[3 * i + 1], mtex[3 * i + 2]};
        Material& m = s->mats[i];  // an isotropic glass or substrate: RoughnessY follows RoughnessX's texture
        if (!m.ownY && (m.type == kMatGlass || m.type == kMatSubstrate)) m.texY = mtex[3 * i + 2];
    }
}
// itype {type, channels, width, height, wrap}; ftype {black[4], white[4], scale}
uint32_t or_scene_add_texture(void* scene, const uint32_t* itype, const float* ftype, const uint8_t* texels) {
    auto* s = static_cast<Scene*>(scene);
    Texture t;
    t.type = itype[0], t.channels = itype[1], t.width = itype[2], t.height = itype[3], t.wrap = itype[4];
    std::memcpy(t.black, ftype, 16), std::memcpy(t.white, ftype + 4, 16);
    t.scale = ftype[8];
    if (t.type == 2) t.texels.assign(texels, texels + static_cast<size_t>(t.width) * t.height * t.channels);
    s->textures.push_back(std::move(t));
    return static_cast<uint32_t>(s->textures.size() - 1);
}
// one texture lookup (for tests)
void or_texture_sample(void* scene, uint32_t tex, float u, float v, float* out) {
    const auto c = sampleTexture(static_cast<Scene*>(scene)->textures[tex], u, v);
    std::memcpy(out, c.data(), 16);
}
float or_fresnel_dielectric(float cosI, float etaI, float etaT) { return fresnelDielectric(cosI, etaI, etaT); }
float or_remap_roughness(float roughness) { return remapRoughness(roughness); }
// the anisotropic D at a unit half vector (x, y, z)
float or_ggx_d_aniso(float x, float y, float z, float ax, float ay) { return ggxD(V3{x, y, z}, Dist{ax, ay}); }
float or_ggx_lambda(float x, float y, float z, float ax, float ay) { return ggxLambda(V3{x, y, z}, Dist{ax, ay}); }
void or_ggx_sample_wh(const float* wo, float ax, float ay, float u1, float u2, float* wh) {
    const V3 h = ggxSampleWh(V3{wo[0], wo[1], wo[2]}, Dist{ax, ay}, u1, u2);
    wh[0] = h.x, wh[1] = h.y, wh[2] = h.z;
}
float or_ggx_d(float cosTheta, float alpha) {
    const float s = std::sqrt(std::fmax(0.0f, 1.0f - cosTheta * cosTheta));
    return ggxD(V3{s, 0.0f, cosTheta}, Dist{alpha, alpha});
}
}

// Layouts shared by the host library code and the gfx950 kernels of libpiper_hip.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace pg {

constexpr uint32_t kSobolDims = 1024;  // ScrambledSobolSampler/Shared.hpp:22
constexpr uint32_t kSobolCols = 52;    // ScrambledSobolSampler/Shared.hpp:23
constexpr uint32_t kMaxLog2Res = 16;   // interval-to-index tables cover resolutions up to 2^16
// Byte-sliced Sobol table for the first kByteDims dimensions: entry [k][b][d] is the XOR of the
// columns 8k..8k+7 of dimension d selected by the bits of byte value b.
constexpr uint32_t kByteDims = 128;
constexpr uint32_t kByteChunks = (kSobolCols + 7) / 8;

constexpr uint32_t kInvalidChild = 0xFFFFFFFFu;
constexpr uint32_t kLeafBit = 0x80000000u;  // child = kLeafBit | (count-1) << 27 | first triangle
constexpr uint32_t kLeafCountShift = 27;
constexpr uint32_t kLeafFirstMask = (1u << kLeafCountShift) - 1u;
constexpr uint32_t kMaxLeafSize = 8;
// Traversal stack entries per lane kept in LDS; deeper entries spill to a per-lane HBM region.
// 10 entries (5 KB per 64-lane wave) let 32 waves share a CU's 160 KB, so registers, not LDS,
// bound the traversal kernels' occupancy (16 entries held them to 20 waves per CU).
constexpr int kLdsStack = 10;
// Upper bound on the resident waves per CU of a traversal kernel (the LDS limit above); the
// per-lane spill region is sized for this many lanes.
constexpr uint32_t kTravWavesPerCuMax = 32;

constexpr uint32_t kHitMiss = 0xFFFFFFFFu;
constexpr uint32_t kHitQuad = 0x80000000u;
constexpr uint32_t kMisFlag = 0x80000000u;  // ShadowRec::path bit: MIS-BSDF light query, not NEE

// Wide BVH node, structure-of-arrays inside the node so that one lane streams the whole node with
// 16-byte loads: W boxes as six planes of W floats, then W child references.
template <int W>
struct alignas(16) WideNode {
    float lox[W], loy[W], loz[W], hix[W], hiy[W], hiz[W];
    uint32_t child[W];
};
static_assert(sizeof(WideNode<8>) == 224, "BVH8 node is 224 B");
static_assert(sizeof(WideNode<4>) == 112, "BVH4 node is 112 B");

// Material types: the public PG_MAT_* values (0 Lambert, 1 conductor, 2 plastic, 3 mirror,
// 4 smooth glass, 5 substrate) plus this device-only one.
constexpr uint32_t kDevOrenNayar = 16;  // matte with roughness > 0
constexpr uint32_t kDevRoughGlass = 17;  // glass with roughness > 0 (alpha: remapped roughness)
constexpr uint32_t kMatTextured = 0x80000000u;  // type flag: a colour input comes from a texture
constexpr uint32_t kTexNone = 0xFFFFFFFFu;
constexpr uint32_t kTexChecker = 1, kTexImage = 2;   // PG_TEX_*
constexpr uint32_t kWrapRepeat = 0, kWrapMirror = 1;  // PG_WRAP_*
struct DevTexture {  // 64 B
    uint32_t type, channels, width, height;
    uint32_t wrap;
    uint32_t offset;  // first texel byte in the scene's texel pool
    float scale, pad;
    float black[4], white[4];
};
struct DevMaterial {  // 48 B
    uint32_t type;
    float albedo[3];   // diffuse reflectance (Lambert, Oren-Nayar, plastic, substrate), conductor F0,
                       // mirror / glass reflection
    float alpha;       // GGX alpha (plastic, substrate, glass: remapped RoughnessX); Oren-Nayar: A
    float spec[3];     // plastic / substrate specular; glass transmission; Oren-Nayar: {B, 0, 0}
    float alphaY;      // glass, substrate: remapped RoughnessY (= alpha when isotropic)
    float roughX, roughY;  // glass, substrate: the RoughnessX / RoughnessY constants as configured,
                           // for a hit whose other roughness input comes from a texture
    uint32_t texY;     // glass, substrate: RoughnessY texture (kTexNone: the constant)
};
static_assert(sizeof(DevMaterial) == 48, "DevMaterial is 48 B");

struct DevQuad {  // Plane primitive (PerPlaneData) + DiffuseAreaLightData, 80 B
    float origin[3], u[3], v[3];
    float normal[3];
    float radiance[3];    // the owning light's radiance
    float area;           // the owning light's area: every primitive of its Plane (DiffuseAreaLightData::area)
    uint32_t maxDetComp;  // PerPlaneData::maxDetComp, BasicShapes.cpp:70-75
    uint32_t light;       // owning area light: quads of one light are consecutive
    float cdf, pdf;       // this primitive's area CDF and probability within its light (BasicShapes.cpp:76-94)
};
// Area light: a Plane of `count` primitives starting at quad `first` (DiffuseAreaLight over Plane).
struct DevAreaLight {
    uint32_t first, count;
};

struct DevCamera {
    float base[3], lookAt[3], up[3];
    float sizeX, sizeY, focal, aperture;
};

// Moving instance (a Group child with a keyframed TransformSRT, Embree.cpp:709-773): its mesh is a
// BLAS of its own in the scene's node and triangle arrays, in the instance's local space, drawn at
// each ray's time. Its leaf triangles carry the instance index + 1 in the w of their second vertex
// (0 on static triangles).
constexpr uint32_t kMaxMotionKeys = 8;
constexpr uint32_t kMaxInstances = 256;
struct DevInstance {
    float lo[3], hi[3];  // world box over every time at which the instance is visible
    float t0, t1;        // visible for t0 <= time <= t1 (shutter-normalised time of the first and last key)
    float rate;          // (keys - 1) / (t1 - t0): key segments per unit time
    uint32_t keys;       // transforms at equal time steps from t0 to t1 (1: fixed)
    uint32_t root;       // BLAS root node
    uint32_t visible;    // 0: absent from the frame
    float parent[12];    // static row-major 3x4 transform of the instance's parent nodes
    float srt[kMaxMotionKeys][16];  // TransformSRT: scale xyz, skew xy xz yz, shift xyz, unit quaternion w x y z, translation xyz
};

struct DevScene {
    const float4* tris;      // 3 float4 per triangle in leaf order: (v0, prim bits) (v1, 0) (v2, 0)
    const uint32_t* triMat;  // material per leaf-order triangle
    const void* nodes;       // WideNode<width>[]
    const DevMaterial* mats;
    const DevQuad* quads;
    const float* points;         // point lights: position xyz, intensity rgb
    const float* triUV;          // texture coordinates of each original triangle's corners (6 floats)
    const DevTexture* textures;
    const uint8_t* texels;
    const uint4* matTex;         // per material: albedo, spec, roughness texture (kTexNone: the constant), base type
    const uint32_t* sobol;       // kSobolDims * kSobolCols
    const uint32_t* sobolT;      // the same matrices bit-major: kSobolCols * kSobolDims
    const uint32_t* sobolB;      // byte-sliced: kByteChunks * 256 * kByteDims
    const uint64_t* sobolInv;    // [kMaxLog2Res+1][32]
    const uint64_t* sobolDelta;  // [kMaxLog2Res+1][64]
    uint32_t numTris, numNodes, numQuads, width;
    uint32_t numPoints;
    const DevAreaLight* lights;  // area lights: lights are {environment, area lights, point lights}
    uint32_t numLights;
    uint32_t textured;  // some material has a texture: shade takes its textured variant
    DevCamera cam;
    float env[3];
    uint32_t hasBvh;              // static triangles exist (their BVH's root is node 0)
    const DevInstance* inst;      // moving instances
    uint32_t numInst;
};

// Queue records written by the shade stage.
struct alignas(16) ShadowRec {  // 48 B
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
    float cr, cg, cb;
    uint32_t path;
};
struct alignas(16) MisRec {  // 48 B
    float ox, oy, oz, bsdfPdf;
    float dx, dy, dz;
    uint32_t light;  // selected light (0 = environment, 1 + l = area light l)
    float cr, cg, cb;
    uint32_t path;
};

// Counters in device memory: a block of u32, then the pool's u64 path numbers.
enum CounterSlot : uint32_t {
    kCntExtCur = 0,     // extension queues (two, alternating by bounce)
    kCntExtNext = 1,
    kCntShadow = 2,
    kCntMis = 3,
    kCntDone = 4,       // free path slots: done[0, count); finished paths are pushed on top
    kCntRecorded = 5,   // done entries below this have had their film records written
    kCntRegenBase = 6,  // this iteration's new paths take the slots done[base, base + count)
    kCntRegenCount = 7,
    kCntRegenExt = 8,   // ... and the extension queue entries [ext, ext + count)
    kCntNum = 12
};
struct PoolPaths {
    unsigned long long next;        // number of the next path to start
    unsigned long long regenFirst;  // number of this iteration's first new path
    unsigned long long limit;       // paths numbered below this may start (set by the host)
};
constexpr uint32_t kCountsBytes = kCntNum * 4 + sizeof(PoolPaths);
enum TotalSlot : uint32_t { kTotExt = 0, kTotShadow = 1, kTotMis = 2, kTotNum = 4 };

struct RenderConsts {
    uint32_t width, height, spp;
    uint32_t rectX, rectY, rectW, rectH;
    uint32_t sppBegin, sppCount;
    uint32_t maxDepth, minRR, flags;
    uint32_t log2Res;
    uint32_t segSpp;  // samples per pixel of one film segment (path numbers run segment by segment)
    float ndc[4];
    uint32_t filter;       // kFilterBox / kFilterTriangle / kFilterGaussian
    float negAlpha, sub;   // gaussian: -Alpha and exp(-Alpha) (FixedSampler.cpp:40-44)
};
constexpr uint32_t kFilterBox = 0, kFilterTriangle = 1, kFilterGaussian = 2;

constexpr uint32_t kCompatD1 = 0x1u, kCompatD3 = 0x2u, kCompatD9 = 0x4u, kCompatD13 = 0x8u;
constexpr uint32_t kFilmFp64 = 0x100u;
constexpr uint32_t kCountTraversal = 0x200u;
constexpr uint32_t kNoStageTiming = 0x400u;
constexpr uint32_t kUseGraph = 0x800u;

}  // namespace pg

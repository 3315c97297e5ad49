// Launchers of the gfx950 kernels (pg_kernels.hip) used by the host library (pg_api.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include "pg_shared.hpp"

namespace pg {

// State of one path slot: exactly one 128 B cache line. The pool hands slots out in whatever
// order paths finish, so the stages touch slots at random; one line per slot keeps each such
// access to a single line instead of one line per field.
struct alignas(128) PathRec {
    float4 o;       // ray origin
    float4 d;       // ray direction
    float4 thr;     // throughput
    float4 L;       // radiance so far
    float4 Lmis;    // MIS-BSDF light hit of the current bounce, added to L after its NEE term
    float4 hit;     // t, hit position bits, u, v
    uint4 meta;     // next Sobol dim, depth, flags, path number within its work item
    uint64_t idx;   // Sobol index
    float2 film;    // film point
};
static_assert(sizeof(PathRec) == 128, "one cache line per path slot");

struct PathState {
    PathRec* p;  // capacity slots
};

// meta.z bits: 0 camera/specular ray, 1 a guided vertex awaits its radiance snapshot,
// 8..15 guided vertices recorded for training, 16..31 work item of a batched render
constexpr uint32_t kMetaPending = 0x2u;

// Guiding state handed to the shade and record stages.
struct GuideShade {
    const float* mix;  // kGMixFloats per extension-queue entry (pg_guide.hpp layout)
    float4* hist;      // per path slot: histDepth records of 6 float4
    float alpha;       // probability of sampling the guide (negative: the network's learned one)
    float3 lo, invExt; // scene box -> [0,1]^3
    uint32_t histDepth, trainStride, trainPhase;
    uint32_t enabled, train;
};

// Training records emitted when a training path finishes.
struct GuideRecordOut {
    float4* rec;       // 4 float4 per record
    uint32_t* count;
    uint32_t capacity;
};

// Guiding of one item of a guided batch: its scene's guide and the scene's settings. Items of
// scenes that share a guide share its key, its weights and its training records.
constexpr uint32_t kMaxBatchGuides = 256;
struct ItemGuide {
    const uint16_t* wb;    // the guide's BF16 weights
    const float* theta;    // and its FP32 parameters (biases)
    GuideRecordOut tr;     // the guide's training records (capacity 0: no training)
    float3 lo, invExt;     // the scene's box -> [0,1]^3
    float alpha;           // probability of sampling the guide (negative: the network's learned one)
    uint32_t key;          // index of the guide among the batch's distinct guides
    uint32_t trainStride, trainPhase;
};

struct Queues {
    uint32_t* ext[2];  // extension queues (current / next bounce)
    ShadowRec* shadow;
    MisRec* mis;
    uint32_t* done;    // free path slots, with finished paths pushed on top (kCntDone)
    uint32_t* counts;  // kCntNum u32
    PoolPaths* pool;
};

// Work items: film segments (whole rect x a sample sub-range) with consecutive path numbers.
// A single-scene render splits its sample range into segments of rc.segSpp samples, so item k
// owns path numbers [k * P, (k + 1) * P) with P = rectW * rectH * segSpp. A batched render lists
// its items in device tables, item k owning [pathOffsets[k], pathOffsets[k + 1]). Either way a
// path carries its item id in meta.z bits 16..31 and its number within the item in meta.w.
constexpr uint32_t kMaxBatchItems = 1u << 16;
struct BatchTable {
    const DevScene* scenes;       // per item: its scene
    const RenderConsts* rc;       // per item: rect, sample sub-range, depth, flags
    const uint64_t* pathOffsets;  // n + 1 entries
    uint32_t n;
    const ItemGuide* guide;       // per item, in guided batches (nullptr otherwise)
    uint32_t motion;              // some scene has moving instances (selects the kernel variants)
};

// Film records live in a ring of `ring` slots: path number G writes slot G mod ring.
struct RecordRing {
    float4* L;       // (L, film x)
    float* y;        // film y
    uint64_t size;
    uint32_t* itemDone;  // paths recorded per item
};

// Launch settings of a scene (pg_autotune picks them; no choice changes any result): the
// traversal stack entries each lane keeps in LDS (deeper ones spill to HBM), the traversal waves
// per CU (0: as many as stay resident), the shade block and the waves per guide-inference block.
constexpr uint32_t kLdsStackMax = 16;
struct LaunchTune {
    uint32_t ldsStack = kLdsStack;  // 4 .. kLdsStackMax
    uint32_t travWaves = 0;
    uint32_t shadeBlock = 256;      // 128 or 256
    uint32_t inferWaves = 8;        // 4, 8 or 16 (guided batches always use 8)
};

// `bt` selects the batched kernels (nullptr: one scene, `s` and `rc` in kernel arguments).
// Batched launches still take the first scene as `s`: all scenes share its BVH width.
//
// Path pool: `capacity` slots. launchPoolInit marks every slot free. Each iteration,
// launchIterBegin (one thread) takes up to pool->limit - pool->next new path numbers for the free
// slots, then launchRegen starts those paths and appends their slots to extension queue curSlot.
void launchPoolInit(hipStream_t st, const Queues& q, uint32_t capacity);
void launchIterBegin(hipStream_t st, const Queues& q, uint32_t curSlot);
void launchRegen(hipStream_t st, const DevScene& s, const PathState& ps, const Queues& q, const RenderConsts& rc,
                 uint32_t curSlot, uint32_t gridThreads, const BatchTable* bt = nullptr);
void launchExtend(hipStream_t st, const DevScene& s, const PathState& ps, const Queues& q, uint32_t curSlot,
                  uint32_t gridLanes, uint2* spill, uint32_t spillPerLane, uint32_t* travCounts,
                  const BatchTable* bt, const LaunchTune& lt);
void launchShade(hipStream_t st, const DevScene& s, const PathState& ps, const Queues& q, const RenderConsts& rc,
                 uint32_t curSlot, uint32_t nextSlot, uint32_t gridThreads, uint64_t* totals, const GuideShade& gs,
                 const BatchTable* bt, bool textured, const LaunchTune& lt);
void launchShadow(hipStream_t st, const DevScene& s, const PathState& ps, const Queues& q, uint32_t gridLanes,
                  uint2* spill, uint32_t spillPerLane, const BatchTable* bt, const LaunchTune& lt);
void launchMis(hipStream_t st, const DevScene& s, const PathState& ps, const Queues& q, uint32_t gridLanes,
               uint2* spill, uint32_t spillPerLane, const BatchTable* bt, const LaunchTune& lt);
// Film records of the paths that finished this iteration (done entries above kCntRecorded).
void launchRecord(hipStream_t st, const PathState& ps, const Queues& q, const RenderConsts& rc, uint32_t gridThreads,
                  const RecordRing& ring, const GuideShade& gs, const GuideRecordOut& tr,
                  const BatchTable* bt = nullptr);
// Resolve one item (rc: its rect and sample sub-range; its first path number `first`) into the
// film. quad: 16 floats (or doubles with film64) per rect pixel.
void launchResolve(hipStream_t st, const RenderConsts& rc, const RecordRing& ring, uint64_t first, void* quad,
                   float* film32, double* film64);
// times: the ray times (moving instances), or nullptr for time 0.
void launchTrace(hipStream_t st, const DevScene& s, const float* rays, uint32_t n, float* t, uint32_t* prim, float* uv,
                 uint8_t* occ, uint32_t gridLanes, uint2* spill, uint32_t spillPerLane, const LaunchTune& lt,
                 const float* times = nullptr);
void launchCameraRays(hipStream_t st, const DevScene& s, const RenderConsts& rc, const uint32_t* pix, uint32_t n,
                      float* rays, float* film);
void launchCopy(hipStream_t st, const void* src, void* dst, size_t bytes);
void launchSobol(hipStream_t st, const DevScene& s, uint32_t m, uint32_t dims, const uint32_t* pix, uint32_t n,
                 float* out, uint64_t* indexOut);

}  // namespace pg

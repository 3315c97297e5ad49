// Launchers of the guiding-network kernels (pg_guide.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pg_kernels.hpp"
#include "pg_shared.hpp"

namespace pg {

constexpr uint32_t kGTiles = 96;       // 16x16 weight-gradient tiles over the five layers
constexpr uint32_t kGMaxChunks = 64;   // sample chunks of the split-K weight gradient
constexpr uint32_t kGParamCount = 24880;
constexpr uint32_t kGMixFloatsHost = 52;  // mixture record floats per sample

struct WGradTile {
    uint16_t rows, cols, mo, ki;  // layer shape and tile coordinates
    uint32_t wOff, bOff;          // parameter offsets
    const uint16_t* dz;
    const uint16_t* x;
};

struct AdamParams {
    float lr, b1, b2, eps;
    float c1, c2;  // bias corrections 1 - b1^t, 1 - b2^t
    uint32_t update;
};

// Sample layout of one training batch of n records: row stride npad (a multiple of 64) split into
// `chunks` chunks of `chunk` samples.
struct GuideBatchLayout {
    uint32_t npad, chunk, chunks;
};
GuideBatchLayout guideBatchLayout(uint32_t n);
size_t guideTrainScratchBytes(const GuideBatchLayout& b);
void buildGuideTiles(void* scratch, const GuideBatchLayout& b, WGradTile* host);

// Mixture records (48 floats per sample) for n samples given as 3 float4 each (pos01, normal, wo).
void launchGuideInferArray(hipStream_t st, const float* in, uint32_t n, const uint16_t* wb, const float* theta,
                           float* mix, uint32_t maxBlocks);
// Mixture records for the surface hits of an extension queue (count read on the device); record i
// belongs to queue entry i. cus: the device's CUs; waves: waves per block (4, 8 or 16).
void launchGuideInferPaths(hipStream_t st, const DevScene& s, const PathState& ps, const uint32_t* queue,
                           const uint32_t* count, uint32_t capacity, float3 lo, float3 invExt, const uint16_t* wb,
                           const float* theta, float* mix, uint32_t cus, uint32_t waves);
// Guided batches: the extension queue mixes scenes with different guides. Each iteration groups
// the queue entries by guide key without global atomics: kSortBlocks blocks count their
// contiguous share of the queue per key, one block scans those counts (every group padded to
// kGroupUnit entries), and the same blocks scatter their entry indices; then every inference block
// runs units of one guide with that guide's weights in LDS. Record i still belongs to queue entry i.
constexpr uint32_t kGroupUnit = 512;
constexpr uint32_t kSortBlocks = 256;
struct GroupBuffers {
    uint32_t* blockCount;  // [kMaxBatchGuides][kSortBlocks]: entries of key k in sort block b
    uint32_t* blockBase;   // same shape: where sort block b's entries of key k start in perm
    uint32_t* keyStart;    // kMaxBatchGuides + 1: padded group starts; [nKeys] = padded total
    uint32_t* keyEnd;      // kMaxBatchGuides: end of each group's real entries
    uint32_t* perm;        // capacity + kMaxBatchGuides * kGroupUnit queue-entry indices
};
size_t groupBufferBytes(uint32_t capacity);
GroupBuffers groupBuffers(void* base, uint32_t capacity);
void launchGuideInferGrouped(hipStream_t st, const BatchTable& bt, const PathState& ps, const uint32_t* queue,
                             const uint32_t* count, uint32_t capacity, uint32_t nKeys, const GroupBuffers& gb,
                             const uint16_t* const* keyWb, const float* const* keyTheta, float* mix,
                             uint32_t maxBlocks);
// Forward + backward over n training records (4 float4 each: (pos01, c) (normal, BSDF pdf) (wo)
// (direction)), then the split-K weight gradients into partial[chunks][kGParamCount]. learnAlpha: the
// loss is that of the one-sample MIS density alpha p_guide + (1 - alpha) p_bsdf with alpha the
// network's own selection output (trained too); otherwise that of p_guide alone.
void launchGuideGradients(hipStream_t st, const float* rec, uint32_t n, const GuideBatchLayout& b, const uint16_t* wb,
                          const float* theta, void* scratch, float* csum, float* partial, float* lossAcc,
                          const WGradTile* tilesDev, uint32_t maxBlocks, bool learnAlpha);
void launchGuideAdam(hipStream_t st, const float* partial, uint32_t chunks, float* theta, float* m, float* v,
                     uint16_t* wb, float* gradOut, const AdamParams& ap);
void launchGuideToBf16(hipStream_t st, const float* theta, uint16_t* wb);
// Raise the dynamic-LDS limit of the network kernels (they use 64-96 KB per block); idempotent.
void guideConfigureLds();

}  // namespace pg

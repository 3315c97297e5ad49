// Host wide-BVH builder interface.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <functional>
#include <mutex>
#include <vector>

#include "pg_shared.hpp"

namespace pg {

struct BvhParams {
    uint32_t width = 8, leafSize = 4, bins = 32;
};

struct WideBvh {
    uint32_t width = 8;
    uint32_t numNodes = 0, numLeaves = 0, maxDepth = 0;
    float pad = 0.0f;
    std::vector<float> nodes;     // WideNode<width> as raw floats (child refs stored as bits)
    std::vector<float> tris;      // 12 floats per triangle in leaf order
    std::vector<uint32_t> order;  // leaf order -> original triangle index
};

// triVerts: 9 floats per triangle (v0, v1, v2), world space.
void buildWideBvh(const float* triVerts, uint32_t numTris, const BvhParams& params, WideBvh& out);

// A BVH built on the device: the same node and triangle layout as WideBvh, in device buffers the
// caller owns and hands out through `alloc` (called once per buffer with its size in bytes).
struct DeviceBvh {
    enum Buffer { kNodes, kTris, kTriMat };
    std::function<void*(Buffer, size_t)> alloc;
    uint32_t width = 8, numNodes = 0, numLeaves = 0, maxDepth = 0;
    float pad = 0.0f;
};

// dTriVerts: 9 floats per triangle on the device; dTriMat: material per triangle. Runs on `st`
// and returns with the stream drained.
void buildWideBvhDevice(hipStream_t st, const float* dTriVerts, const uint32_t* dTriMat, uint32_t numTris,
                        const BvhParams& params, DeviceBvh& out);

// Sobol generator matrices (kSobolDims x kSobolCols, u32) and interval-to-index tables
// inv[(kMaxLog2Res+1) x 32], delta[(kMaxLog2Res+1) x 64].
void buildSobolTables(std::vector<uint32_t>& matrices, std::vector<uint64_t>& inv, std::vector<uint64_t>& delta);
// Byte-sliced table (kByteChunks x 256 x kByteDims) from the matrices.
void buildSobolByteTable(const std::vector<uint32_t>& matrices, std::vector<uint32_t>& table);

}  // namespace pg

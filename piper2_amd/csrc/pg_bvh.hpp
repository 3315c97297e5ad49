// Host wide-BVH builder interface.
#pragma once
#include <cstdint>
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

// Sobol generator matrices (kSobolDims x kSobolCols, u32) and interval-to-index tables
// inv[(kMaxLog2Res+1) x 32], delta[(kMaxLog2Res+1) x 64].
void buildSobolTables(std::vector<uint32_t>& matrices, std::vector<uint64_t>& inv, std::vector<uint64_t>& delta);
// Byte-sliced table (kByteChunks x 256 x kByteDims) from the matrices.
void buildSobolByteTable(const std::vector<uint32_t>& matrices, std::vector<uint32_t>& table);

}  // namespace pg

// Sobol generator matrices and interval-to-index tables for the device sampler.
//
// The reference samples with pbrt-v3-style 32-bit Sobol matrices built from the Joe-Kuo
// "new-joe-kuo-6.21201" direction numbers, 1024 dimensions x 52 columns
// (Sampler/ScrambledSobolSampler/Shared.hpp:22-23, ScrambledSobolSampler.cpp:26-27); its generated
// table is not in the reference tree, so it is rebuilt here from the published direction numbers.
// sobolIntervalToIndex (Kernel.cpp:31-50) needs, per resolution 2^m, the inverse of the GF(2) map
// from the low 2m index bits to the top m bits of dimensions 0 and 1; that inverse is computed
// here by Gaussian elimination over GF(2).
#include <stdexcept>

#include "pg_bvh.hpp"

namespace pg {
namespace {

struct DirectionNumbers {
    uint32_t degree, coeffs;
    uint32_t m[18];
};
const DirectionNumbers kDirections[kSobolDims] = {
#include "joe_kuo_1024.inc"
};

}  // namespace

void buildSobolTables(std::vector<uint32_t>& mat, std::vector<uint64_t>& inv, std::vector<uint64_t>& delta) {
    mat.assign(kSobolDims * kSobolCols, 0u);
    for (uint32_t d = 0; d < kSobolDims; ++d) {
        // 64-bit direction numbers v_j = m_j / 2^(j+1) as fixed point; keep the high 32 bits.
        uint64_t v[kSobolCols];
        const uint32_t s = kDirections[d].degree;
        for (uint32_t j = 0; j < kSobolCols; ++j) {
            if (d == 0) {
                v[j] = 1ull << (63 - j);  // van der Corput
            } else if (j < s) {
                v[j] = static_cast<uint64_t>(kDirections[d].m[j]) << (63 - j);
            } else {
                uint64_t x = v[j - s] ^ (v[j - s] >> s);
                for (uint32_t k = 1; k < s; ++k)
                    if ((kDirections[d].coeffs >> (s - 1 - k)) & 1u) x ^= v[j - k];
                v[j] = x;
            }
            mat[d * kSobolCols + j] = static_cast<uint32_t>(v[j] >> 32);
        }
    }

    inv.assign((kMaxLog2Res + 1) * 32, 0ull);
    delta.assign((kMaxLog2Res + 1) * 64, 0ull);
    for (uint32_t m = 1; m <= kMaxLog2Res; ++m) {
        const uint32_t n = 2 * m;
        // image of index bit j in the 2m output bits: (top m bits of dim 0) << m | top m of dim 1
        auto image = [&](uint32_t j) -> uint64_t {
            return (static_cast<uint64_t>(mat[j] >> (32 - m)) << m) | (mat[kSobolCols + j] >> (32 - m));
        };
        // Invert A (columns = image(j), j < n) by elimination on [A | I] kept column-wise:
        // cols[j] = (A column j, I column j). Reduce A's columns to unit vectors.
        uint64_t colA[32], colI[32];
        for (uint32_t j = 0; j < n; ++j) colA[j] = image(j), colI[j] = 1ull << j;
        for (uint32_t r = 0; r < n; ++r) {
            uint32_t p = r;
            while (p < n && !((colA[p] >> r) & 1ull)) ++p;
            if (p == n) throw std::runtime_error("Sobol interval map is singular");
            std::swap(colA[p], colA[r]);
            std::swap(colI[p], colI[r]);
            for (uint32_t j = 0; j < n; ++j)
                if (j != r && ((colA[j] >> r) & 1ull)) colA[j] ^= colA[r], colI[j] ^= colI[r];
        }
        // Now column r of A is e_r and colI[r] is the set of index bits that produce output bit r.
        for (uint32_t r = 0; r < n; ++r) inv[m * 32 + r] = colI[r];
        for (uint32_t c = 0; c < 64 && n + c < kSobolCols; ++c) delta[m * 64 + c] = image(n + c);
    }
}

void buildSobolByteTable(const std::vector<uint32_t>& mat, std::vector<uint32_t>& table) {
    table.assign(static_cast<size_t>(kByteChunks) * 256 * kByteDims, 0u);
    for (uint32_t k = 0; k < kByteChunks; ++k)
        for (uint32_t b = 0; b < 256; ++b)
            for (uint32_t d = 0; d < kByteDims; ++d) {
                uint32_t v = 0;
                for (uint32_t j = 0; j < 8; ++j)
                    if (((b >> j) & 1u) && 8 * k + j < kSobolCols) v ^= mat[d * kSobolCols + 8 * k + j];
                table[(static_cast<size_t>(k) * 256 + b) * kByteDims + d] = v;
            }
}

}  // namespace pg

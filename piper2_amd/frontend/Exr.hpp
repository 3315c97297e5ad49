// OpenEXR output for rendered frames (the reference saves with Imf::RgbaOutputFile in WRITE_RGB
// mode, PiperRenderer.cpp:271-291): a single-part scanline file with HALF channels B, G, R (stored
// in name order), one line per block, no compression. The layout follows the published OpenEXR
// file-format description: magic 20000630, version 2, the eight required header attributes, the
// line-offset table, then (y, byte count, channel planes) per line.
#pragma once
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace PiperFrontend {

// IEEE 754 binary32 -> binary16, round to nearest even; overflow gives infinity, NaN stays NaN.
inline uint16_t floatToHalf(float value) {
    uint32_t x;
    std::memcpy(&x, &value, 4);
    const uint16_t sign = static_cast<uint16_t>((x >> 16) & 0x8000u);
    const uint32_t a = x & 0x7FFFFFFFu;
    if (a >= 0x7F800000u) return static_cast<uint16_t>(sign | (a > 0x7F800000u ? 0x7E00u : 0x7C00u));
    if (a >= 0x477FF000u) return static_cast<uint16_t>(sign | 0x7C00u);  // >= 65520 rounds past 65504
    if (a >= 0x38800000u) {                                                 // normal half
        uint32_t h = (a >> 13) - (112u << 10);                             // rebias 127 -> 15
        const uint32_t rest = a & 0x1FFFu;
        if (rest > 0x1000u || (rest == 0x1000u && (h & 1u))) ++h;
        return static_cast<uint16_t>(sign | h);
    }
    if (a < 0x33000000u) return sign;  // below 2^-25: rounds to zero
    // subnormal half: m * 2^-24 with m = significand >> (126 - exponent)
    const uint32_t e = a >> 23;
    const uint32_t sig = (a & 0x7FFFFFu) | 0x800000u;
    const uint32_t shift = 126u - e;
    uint32_t m = sig >> shift;
    const uint32_t rest = sig & ((1u << shift) - 1u), halfway = 1u << (shift - 1u);
    if (rest > halfway || (rest == halfway && (m & 1u))) ++m;
    return static_cast<uint16_t>(sign | m);
}

namespace detail {
struct ByteWriter {
    std::vector<uint8_t> b;
    void u8(uint8_t v) { b.push_back(v); }
    void i32(int32_t v) {
        for (int k = 0; k < 4; ++k) b.push_back(static_cast<uint8_t>((static_cast<uint32_t>(v) >> (8 * k)) & 0xFFu));
    }
    void u64(uint64_t v) {
        for (int k = 0; k < 8; ++k) b.push_back(static_cast<uint8_t>((v >> (8 * k)) & 0xFFu));
    }
    void f32(float v) {
        uint32_t u;
        std::memcpy(&u, &v, 4);
        i32(static_cast<int32_t>(u));
    }
    void str(const std::string& s) {
        b.insert(b.end(), s.begin(), s.end());
        b.push_back(0);
    }
    void attribute(const std::string& name, const std::string& type, const std::vector<uint8_t>& value) {
        str(name);
        str(type);
        i32(static_cast<int32_t>(value.size()));
        b.insert(b.end(), value.begin(), value.end());
    }
};
}  // namespace detail

// Pixels whose value is not finite or is negative are written as (1, 0, 0), as the reference's
// saveImage does; returns how many there were.
inline size_t writeExrRgb(const std::string& path, uint32_t width, uint32_t height,
                          const std::vector<std::array<float, 3>>& rgb) {
    if (width == 0 || height == 0 || rgb.size() != static_cast<size_t>(width) * height)
        throw std::runtime_error("writeExrRgb: pixel count does not match the image size");
    detail::ByteWriter w;
    w.i32(20000630);  // magic
    w.i32(2);         // version 2, single-part scanline
    {
        detail::ByteWriter ch;
        for (const char* name : {"B", "G", "R"}) {
            ch.str(name);
            ch.i32(1);  // HALF
            ch.u8(0);   // pLinear
            ch.u8(0), ch.u8(0), ch.u8(0);
            ch.i32(1), ch.i32(1);  // x / y sampling
        }
        ch.u8(0);
        w.attribute("channels", "chlist", ch.b);
    }
    w.attribute("compression", "compression", {0});  // NO_COMPRESSION
    detail::ByteWriter box;
    box.i32(0), box.i32(0), box.i32(static_cast<int32_t>(width) - 1), box.i32(static_cast<int32_t>(height) - 1);
    w.attribute("dataWindow", "box2i", box.b);
    w.attribute("displayWindow", "box2i", box.b);
    w.attribute("lineOrder", "lineOrder", {0});  // INCREASING_Y
    detail::ByteWriter one, center;
    one.f32(1.0f);
    center.f32(0.0f), center.f32(0.0f);
    w.attribute("pixelAspectRatio", "float", one.b);
    w.attribute("screenWindowCenter", "v2f", center.b);
    w.attribute("screenWindowWidth", "float", one.b);
    w.u8(0);  // end of header

    const size_t lineBytes = static_cast<size_t>(width) * 3 * 2;
    const uint64_t tableAt = w.b.size(), first = tableAt + 8ull * height;
    for (uint32_t y = 0; y < height; ++y) w.u64(first + static_cast<uint64_t>(y) * (8 + lineBytes));
    size_t invalid = 0;
    std::vector<uint16_t> plane(static_cast<size_t>(width) * 3);
    for (uint32_t y = 0; y < height; ++y) {
        for (uint32_t x = 0; x < width; ++x) {
            std::array<float, 3> p = rgb[static_cast<size_t>(y) * width + x];
            if (!(std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]) &&
                  std::fmin(p[0], std::fmin(p[1], p[2])) >= 0.0f)) {
                p = {1.0f, 0.0f, 0.0f};
                ++invalid;
            }
            plane[x] = floatToHalf(p[2]);              // B
            plane[width + x] = floatToHalf(p[1]);      // G
            plane[2 * width + x] = floatToHalf(p[0]);  // R
        }
        w.i32(static_cast<int32_t>(y));
        w.i32(static_cast<int32_t>(lineBytes));
        for (uint16_t h : plane) w.u8(static_cast<uint8_t>(h & 0xFFu)), w.u8(static_cast<uint8_t>(h >> 8));
    }
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + path);
    f.write(reinterpret_cast<const char*>(w.b.data()), static_cast<std::streamsize>(w.b.size()));
    if (!f) throw std::runtime_error("cannot write " + path);
    return invalid;
}

}  // namespace PiperFrontend

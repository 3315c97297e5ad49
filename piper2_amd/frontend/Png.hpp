// PNG reader for sampled textures (the reference decodes images with stb_image,
// SampledTexture.cpp:132-152): 8-bit greyscale, grey + alpha, RGB and RGBA, non-interlaced, all
// five scanline filters, zlib/DEFLATE decoded here (RFC 1950 / 1951: stored, fixed and dynamic
// Huffman blocks). Palette, 16-bit and interlaced files are rejected with a message.
#pragma once
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace PiperFrontend {

namespace png_detail {

struct BitReader {
    const uint8_t* p;
    size_t n, pos = 0;
    uint32_t bit = 0;
    uint32_t bits(uint32_t count) {  // LSB-first
        uint32_t v = 0;
        for (uint32_t i = 0; i < count; ++i) {
            if (pos >= n) throw std::runtime_error("PNG: truncated compressed data");
            v |= static_cast<uint32_t>((p[pos] >> bit) & 1u) << i;
            if (++bit == 8) bit = 0, ++pos;
        }
        return v;
    }
    void alignByte() {
        if (bit) bit = 0, ++pos;
    }
};

// canonical Huffman decoding table: counts per length and symbols in code order
struct Huffman {
    uint16_t count[16] = {};
    std::vector<uint16_t> symbol;
    void build(const uint8_t* lengths, size_t n) {
        symbol.assign(n, 0);
        for (auto& c : count) c = 0;
        for (size_t i = 0; i < n; ++i) ++count[lengths[i]];
        count[0] = 0;
        uint16_t offs[16] = {};
        for (int l = 1; l < 16; ++l) offs[l] = static_cast<uint16_t>(offs[l - 1] + count[l - 1]);
        for (size_t i = 0; i < n; ++i)
            if (lengths[i]) symbol[offs[lengths[i]]++] = static_cast<uint16_t>(i);
    }
    uint32_t decode(BitReader& br) const {
        int code = 0, first = 0, index = 0;
        for (int l = 1; l < 16; ++l) {
            code |= static_cast<int>(br.bits(1));
            const int c = count[l];
            if (code - c < first) return symbol[static_cast<size_t>(index + (code - first))];
            index += c;
            first += c;
            first <<= 1;
            code <<= 1;
        }
        throw std::runtime_error("PNG: bad Huffman code");
    }
};

inline std::vector<uint8_t> inflate(const std::vector<uint8_t>& z) {
    if (z.size() < 2 || (z[0] & 0x0F) != 8 || ((z[0] << 8) | z[1]) % 31 != 0 || (z[1] & 0x20))
        throw std::runtime_error("PNG: bad zlib header");
    BitReader br{z.data() + 2, z.size() - 2};
    std::vector<uint8_t> out;
    static const uint16_t lenBase[29] = {3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27, 31, 35, 43, 51, 59,
                                         67, 83, 99, 115, 131, 163, 195, 227, 258};
    static const uint8_t lenExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
    static const uint16_t distBase[30] = {1, 2, 3, 4, 5, 7, 9, 13, 17, 25, 33, 49, 65, 97, 129, 193, 257, 385, 513, 769,
                                          1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
    static const uint8_t distExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};
    bool last = false;
    while (!last) {
        last = br.bits(1) != 0;
        const uint32_t type = br.bits(2);
        if (type == 0) {  // stored
            br.alignByte();
            if (br.pos + 4 > br.n) throw std::runtime_error("PNG: truncated stored block");
            const uint32_t len = br.p[br.pos] | (br.p[br.pos + 1] << 8);
            br.pos += 4;
            if (br.pos + len > br.n) throw std::runtime_error("PNG: truncated stored block");
            out.insert(out.end(), br.p + br.pos, br.p + br.pos + len);
            br.pos += len;
            continue;
        }
        if (type == 3) throw std::runtime_error("PNG: bad block type");
        Huffman lit, dist;
        if (type == 1) {  // fixed codes
            uint8_t l[288], d[30];
            for (int i = 0; i < 288; ++i) l[i] = i < 144 ? 8 : i < 256 ? 9 : i < 280 ? 7 : 8;
            for (auto& x : d) x = 5;
            lit.build(l, 288);
            dist.build(d, 30);
        } else {  // dynamic codes
            const uint32_t hlit = br.bits(5) + 257, hdist = br.bits(5) + 1, hclen = br.bits(4) + 4;
            static const uint8_t order[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
            uint8_t cl[19] = {};
            for (uint32_t i = 0; i < hclen; ++i) cl[order[i]] = static_cast<uint8_t>(br.bits(3));
            Huffman clh;
            clh.build(cl, 19);
            std::vector<uint8_t> lengths;
            while (lengths.size() < hlit + hdist) {
                const uint32_t sym = clh.decode(br);
                if (sym < 16) {
                    lengths.push_back(static_cast<uint8_t>(sym));
                } else if (sym == 16) {
                    if (lengths.empty()) throw std::runtime_error("PNG: repeat with no previous length");
                    const uint8_t prev = lengths.back();
                    for (uint32_t r = 3 + br.bits(2); r; --r) lengths.push_back(prev);
                } else {
                    const uint32_t r = sym == 17 ? 3 + br.bits(3) : 11 + br.bits(7);
                    lengths.insert(lengths.end(), r, 0);
                }
            }
            if (lengths.size() != hlit + hdist) throw std::runtime_error("PNG: bad code lengths");
            lit.build(lengths.data(), hlit);
            dist.build(lengths.data() + hlit, hdist);
        }
        while (true) {
            const uint32_t sym = lit.decode(br);
            if (sym < 256) {
                out.push_back(static_cast<uint8_t>(sym));
            } else if (sym == 256) {
                break;
            } else {
                if (sym > 285) throw std::runtime_error("PNG: bad length symbol");
                const uint32_t len = lenBase[sym - 257] + br.bits(lenExtra[sym - 257]);
                const uint32_t ds = dist.decode(br);
                if (ds > 29) throw std::runtime_error("PNG: bad distance symbol");
                const uint32_t d = distBase[ds] + br.bits(distExtra[ds]);
                if (d > out.size()) throw std::runtime_error("PNG: distance before the start");
                const size_t from = out.size() - d;
                for (uint32_t k = 0; k < len; ++k) out.push_back(out[from + k]);
            }
        }
    }
    return out;
}

inline uint32_t be32(const uint8_t* p) { return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | p[3]; }

}  // namespace png_detail

struct DecodedImage {
    uint32_t width = 0, height = 0, channels = 0;
    std::vector<uint8_t> texels;  // rows top to bottom, channels interleaved
};

inline DecodedImage loadPng(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open image " + path);
    std::vector<uint8_t> d((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    if (d.size() < 8 || !std::equal(sig, sig + 8, d.begin())) throw std::runtime_error("not a PNG file: " + path);
    DecodedImage img;
    uint32_t depth = 0, colour = 0, interlace = 0;
    std::vector<uint8_t> idat;
    for (size_t p = 8; p + 12 <= d.size();) {
        const uint32_t len = png_detail::be32(&d[p]);
        if (p + 12 + static_cast<size_t>(len) > d.size()) throw std::runtime_error("truncated PNG chunk in " + path);
        const std::string type(reinterpret_cast<const char*>(&d[p + 4]), 4);
        const uint8_t* body = &d[p + 8];
        if (type == "IHDR") {
            img.width = png_detail::be32(body), img.height = png_detail::be32(body + 4);
            depth = body[8], colour = body[9], interlace = body[12];
        } else if (type == "IDAT") {
            idat.insert(idat.end(), body, body + len);
        } else if (type == "IEND") {
            break;
        }
        p += 12 + len;
    }
    const uint32_t ch = colour == 0 ? 1 : colour == 4 ? 2 : colour == 2 ? 3 : colour == 6 ? 4 : 0;
    if (depth != 8 || ch == 0 || interlace != 0 || img.width == 0 || img.height == 0)
        throw std::runtime_error("unsupported PNG (8-bit grey / grey-alpha / RGB / RGBA, not interlaced): " + path);
    const std::vector<uint8_t> raw = png_detail::inflate(idat);
    const size_t stride = static_cast<size_t>(img.width) * ch;
    if (raw.size() < (stride + 1) * img.height) throw std::runtime_error("PNG: too little image data in " + path);
    img.channels = ch;
    img.texels.assign(stride * img.height, 0);
    for (uint32_t y = 0; y < img.height; ++y) {  // undo the scanline filters
        const uint8_t filter = raw[y * (stride + 1)];
        const uint8_t* src = &raw[y * (stride + 1) + 1];
        uint8_t* row = &img.texels[y * stride];
        const uint8_t* up = y ? &img.texels[(y - 1) * stride] : nullptr;
        for (size_t x = 0; x < stride; ++x) {
            const int a = x >= ch ? row[x - ch] : 0, b = up ? up[x] : 0, c = (up && x >= ch) ? up[x - ch] : 0;
            int v;
            switch (filter) {
                case 0: v = 0; break;
                case 1: v = a; break;
                case 2: v = b; break;
                case 3: v = (a + b) / 2; break;
                case 4: {  // Paeth
                    const int pp = a + b - c, pa = std::abs(pp - a), pb = std::abs(pp - b), pc = std::abs(pp - c);
                    v = (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
                    break;
                }
                default: throw std::runtime_error("PNG: bad filter type in " + path);
            }
            row[x] = static_cast<uint8_t>(src[x] + v);
        }
    }
    return img;
}

}  // namespace PiperFrontend

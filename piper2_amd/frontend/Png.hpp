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
    // the next `count` stream bits (LSB-first) without consuming them; zeros past the end
    uint32_t peek(uint32_t count) const {
        uint32_t v = 0;
        size_t q = pos;
        uint32_t b = bit;
        for (uint32_t i = 0; i < count && q < n; ++i) {
            v |= static_cast<uint32_t>((p[q] >> b) & 1u) << i;
            if (++b == 8) b = 0, ++q;
        }
        return v;
    }
    void skip(uint32_t count) {
        const size_t at = pos * 8 + bit + count;
        if (at > n * 8) throw std::runtime_error("PNG: truncated compressed data");
        pos = at / 8, bit = static_cast<uint32_t>(at % 8);
    }
    uint32_t bits(uint32_t count) {
        const uint32_t v = peek(count);
        skip(count);
        return v;
    }
    void alignByte() {
        if (bit) bit = 0, ++pos;
    }
};

// Prefix code of RFC 1951 section 3.2.2, decoded through a table indexed by the next 15 stream
// bits: codes of each length are consecutive integers in symbol order, starting where the
// shorter lengths' codes end; a code is sent most significant bit first, so in the LSB-first
// stream a code of length l occupies every table index whose low l bits are the code reversed.
struct Huffman {
    static constexpr uint32_t kMaxBits = 15;
    std::vector<uint32_t> table;  // symbol << 4 | code length; 0 where no code starts
    void build(const uint8_t* lengths, size_t n) {
        uint32_t lengthCount[kMaxBits + 1] = {};
        for (size_t i = 0; i < n; ++i) ++lengthCount[lengths[i]];
        lengthCount[0] = 0;
        uint32_t nextCode[kMaxBits + 1] = {};
        for (uint32_t l = 1, code = 0; l <= kMaxBits; ++l) {
            code = (code + lengthCount[l - 1]) << 1;
            nextCode[l] = code;
        }
        table.assign(size_t(1) << kMaxBits, 0);
        for (size_t sym = 0; sym < n; ++sym) {
            const uint32_t l = lengths[sym];
            if (!l) continue;
            const uint32_t code = nextCode[l]++;
            if (code >> l) throw std::runtime_error("PNG: over-subscribed code lengths");
            uint32_t rev = 0;
            for (uint32_t b = 0; b < l; ++b) rev |= ((code >> b) & 1u) << (l - 1 - b);
            for (uint32_t i = rev; i < (1u << kMaxBits); i += 1u << l) table[i] = static_cast<uint32_t>(sym) << 4 | l;
        }
    }
    uint32_t decode(BitReader& br) const {
        const uint32_t e = table[br.peek(kMaxBits)];
        if (!e) throw std::runtime_error("PNG: bad Huffman code");
        br.skip(e & 15u);
        return e >> 4;
    }
};

// maxOut: the decoded size the image header allows (anything larger is an error)
inline std::vector<uint8_t> inflate(const std::vector<uint8_t>& z, size_t maxOut) {
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
            const uint32_t nlen = br.p[br.pos + 2] | (br.p[br.pos + 3] << 8);
            if (nlen != (~len & 0xFFFFu)) throw std::runtime_error("PNG: stored block length check failed");
            br.pos += 4;
            if (br.pos + len > br.n) throw std::runtime_error("PNG: truncated stored block");
            out.insert(out.end(), br.p + br.pos, br.p + br.pos + len);
            br.pos += len;
            if (out.size() > maxOut) throw std::runtime_error("PNG: more image data than the header allows");
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
            if (out.size() > maxOut) throw std::runtime_error("PNG: more image data than the header allows");
        }
    }
    // RFC 1950: the Adler-32 checksum of the uncompressed data follows the last block
    br.alignByte();
    if (br.pos + 4 > br.n) throw std::runtime_error("PNG: missing zlib checksum");
    uint32_t a = 1, b = 0;
    for (const uint8_t v : out) a = (a + v) % 65521u, b = (b + a) % 65521u;
    const uint32_t want = (uint32_t(br.p[br.pos]) << 24) | (uint32_t(br.p[br.pos + 1]) << 16) |
                          (uint32_t(br.p[br.pos + 2]) << 8) | br.p[br.pos + 3];
    if (want != (b << 16 | a)) throw std::runtime_error("PNG: zlib checksum mismatch");
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
    bool header = false;
    std::vector<uint8_t> idat;
    for (size_t p = 8; p + 12 <= d.size();) {
        const uint32_t len = png_detail::be32(&d[p]);
        if (p + 12 + static_cast<size_t>(len) > d.size()) throw std::runtime_error("truncated PNG chunk in " + path);
        const std::string type(reinterpret_cast<const char*>(&d[p + 4]), 4);
        const uint8_t* body = &d[p + 8];
        if ((type == "IHDR") == header) throw std::runtime_error("PNG: IHDR must be the first chunk and appear once: " + path);
        if (type == "IHDR") {
            if (len != 13) throw std::runtime_error("PNG: IHDR of " + std::to_string(len) + " bytes in " + path);
            img.width = png_detail::be32(body), img.height = png_detail::be32(body + 4);
            depth = body[8], colour = body[9], interlace = body[12];
            header = true;
        } else if (type == "IDAT") {
            idat.insert(idat.end(), body, body + len);
        } else if (type == "IEND") {
            break;
        }
        p += 12 + len;
    }
    const uint32_t ch = colour == 0 ? 1 : colour == 4 ? 2 : colour == 2 ? 3 : colour == 6 ? 4 : 0;
    if (!header) throw std::runtime_error("PNG without IHDR: " + path);
    if (depth != 8 || ch == 0 || interlace != 0 || img.width == 0 || img.height == 0)
        throw std::runtime_error("unsupported PNG (8-bit grey / grey-alpha / RGB / RGBA, not interlaced): " + path);
    // the PNG limit is 2^31 - 1 per side; the texel count is held to 2^28 so no size below wraps
    constexpr uint64_t kMaxTexelBytes = uint64_t(1) << 30;
    if (img.width > 0x7FFFFFFFu || img.height > 0x7FFFFFFFu ||
        static_cast<uint64_t>(img.width) * img.height * ch > kMaxTexelBytes)
        throw std::runtime_error("PNG image too large (" + std::to_string(img.width) + " x " +
                                 std::to_string(img.height) + "): " + path);
    const size_t stride = static_cast<size_t>(img.width) * ch;
    const std::vector<uint8_t> raw = png_detail::inflate(idat, (stride + 1) * img.height);
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

// 8-bit image files for sampled textures (the reference loads them through stb_image,
// SampledTexture.cpp:120-152): PNG (Png.hpp) and binary netpbm (P5 grey, P6 RGB, P7 PAM with DEPTH
// 1..4). Three channels are widened to four with alpha 255, as the reference asks stb_image for
// RGBA when a file holds RGB (SampledTexture.cpp:136-138).
#pragma once
#include <cctype>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "Png.hpp"

namespace PiperFrontend {

inline DecodedImage widenRgb(DecodedImage img) {
    if (img.channels != 3) return img;
    std::vector<uint8_t> rgba(static_cast<size_t>(img.width) * img.height * 4);
    for (size_t i = 0; i < static_cast<size_t>(img.width) * img.height; ++i) {
        std::memcpy(&rgba[4 * i], &img.texels[3 * i], 3);
        rgba[4 * i + 3] = 255;
    }
    img.texels = std::move(rgba);
    img.channels = 4;
    return img;
}

inline DecodedImage loadNetpbm(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open image " + path);
    auto token = [&]() {
        std::string t;
        char c;
        while (f.get(c)) {
            if (c == '#') {
                std::string skip;
                std::getline(f, skip);
                continue;
            }
            if (std::isspace(static_cast<unsigned char>(c))) {
                if (!t.empty()) break;
                continue;
            }
            t += c;
        }
        return t;
    };
    auto number = [&](const std::string& t) {
        if (t.empty() || t.size() > 9 || t.find_first_not_of("0123456789") != std::string::npos)
            throw std::runtime_error("bad number in the header of " + path);
        return static_cast<uint32_t>(std::stoul(t));
    };
    const std::string magic = token();
    DecodedImage img;
    uint32_t maxval = 0;
    if (magic == "P5" || magic == "P6") {
        img.width = number(token()), img.height = number(token());
        maxval = number(token());
        img.channels = magic == "P5" ? 1 : 3;
    } else if (magic == "P7") {
        for (std::string t = token(); t != "ENDHDR"; t = token()) {
            if (t.empty()) throw std::runtime_error("truncated PAM header in " + path);
            if (t == "WIDTH") img.width = number(token());
            else if (t == "HEIGHT") img.height = number(token());
            else if (t == "DEPTH") img.channels = number(token());
            else if (t == "MAXVAL") maxval = number(token());
            else if (t == "TUPLTYPE") (void)token();
        }
    } else {
        throw std::runtime_error("unsupported image format (PNG, binary PPM, PGM and PAM are read): " + path);
    }
    if (maxval != 255 || img.width == 0 || img.height == 0 || img.channels < 1 || img.channels > 4)
        throw std::runtime_error("unsupported image layout in " + path);
    img.texels.resize(static_cast<size_t>(img.width) * img.height * img.channels);
    if (!f.read(reinterpret_cast<char*>(img.texels.data()), static_cast<std::streamsize>(img.texels.size())))
        throw std::runtime_error("truncated image data in " + path);
    return img;
}

// An image file by its extension (.png, otherwise netpbm), RGB widened to RGBA.
inline DecodedImage loadImage(const std::string& path) {
    std::string ext;
    const size_t dot = path.find_last_of('.');
    if (dot != std::string::npos) ext = path.substr(dot);
    for (char& ch : ext) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
    return widenRgb(ext == ".png" ? loadPng(path) : loadNetpbm(path));
}

}  // namespace PiperFrontend

// Dependency-free triangle-mesh loaders, standing in for the Assimp import of
// Geometry/TriangleMesh (TriangleMesh.cpp:47-119, aiProcess_Triangulate): Wavefront OBJ and PLY
// (ascii and binary_little_endian), positions and the first set of texture coordinates. An OBJ
// whose faces name texture coordinates gets one vertex per distinct (position, texcoord) pair, as
// an importer that joins identical vertices produces. Polygons are fanned into triangles around
// their first vertex, which is what a triangulating importer does for convex faces. Errors throw
// std::runtime_error naming the file and line.
#pragma once
#include <cstdint>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace PiperFrontend {

struct TriangleMeshData {
    std::vector<float> verts;     // float3 per vertex
    std::vector<uint32_t> index;  // uint3 per triangle
    std::vector<float> uv;        // float2 per vertex, or empty
};

namespace detail {
inline std::string lower(std::string s) {
    for (char& c : s)
        if (c >= 'A' && c <= 'Z') c = static_cast<char>(c - 'A' + 'a');
    return s;
}
inline bool endsWith(const std::string& s, const std::string& suffix) {
    return s.size() >= suffix.size() && s.compare(s.size() - suffix.size(), suffix.size(), suffix) == 0;
}
}  // namespace detail

// OBJ: "v x y z [w]" and "f a b c ..." where each corner is v, v/vt, v//vn or v/vt/vn; indices
// are 1-based, negative ones count back from the latest vertex. Other statements are ignored.
inline TriangleMeshData loadObj(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open mesh " + path);
    std::vector<float> pos, tex;                 // as listed in the file
    std::vector<std::pair<long, long>> corners;  // (position, texcoord or -1) per triangle corner
    bool anyTex = false;
    std::string line;
    size_t lineNo = 0;
    std::vector<std::pair<long, long>> poly;
    auto fail = [&](const std::string& msg) { throw std::runtime_error(path + ":" + std::to_string(lineNo) + ": " + msg); };
    auto resolve = [&](const std::string& s, long n) -> long {
        char* end = nullptr;
        const long k = std::strtol(s.c_str(), &end, 10);
        if (s.empty() || *end != '\0' || k == 0) fail("bad face corner index \"" + s + "\"");
        const long idx = k > 0 ? k - 1 : n + k;
        if (idx < 0 || idx >= n) fail("face index out of range");
        return idx;
    };
    while (std::getline(f, line)) {
        ++lineNo;
        std::istringstream in(line);
        std::string tag;
        if (!(in >> tag)) continue;
        if (tag == "v") {
            float x, y, z;
            if (!(in >> x >> y >> z)) fail("bad vertex");
            pos.insert(pos.end(), {x, y, z});
        } else if (tag == "vt") {
            float u, v = 0.0f;
            if (!(in >> u)) fail("bad texture coordinate");
            in >> v;
            tex.insert(tex.end(), {u, v});
        } else if (tag == "f") {
            poly.clear();
            std::string corner;
            while (in >> corner) {
                const size_t s1 = corner.find('/');
                const long p = resolve(corner.substr(0, s1), static_cast<long>(pos.size() / 3));
                long t = -1;
                if (s1 != std::string::npos) {
                    const size_t s2 = corner.find('/', s1 + 1);
                    const std::string ts = corner.substr(s1 + 1, s2 == std::string::npos ? std::string::npos : s2 - s1 - 1);
                    if (!ts.empty()) {
                        t = resolve(ts, static_cast<long>(tex.size() / 2));
                        anyTex = true;
                    }
                }
                poly.push_back({p, t});
            }
            if (poly.size() < 3) fail("face with fewer than 3 corners");
            for (size_t t = 1; t + 1 < poly.size(); ++t) corners.insert(corners.end(), {poly[0], poly[t], poly[t + 1]});
        }
    }
    TriangleMeshData m;
    if (!anyTex) {  // positions as listed, faces index them directly
        m.verts = std::move(pos);
        for (const auto& c : corners) m.index.push_back(static_cast<uint32_t>(c.first));
        return m;
    }
    std::map<std::pair<long, long>, uint32_t> lookup;  // one vertex per distinct (position, texcoord)
    for (const auto& c : corners) {
        auto it = lookup.find(c);
        if (it == lookup.end()) {
            it = lookup.emplace(c, static_cast<uint32_t>(m.verts.size() / 3)).first;
            m.verts.insert(m.verts.end(), {pos[3 * c.first], pos[3 * c.first + 1], pos[3 * c.first + 2]});
            if (c.second >= 0) m.uv.insert(m.uv.end(), {tex[2 * c.second], tex[2 * c.second + 1]});
            else m.uv.insert(m.uv.end(), {0.0f, 0.0f});
        }
        m.index.push_back(it->second);
    }
    return m;
}

// PLY with an "element vertex" carrying float/double x, y, z (other properties skipped) and an
// "element face" whose first property is a list of vertex indices.
inline TriangleMeshData loadPly(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open mesh " + path);
    auto fail = [&](const std::string& msg) { throw std::runtime_error(path + ": " + msg); };
    std::string line;
    if (!std::getline(f, line) || line.rfind("ply", 0) != 0) fail("not a PLY file");
    struct Prop {
        std::string name, type, countType;  // countType set for lists
    };
    struct Element {
        std::string name;
        size_t count = 0;
        std::vector<Prop> props;
    };
    std::vector<Element> elems;
    std::string format;
    while (std::getline(f, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        std::istringstream in(line);
        std::string tag;
        in >> tag;
        if (tag == "format") {
            in >> format;
        } else if (tag == "element") {
            Element e;
            in >> e.name >> e.count;
            elems.push_back(e);
        } else if (tag == "property") {
            if (elems.empty()) fail("property before element");
            Prop p;
            std::string t;
            in >> t;
            if (t == "list") in >> p.countType >> p.type >> p.name;
            else p.type = t, in >> p.name;
            elems.back().props.push_back(p);
        } else if (tag == "end_header") {
            break;
        }
    }
    const bool ascii = format == "ascii";
    if (!ascii && format != "binary_little_endian") fail("unsupported PLY format " + format);
    auto sizeOf = [&](const std::string& t) -> size_t {
        if (t == "char" || t == "uchar" || t == "int8" || t == "uint8") return 1;
        if (t == "short" || t == "ushort" || t == "int16" || t == "uint16") return 2;
        if (t == "int" || t == "uint" || t == "int32" || t == "uint32" || t == "float" || t == "float32") return 4;
        if (t == "double" || t == "float64") return 8;
        fail("unknown PLY type " + t);
        return 0;
    };
    auto readBin = [&](const std::string& t) -> double {
        unsigned char b[8] = {};
        const size_t n = sizeOf(t);
        if (!f.read(reinterpret_cast<char*>(b), static_cast<std::streamsize>(n))) fail("truncated binary data");
        if (t == "char" || t == "int8") return static_cast<int8_t>(b[0]);
        if (t == "uchar" || t == "uint8") return b[0];
        if (t == "short" || t == "int16") { int16_t v; std::memcpy(&v, b, 2); return v; }
        if (t == "ushort" || t == "uint16") { uint16_t v; std::memcpy(&v, b, 2); return v; }
        if (t == "int" || t == "int32") { int32_t v; std::memcpy(&v, b, 4); return v; }
        if (t == "uint" || t == "uint32") { uint32_t v; std::memcpy(&v, b, 4); return v; }
        if (t == "float" || t == "float32") { float v; std::memcpy(&v, b, 4); return v; }
        double v;
        std::memcpy(&v, b, 8);
        return v;
    };
    std::istringstream* text = nullptr;
    std::string body;
    std::istringstream asciiIn;
    if (ascii) {
        std::ostringstream rest;
        rest << f.rdbuf();
        body = rest.str();
        asciiIn.str(body);
        text = &asciiIn;
    }
    auto next = [&](const std::string& t) -> double {
        if (!ascii) return readBin(t);
        double v;
        if (!(*text >> v)) fail("truncated ascii data");
        return v;
    };
    TriangleMeshData m;
    size_t nv = 0;
    for (const Element& e : elems) {
        if (e.name == "vertex") {
            int ix = -1, iy = -1, iz = -1, iu = -1, iv = -1;
            for (size_t k = 0; k < e.props.size(); ++k) {
                const std::string& n = e.props[k].name;
                if (n == "x") ix = static_cast<int>(k);
                if (n == "y") iy = static_cast<int>(k);
                if (n == "z") iz = static_cast<int>(k);
                if (n == "u" || n == "s" || n == "texture_u" || n == "texture_s") iu = static_cast<int>(k);
                if (n == "v" || n == "t" || n == "texture_v" || n == "texture_t") iv = static_cast<int>(k);
            }
            const bool hasUV = iu >= 0 && iv >= 0;
            if (hasUV) m.uv.resize(2 * e.count);
            if (ix < 0 || iy < 0 || iz < 0) fail("vertex element without x, y, z");
            nv = e.count;
            m.verts.resize(3 * nv);
            std::vector<double> row(e.props.size());
            for (size_t v = 0; v < e.count; ++v) {
                for (size_t k = 0; k < e.props.size(); ++k) {
                    if (!e.props[k].countType.empty()) fail("list property in vertex element");
                    row[k] = next(e.props[k].type);
                }
                m.verts[3 * v] = static_cast<float>(row[static_cast<size_t>(ix)]);
                m.verts[3 * v + 1] = static_cast<float>(row[static_cast<size_t>(iy)]);
                m.verts[3 * v + 2] = static_cast<float>(row[static_cast<size_t>(iz)]);
                if (hasUV) {
                    m.uv[2 * v] = static_cast<float>(row[static_cast<size_t>(iu)]);
                    m.uv[2 * v + 1] = static_cast<float>(row[static_cast<size_t>(iv)]);
                }
            }
        } else if (e.name == "face") {
            if (e.props.empty() || e.props[0].countType.empty()) fail("face element without an index list");
            std::vector<uint32_t> poly;
            for (size_t fi = 0; fi < e.count; ++fi) {
                for (size_t k = 0; k < e.props.size(); ++k) {
                    const Prop& p = e.props[k];
                    if (p.countType.empty()) {
                        (void)next(p.type);
                        continue;
                    }
                    const size_t n = static_cast<size_t>(next(p.countType));
                    poly.clear();
                    for (size_t c = 0; c < n; ++c) {
                        const double idx = next(p.type);
                        if (k == 0) {
                            if (idx < 0 || idx >= static_cast<double>(nv)) fail("face index out of range");
                            poly.push_back(static_cast<uint32_t>(idx));
                        }
                    }
                    if (k == 0) {
                        if (poly.size() < 3) fail("face with fewer than 3 corners");
                        for (size_t t = 1; t + 1 < poly.size(); ++t)
                            m.index.insert(m.index.end(), {poly[0], poly[t], poly[t + 1]});
                    }
                }
            }
        } else {  // skip an element we do not use
            for (size_t r = 0; r < e.count; ++r)
                for (const Prop& p : e.props) {
                    if (p.countType.empty()) {
                        (void)next(p.type);
                    } else {
                        const size_t n = static_cast<size_t>(next(p.countType));
                        for (size_t c = 0; c < n; ++c) (void)next(p.type);
                    }
                }
        }
    }
    return m;
}

inline TriangleMeshData loadMesh(const std::string& path) {
    const std::string l = detail::lower(path);
    if (detail::endsWith(l, ".obj")) return loadObj(path);
    if (detail::endsWith(l, ".ply")) return loadPly(path);
    throw std::runtime_error("unsupported mesh format (OBJ and PLY are read): " + path);
}

}  // namespace PiperFrontend

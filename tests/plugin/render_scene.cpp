// Test driver for the C++ component pair: renders a scene file through HIPTracer / HIPAccelerator
// the way Piper2 drives a Tracer plugin (PiperRenderer builds the scene graph, Tracer::buildPipeline,
// Pipeline::prepare, FixedSampler::renderFrame), and writes the film.
//
//   render_scene SCENE OUT.rgbw [FILTER [ALPHA]]   render with the FixedSampler's FilterType (box,
//                                      triangle, gaussian); OUT holds width * height RGBW floats
//   render_scene --errors SCENE        exercise the error paths; prints one line per case
//
// Scene file (little endian): "PGSC", u32 version 1, u32 nv nt no nm nq width height spp maxDepth
// minRR compatFlags fitMode, f32 vertices[3 nv], u32 indices[3 nt], u32 triObject[nt],
// u32 objectMaterial[no], f32 materials[8 nm] (type, rgb, alpha, specular rgb), f32 quads[12 nq]
// (origin, U, V, radiance), f32 camera[14] (position, lookAt, up, sensor w h, focal, aperture, 0),
// f32 env[3], f32 srt[16 no] (TransformSRT per object, applied to its triangles).
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>

#include "PiperHIP.hpp"

using namespace Piper;

namespace {

struct SceneFile {
    uint32_t nv, nt, no, nm, nq, width, height, spp, maxDepth, minRR, compat, fit;
    std::vector<float> verts, materials, quads, camera, env, srt, points;
    std::vector<uint32_t> index, triObject, objectMaterial;
};

template <typename T>
void readArr(std::ifstream& f, std::vector<T>& v, size_t n) {
    v.resize(n);
    f.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(n * sizeof(T)));
}

SceneFile load(const char* path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw PiperException(std::string("cannot open ") + path);
    char magic[4];
    uint32_t version;
    f.read(magic, 4);
    f.read(reinterpret_cast<char*>(&version), 4);
    if (std::string(magic, 4) != "PGSC" || version != 1) throw PiperException("not a scene file");
    SceneFile s{};
    uint32_t* h = &s.nv;
    f.read(reinterpret_cast<char*>(h), 12 * 4);
    readArr(f, s.verts, 3ull * s.nv);
    readArr(f, s.index, 3ull * s.nt);
    readArr(f, s.triObject, s.nt);
    readArr(f, s.objectMaterial, s.no);
    readArr(f, s.materials, 8ull * s.nm);
    readArr(f, s.quads, 12ull * s.nq);
    readArr(f, s.camera, 14);
    readArr(f, s.env, 3);
    readArr(f, s.srt, 16ull * s.no);
    if (!f) throw PiperException("truncated scene file");
    uint32_t np = 0;  // optional tail: point lights {position, intensity}
    if (f.read(reinterpret_cast<char*>(&np), 4)) readArr(f, s.points, 6ull * np);
    if (np && !f) throw PiperException("truncated scene file");
    return s;
}

TransformInfo staticTransform(const TransformSRT& t) {
    TransformInfo info;
    info.transforms.push_back({0u, t});
    return info;
}

struct Built {
    std::shared_ptr<Node> root, sensor;
};

// One Geometry per object (its triangles in file order, local vertices), one GSMInstance with the
// object's material, one child with the object's transform; then lights, environment, sensor.
Built buildScene(HIPTracer& tracer, const SceneFile& s) {
    Accelerator& acc = tracer.getAccelerator();
    std::vector<std::pair<TransformInfo, std::shared_ptr<Node>>> children;
    for (uint32_t o = 0; o < s.no; ++o) {
        std::vector<float> verts;
        std::vector<uint32_t> index;
        std::map<uint32_t, uint32_t> local;
        for (uint32_t t = 0; t < s.nt; ++t) {
            if (s.triObject[t] != o) continue;
            for (int k = 0; k < 3; ++k) {
                const uint32_t v = s.index[3 * t + k];
                auto it = local.find(v);
                if (it == local.end()) {
                    it = local.emplace(v, static_cast<uint32_t>(verts.size() / 3)).first;
                    verts.insert(verts.end(), &s.verts[3 * v], &s.verts[3 * v] + 3);
                }
                index.push_back(it->second);
            }
        }
        if (index.empty()) continue;
        const size_t vBytes = verts.size() * 4, iOff = (vBytes + 15) / 16 * 16, size = iOff + index.size() * 4;
        auto buffer = acc.createBuffer(size, 16, [&](std::byte* p) {
            std::memcpy(p, verts.data(), vBytes);
            std::memcpy(p + iOff, index.data(), index.size() * 4);
        });
        GeometryDesc gd;
        gd.desc = TriangleIndexedGeometryDesc{static_cast<uint32_t>(verts.size() / 3), static_cast<uint32_t>(index.size() / 3),
                                              buffer, size, 0, iOff};
        auto geometry = std::make_shared<Geometry>(gd);
        const float* m = &s.materials[8 * s.objectMaterial[o]];
        std::shared_ptr<Surface> surface;
        if (m[0] == 2.0f) {
            surface = Surface::plastic({m[1], m[2], m[3]}, {m[5], m[6], m[7]}, m[4]);
        } else if (m[0] == 3.0f) {
            surface = Surface::mirror({m[1], m[2], m[3]});
        } else if (m[0] == 4.0f) {
            surface = Surface::glass({m[1], m[2], m[3]}, {m[5], m[6], m[7]}, m[4], m[4]);
        } else if (m[0] == 5.0f) {
            surface = Surface::substrate({m[1], m[2], m[3]}, {m[5], m[6], m[7]}, m[4], m[4]);
        } else {
            surface = std::make_shared<Surface>();  // alpha as stored (Surface::conductor squares a roughness)
            surface->kind = m[0] == 0.0f ? Surface::Kind::Matte : Surface::Kind::Conductor;
            surface->color = {m[1], m[2], m[3]};
            surface->alpha = m[4];
        }
        auto inst = tracer.buildGSMInstance(geometry, surface, nullptr);
        TransformSRT srt;
        std::memcpy(&srt, &s.srt[16 * o], sizeof(srt));
        children.push_back({staticTransform(srt), tracer.buildNode(inst)});
    }
    for (uint32_t q = 0; q < s.nq; ++q) {
        const float* p = &s.quads[12 * q];
        auto light = Light::areaPlane({p[0], p[1], p[2]}, {p[3], p[4], p[5]}, {p[6], p[7], p[8]}, {p[9], p[10], p[11]});
        children.push_back({TransformInfo{}, tracer.buildNode(light)});
    }
    for (size_t p = 0; p + 6 <= s.points.size(); p += 6) {  // a point light sits at its node's origin
        TransformSRT at = identitySRT();
        at.transX = s.points[p], at.transY = s.points[p + 1], at.transZ = s.points[p + 2];
        children.push_back({staticTransform(at), tracer.buildNode(Light::point({s.points[p + 3], s.points[p + 4], s.points[p + 5]}))});
    }
    if (s.env[0] > 0.0f || s.env[1] > 0.0f || s.env[2] > 0.0f)
        children.push_back({TransformInfo{}, tracer.buildNode(Light::environment({s.env[0], s.env[1], s.env[2]}))});
    auto sensor = std::make_shared<Sensor>();
    sensor->lookAt = {s.camera[3], s.camera[4], s.camera[5]};
    sensor->upRef = {s.camera[6], s.camera[7], s.camera[8]};
    sensor->size = {s.camera[9], s.camera[10]};
    sensor->focalLength = s.camera[11];
    sensor->apertureRadius = s.camera[12];
    TransformSRT at = identitySRT();
    at.transX = s.camera[0], at.transY = s.camera[1], at.transZ = s.camera[2];
    Built b;
    b.sensor = tracer.buildNode(sensor);
    children.push_back({staticTransform(at), b.sensor});
    b.root = tracer.buildNode(children);
    return b;
}

int render(const char* scenePath, const char* outPath, const std::string& filterType, float alpha) {
    const SceneFile s = load(scenePath);
    auto acc = std::make_shared<HIPAccelerator>(0);
    HIPTracer tracer(acc);
    std::cout << "platform: " << acc->getNativePlatform() << "\n";
    const Built b = buildScene(tracer, s);
    Integrator integrator;
    integrator.maxDepth = s.maxDepth, integrator.minDepthForRussianRoulette = s.minRR, integrator.compatFlags = s.compat;
    auto sampler = std::make_shared<Sampler>();
    sampler->samplesPerPixel = s.spp;
    FixedSampler driver;
    driver.filter.filterType = filterType;
    driver.filter.alpha = alpha;
    LightSampler lights;
    auto pipeline = tracer.buildPipeline(b.root, integrator, driver, lights, sampler);
    auto launcher = pipeline->prepare(b.sensor, s.width, s.height, s.fit ? FitMode::OverScan : FitMode::Fill);
    const RenderRECT r = launcher->getRenderRECT();
    std::cout << "rect: " << r.left << " " << r.top << " " << r.width << " " << r.height << "\n";
    const auto rgb = driver.renderFrame(tracer, *launcher);
    std::cout << pipeline->generateStatisticsReport() << "\n";
    std::ofstream out(outPath, std::ios::binary);
    out.write(reinterpret_cast<const char*>(driver.lastFilm.data()),
              static_cast<std::streamsize>(driver.lastFilm.size() * 4));
    return out ? 0 : 1;
}

// Each case must raise PiperException (the reference would hand it to raiseException).
int errors(const char* scenePath) {
    const SceneFile s = load(scenePath);
    auto acc = std::make_shared<HIPAccelerator>(0);
    HIPTracer tracer(acc);
    int raised = 0, cases = 0;
    auto expect = [&](const char* name, const std::function<void()>& f) {
        ++cases;
        try {
            f();
            std::cout << name << ": no error\n";
        } catch (const PiperException& e) {
            ++raised;
            std::cout << name << ": raised: " << e.what() << "\n";
        }
    };
    expect("unknown symbol", [&] { (void)tracer.buildProgram({"HIP-gfx950-builtin", {"orenNayarInit"}}, "orenNayarInit"); });
    expect("foreign format", [&] { (void)tracer.buildProgram({"LLVM IR", {"trace"}}, "trace"); });
    expect("motion blur", [&] {
        TransformInfo info;
        info.transforms = {{0u, identitySRT()}, {1u, identitySRT()}};
        (void)tracer.buildNode({{info, tracer.buildNode(Light::environment({1, 1, 1}))}});
    });
    expect("medium", [&] {
        auto buf = acc->createBuffer(48, 16, nullptr);
        GeometryDesc gd;
        gd.desc = TriangleIndexedGeometryDesc{3, 1, buf, 48, 0, 36};
        (void)tracer.buildGSMInstance(std::make_shared<Geometry>(gd), Surface::matte({1, 1, 1}), std::make_shared<Medium>());
    });
    expect("index out of range", [&] {
        auto buf = acc->createBuffer(48, 16, [](std::byte* p) {
            const uint32_t idx[3] = {0, 1, 7};
            std::memset(p, 0, 36);
            std::memcpy(p + 36, idx, 12);
        });
        GeometryDesc gd;
        gd.desc = TriangleIndexedGeometryDesc{3, 1, buf, 48, 0, 36};
        (void)tracer.buildAcceleration(gd);
    });
    expect("foreign output", [&] {
        const Built b = buildScene(tracer, s);
        Integrator integrator;
        auto sampler = std::make_shared<Sampler>();
        FixedSampler driver;
        LightSampler lights;
        auto pipeline = tracer.buildPipeline(b.root, integrator, driver, lights, sampler);
        auto launcher = pipeline->prepare(b.sensor, s.width, s.height, FitMode::Fill);
        const std::vector<std::shared_ptr<Resource>> res{acc->createBuffer(16, 16, nullptr)};
        (void)launcher->launch({0, 0, 8, 8}, [&](uint32_t) { return packSBTPayload(LaunchData{0, s.width, s.height}); }, res);
    });
    expect("unknown filter", [&] {
        const Built b = buildScene(tracer, s);
        Integrator integrator;
        auto sampler = std::make_shared<Sampler>();
        FixedSampler driver;
        driver.filter.filterType = "mitchell";
        LightSampler lights;
        (void)tracer.buildPipeline(b.root, integrator, driver, lights, sampler);
    });
    std::cout << "raised " << raised << " of " << cases << "\n";
    return raised == cases ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        if (argc == 3 && std::string(argv[1]) == "--errors") return errors(argv[2]);
        if (argc >= 3 && argc <= 5 && std::string(argv[1]) != "--errors")
            return render(argv[1], argv[2], argc >= 4 ? argv[3] : "box", argc == 5 ? std::stof(argv[4]) : 0.0f);
        std::cerr << "usage: render_scene SCENE OUT [FILTER [ALPHA]] | render_scene --errors SCENE\n";
        return 2;
    } catch (const PiperException& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 3;
    }
}

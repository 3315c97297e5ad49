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
#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>

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
        } else if (m[0] == 0.0f) {
            surface = Surface::matte({m[1], m[2], m[3]}, m[4]);
        } else {
            surface = Surface::conductorAlpha({m[1], m[2], m[3]}, m[4]);  // the file holds GGX alpha
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

// The host side of module loading (PiperCore.cpp:785-796): dlopen the plugin library, check its
// protocol, build the module with the host's allocator and ask it for a class.
struct HostAllocator final : Allocator {
    void* alloc(size_t size, size_t align) override { return ::operator new(size, std::align_val_t(align)); }
    void free(void* p) override { ::operator delete(p, std::align_val_t(alignof(std::max_align_t))); }
};
struct LoadedModule {
    void* handle = nullptr;
    HostAllocator allocator;
    PiperContext context;
    Module* module = nullptr;
    explicit LoadedModule(const char* path) {
        handle = dlopen(path, RTLD_NOW | RTLD_LOCAL);
        if (!handle) throw PiperException(std::string("dlopen: ") + dlerror());
        auto protocol = reinterpret_cast<const char* (*)()>(dlsym(handle, "piperGetProtocol"));
        auto init = reinterpret_cast<Module* (*)(PiperContext&, Allocator&, const char*)>(dlsym(handle, "piperInitModule"));
        if (!protocol || !init) throw PiperException("the library exports no Piper module");
        if (std::string(protocol()) != "GCC@STD@0") throw PiperException(std::string("protocol mismatch: ") + protocol());
        module = init(context, allocator, path);
    }
    ~LoadedModule() {
        module->~Module();
        allocator.free(module);
    }
};
std::string pluginPath() {
    Dl_info info{};
    if (!dladdr(reinterpret_cast<const void*>(&piperInitModule), &info) || !info.dli_fname)
        throw PiperException("cannot locate the plugin library");
    return info.dli_fname;
}

int render(const char* scenePath, const char* outPath, const std::string& filterType, float alpha, bool viaModule) {
    const SceneFile s = load(scenePath);
    std::unique_ptr<LoadedModule> mod;
    std::shared_ptr<HIPTracer> tracerRef;
    if (viaModule) {  // the pair by class name, as a Piper host loads a Tracer plugin
        mod = std::make_unique<LoadedModule>(pluginPath().c_str());
        tracerRef = std::dynamic_pointer_cast<HIPTracer>(
            mod->module->newInstance("Tracer", Config::parse(R"({"Accelerator": {"Device": 0}})")));
        if (!tracerRef) throw PiperException("the module's Tracer is not a HIPTracer");
        std::cout << "module: " << pluginPath() << " protocol " << piperGetProtocol() << "\n";
    } else {
        tracerRef = std::make_shared<HIPTracer>(std::make_shared<HIPAccelerator>(0));
    }
    HIPTracer& tracer = *tracerRef;
    std::cout << "platform: " << tracer.getAccelerator().getNativePlatform() << "\n";
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
// Payloads of every built-in component, recorded through a MaterializeContext, one line each:
// "<component> <program symbol> <payload hex>"; calls made by a payload are listed before it.
int payloads() {
    auto acc = HIPAccelerator::hostOnly();
    HIPTracer tracer(acc);
    std::vector<std::pair<std::string, SBTPayload>> calls;
    std::vector<std::shared_ptr<Resource>> resources;
    auto hex = [](const SBTPayload& p) {
        std::ostringstream o;
        for (std::byte b : p) o << std::hex << std::setw(2) << std::setfill('0') << static_cast<int>(b);
        return o.str();
    };
    std::map<const RTProgram*, std::string> programSymbol;  // symbol of each program handed out
    auto registerCall = [&](const std::shared_ptr<RTProgram>& p, const SBTPayload& payload) {
        calls.push_back({programSymbol.count(p.get()) ? programSymbol[p.get()] : "?", payload});
        std::cout << "call " << calls.size() << " " << calls.back().first << " " << hex(payload) << "\n";
        return reinterpret_cast<CallHandle>(static_cast<uintptr_t>(calls.size()));
    };
    struct RecordingTracer final : Tracer {
        HIPTracer& inner;
        std::map<const RTProgram*, std::string>& names;
        RecordingTracer(HIPTracer& t, std::map<const RTProgram*, std::string>& n) : inner(t), names(n) {}
        std::shared_ptr<AccelerationStructure> buildAcceleration(const GeometryDesc& d) override { return inner.buildAcceleration(d); }
        std::shared_ptr<GSMInstance> buildGSMInstance(std::shared_ptr<Geometry> g, std::shared_ptr<Surface> s,
                                                      std::shared_ptr<Medium> m) override {
            return inner.buildGSMInstance(g, s, m);
        }
        std::shared_ptr<Node> buildNode(const std::shared_ptr<Object>& o) override { return inner.buildNode(o); }
        std::shared_ptr<Node> buildNode(const std::vector<std::pair<TransformInfo, std::shared_ptr<Node>>>& c) override {
            return inner.buildNode(c);
        }
        std::shared_ptr<RTProgram> buildProgram(const LinkableProgram& l, std::string symbol) override {
            auto p = inner.buildProgram(l, symbol);
            names[p.get()] = symbol;
            return p;
        }
        std::unique_ptr<Pipeline> buildPipeline(const std::shared_ptr<Node>& s, Integrator& i, RenderDriver& r,
                                                LightSampler& l, std::shared_ptr<Sampler> sm) override {
            return inner.buildPipeline(s, i, r, l, sm);
        }
        std::shared_ptr<Texture> generateTexture(const Config& d, uint32_t c) const override { return inner.generateTexture(d, c); }
        size_t getAlignmentRequirement(AlignmentRequirement r) const noexcept override { return inner.getAlignmentRequirement(r); }
        Accelerator& getAccelerator() const noexcept override { return inner.getAccelerator(); }
    } rec(tracer, programSymbol);
    MaterializeContext* ctxPtr = nullptr;
    const MaterializeContext ctx{rec, *acc,
                                 [&](std::shared_ptr<Resource> r) {
                                     resources.push_back(r);
                                     return static_cast<uint32_t>(resources.size() - 1);
                                 },
                                 registerCall,
                                 [&](const Config& desc, uint32_t channel) {
                                     const auto t = rec.generateTexture(desc, channel);
                                     const auto prog = t->materialize(*ctxPtr);
                                     return registerCall(prog.sample, prog.payload);
                                 }};
    ctxPtr = const_cast<MaterializeContext*>(&ctx);
    auto show = [&](const char* what, const std::shared_ptr<RTProgram>& p, const SBTPayload& payload) {
        std::cout << what << " " << programSymbol[p.get()] << " " << hex(payload) << "\n";
    };
    const auto traversal = reinterpret_cast<TraversalHandle>(uintptr_t(0x1122334455667788ull));
    {
        auto m = Surface::matte({0.25f, 0.5f, 0.75f}, 0.125f);
        const auto p = m->materialize(ctx);
        show("matte", p.init, p.payload);
    }
    {
        auto g = Surface::glass({0.9f, 0.8f, 0.7f}, {0.6f, 0.5f, 0.4f}, 0.3f, 0.3f);
        const auto p = g->materialize(ctx);
        show("glass", p.init, p.payload);
    }
    {
        auto cb = std::make_shared<Surface>(Surface::Kind::Matte,
                                            std::vector<Config>{checkerboardConfig({0.1f, 0.2f, 0.3f, 1.0f}, {0.9f, 0.8f, 0.7f, 1.0f}, 4.0f),
                                                                scalarConfig(0.0f)});
        const auto p = cb->materialize(ctx);
        show("checkered", p.init, p.payload);
    }
    {
        auto l = Light::area({{{1, 2, 3}, {0.5f, 0, 0}, {0, 0, 2}}, {{4, 5, 6}, {0, 1.5f, 0}, {1, 0, 0}}}, {10, 20, 30});
        const auto p = l->materialize(traversal, ctx);
        show("area", p.init, p.payload);
    }
    {
        const auto p = Light::environment({0.1f, 0.2f, 0.3f})->materialize(traversal, ctx);
        show("environment", p.init, p.payload);
        const auto q = Light::point({4, 5, 6})->materialize(traversal, ctx);
        show("point", q.init, q.payload);
    }
    {
        Sensor cam;
        cam.lookAt = {1, 2, 3}, cam.upRef = {0, 1, 0}, cam.size = {0.036f, 0.024f}, cam.focalLength = 0.05f;
        cam.apertureRadius = 0.0125f;
        const auto p = cam.materialize(traversal, ctx);
        show("sensor", p.rayGen, p.payload);
    }
    {
        Integrator in;
        in.maxDepth = 12, in.minDepthForRussianRoulette = 3;
        const auto p = in.materialize(ctx);
        show("integrator", p.trace, p.payload);
    }
    {
        Sampler smp;
        smp.samplesPerPixel = 48;
        const auto a = smp.generatePayload(1920, 1080);
        std::cout << "sampler " << a.samplesPerPixel << " " << a.maxDimension << " " << hex(a.payload) << "\n";
    }
    {
        LightSampler ls;
        ls.preprocess({Light::environment({0, 0, 0}), Light::areaPlane({0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 1}),
                       Light::point({1, 1, 1})});
        const auto p = ls.materialize(ctx);
        show("lightsampler", p.select, p.payload);
    }
    {
        FixedSampler fs;
        fs.filter.filterType = "gaussian", fs.filter.alpha = 2.0f;
        const auto p = fs.materialize(ctx);
        show("driver", p.accumulate, p.payload);
    }
    std::cout << "resources " << resources.size() << "\n";
    return 0;
}

// A native kernel through the accelerator's kernel half (Accelerator.hpp:141-160): compiled from a
// "Native" table with a static redirect and a dynamic symbol, launched with trivially copyable
// arguments and a resource look-up table.
struct KernelProbe {
    Dim3 size{};
    uint32_t a = 0;
    double b = 0.0;
    size_t resources = 0;
    const void* dynamicSymbol = nullptr;
    bool context = false;
};
KernelProbe g_probe;
int g_marker = 0;
void probeKernel(const NativeLaunch& l) {
    g_probe.size = l.size;
    g_probe.a = l.args.get<uint32_t>(0);
    g_probe.b = l.args.get<double>(1);
    g_probe.resources = l.resources.size();
    g_probe.dynamicSymbol = l.dynamicSymbols.at(0);
    auto lock = l.context.makeCurrent();
    g_probe.context = true;
}
int kernel() {
    auto acc = HIPAccelerator::hostOnly();
    const LinkableProgram table = makeNativeLinkable({{"probe", reinterpret_cast<const void*>(&probeKernel)},
                                                      {"marker", &g_marker}},
                                                     7);
    auto k = acc->compileKernel({table}, {{"main", "probe"}}, {"marker"}, "main");
    auto lut = acc->createResourceLUT({acc->createBuffer(64, 16, nullptr), acc->createBuffer(32, 16, nullptr)});
    auto f = acc->launchKernel(Dim3{3, 4, 5}, k, lut, uint32_t{77}, 2.5);
    if (g_probe.a != 0) return 1;  // nothing runs before the future is polled
    f.getSync();
    const bool ok = g_probe.size.x == 3 && g_probe.size.y == 4 && g_probe.size.z == 5 && g_probe.a == 77 &&
                    g_probe.b == 2.5 && g_probe.resources == 2 && g_probe.dynamicSymbol == &g_marker && g_probe.context;
    std::cout << (ok ? "kernel ok" : "kernel mismatch") << " contexts " << acc->enumerateContexts().size() << "\n";
    return ok ? 0 : 1;
}

// Module entry points on the host: protocol, classes, an unknown class.
int module() {
    LoadedModule mod(pluginPath().c_str());
    auto acc = std::dynamic_pointer_cast<HIPAccelerator>(
        mod.module->newInstance("Accelerator", Config::parse(R"({"HostOnly": true})")));
    auto tracer = std::dynamic_pointer_cast<HIPTracer>(
        mod.module->newInstance("Tracer", Config::parse(R"({"Accelerator": {"HostOnly": true}})")));
    std::cout << "protocol " << piperGetProtocol() << "\n";
    std::cout << "accelerator " << (acc ? acc->getNativePlatform() : "none") << " formats "
              << (acc ? acc->getSupportedLinkableFormat().at(0) : "") << "\n";
    std::cout << "tracer " << (tracer ? "HIPTracer" : "none") << "\n";
    try {
        (void)mod.module->newInstance("Denoiser", Config::makeObject({}));
        std::cout << "unknown class: no error\n";
    } catch (const PiperException& e) {
        std::cout << "unknown class: raised: " << e.what() << "\n";
    }
    return acc && tracer ? 0 : 1;
}

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
    expect("unknown symbol", [&] {
        int marker = 0;
        (void)tracer.buildProgram(makeNativeLinkable({{"orenNayarInit", &marker}}, 1), "orenNayarInit");
    });
    expect("missing symbol", [&] { (void)tracer.buildProgram(hipBuiltinProgram(), "disneyInit"); });
    expect("foreign format", [&] {
        LinkableProgram ir;
        ir.format = "LLVM IR";
        (void)tracer.buildProgram(ir, "trace");
    });
    expect("corrupt native table", [&] {
        LinkableProgram p = makeNativeLinkable({{"trace", nullptr}}, 2);
        p.exchange.resize(p.exchange.size() - 3);
        (void)tracer.buildProgram(p, "trace");
    });
    expect("kernel of a foreign format", [&] {
        LinkableProgram ir;
        ir.format = "LLVM IR";
        (void)acc->compileKernel({ir}, {}, {}, "piperMain");
    });
    expect("undefined kernel entry", [&] { (void)acc->compileKernel({hipBuiltinProgram()}, {}, {}, "piperMain"); });
    expect("built-in as kernel entry", [&] { (void)acc->compileKernel({hipBuiltinProgram()}, {}, {}, "trace"); });
    expect("texture channels", [&] { (void)tracer.generateTexture(colourConfig({1, 1, 1}), 1); });
    expect("negative glass RoughnessY", [&] {
        auto glass = std::make_shared<Surface>(Surface::Kind::Glass,
                                               std::vector<Config>{colourConfig({1, 1, 1}), colourConfig({1, 1, 1}),
                                                                   scalarConfig(0.1f), scalarConfig(-0.2f)});
        auto buf = acc->createBuffer(48, 16, [](std::byte* p) {
            const float v[9] = {0, 0, 0, 1, 0, 0, 0, 1, 0};
            const uint32_t idx[3] = {0, 1, 2};
            std::memcpy(p, v, 36);
            std::memcpy(p + 36, idx, 12);
        });
        GeometryDesc gd;
        gd.desc = TriangleIndexedGeometryDesc{3, 1, buf, 48, 0, 36};
        (void)exportScene(tracer, tracer.buildNode({{TransformInfo{}, tracer.buildNode(tracer.buildGSMInstance(
                                                                          std::make_shared<Geometry>(gd), glass, nullptr))}}));
    });
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
        if (argc == 2 && std::string(argv[1]) == "--payloads") return payloads();
        if (argc == 2 && std::string(argv[1]) == "--kernel") return kernel();
        if (argc == 2 && std::string(argv[1]) == "--module") return module();
        const bool viaModule = argc >= 2 && std::string(argv[1]) == "--via-module";
        const int a0 = viaModule ? 2 : 1;
        if (argc - a0 >= 2 && argc - a0 <= 4 && std::string(argv[1]) != "--errors")
            return render(argv[a0], argv[a0 + 1], argc - a0 >= 3 ? argv[a0 + 2] : "box",
                          argc - a0 == 4 ? std::stof(argv[a0 + 3]) : 0.0f, viaModule);
        std::cerr << "usage: render_scene [--via-module] SCENE OUT [FILTER [ALPHA]] | render_scene --errors SCENE | "
                     "render_scene --payloads | --kernel | --module\n";
        return 2;
    } catch (const PiperException& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 3;
    }
}

// piper_render: the Render operator (PiperRenderer.cpp:231-357) over the HIP component pair.
//
//   piper_render OPTIONS.json [--export DIR]
//   piper_render --write-exr RGB.f32 WIDTH HEIGHT OUT.exr
//
// OPTIONS.json carries the operator's own keys: ImageSize [w, h], SceneDesc (path of the scene
// file, relative to OPTIONS.json), SceneParser and Tracer (accepted; this build parses JSON and
// renders with the HIP tracer), OutputDir. The scene file follows the reference's schema:
// RenderDriver, Actions, Objects, Scene, LightSampler, Integrator, Sampler, with components named
// by ClassID "<module>.<class>" as in Src/Module.json. Each action renders FrameCount frames from
// the node named ActiveSensor and writes OutputDir/Action<i>/Frame<j>.exr.
//
// --export DIR builds the same scene graph on an accelerator without a device and writes, per
// action, DIR/Action<i>.pgsc: the flattened scene the pipeline would upload, with the camera of
// the active sensor (the scene-file layout of tests/plugin/render_scene.cpp).
// --write-exr converts raw float32 RGB (row-major, top row first) into an EXR file.
//
// Mesh and image paths are relative to the scene file. Colour and roughness inputs take
// ConstantTexture, CheckBoard or a sampled image ({WrapMode, Image: {Path}}, read from PNG or binary
// PPM / PGM / PAM with 8-bit samples, RGB widened to RGBA as the reference's loader does). Components of
// the reference that the HIP path does not implement (keyframes nested in keyframes, other image
// formats) are rejected with a message.
#include <array>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../plugin/PiperHIP.hpp"
#include "Exr.hpp"
#include "Json.hpp"
#include "MeshIO.hpp"

using namespace Piper;
using PiperFrontend::Json;
namespace fs = std::filesystem;

namespace {

[[noreturn]] void fail(const std::string& msg) { throw PiperException(msg); }

std::string readFile(const fs::path& p) {
    std::ifstream f(p, std::ios::binary);
    if (!f) fail("cannot open " + p.string());
    std::ostringstream s;
    s << f.rdbuf();
    return s.str();
}

float num(const Json& j) { return static_cast<float>(j.number()); }
std::array<float, 3> vec3(const Json& j) {  // parseVector / parseSpectrum: exactly 3 numbers
    const auto& a = j.array();
    if (a.size() != 3) fail("expected 3 numbers" + j.where());
    return {num(a[0]), num(a[1]), num(a[2])};
}
std::array<float, 2> vec2(const Json& j) {
    const auto& a = j.array();
    if (a.size() != 2) fail("expected 2 numbers" + j.where());
    return {num(a[0]), num(a[1])};
}
uint32_t count(const Json& j) {
    const double v = j.number();
    if (!(v >= 0.0 && v <= 4294967295.0) || v != std::floor(v)) fail("expected an unsigned integer" + j.where());
    return static_cast<uint32_t>(v);
}

// ClassID "<module>.<class>": the module is everything before the last '.' (PiperCore.cpp:366-376)
std::pair<std::string, std::string> classOf(const Json& cfg) {
    const std::string& id = cfg.at("ClassID").string();
    const size_t dot = id.find_last_of('.');
    if (dot == std::string::npos) fail("invalid ClassID \"" + id + "\"" + cfg.where());
    return {id.substr(0, dot), id.substr(dot + 1)};
}
bool isClass(const Json& cfg, const char* module, const char* cls) {
    const auto [m, c] = classOf(cfg);
    return m == module && c == cls;
}
void requireClass(const Json& cfg, const char* module, const char* cls) {
    if (!isClass(cfg, module, cls))
        fail(std::string("expected ClassID ") + module + "." + cls + ", got " + cfg.at("ClassID").string() + cfg.where());
}

// parseTransform (PiperRenderer.cpp:58-137): Translation, Scale (number, [x, y, z] or
// {Scale, Skew, Shift}), Rotation (Euler [pitch, yaw, roll] in radians, or quaternion [w, x, y, z]).
TransformSRT parseTransform(const Json& cfg) {
    TransformSRT t = identitySRT();
    if (const Json* tr = cfg.find("Translation")) {
        const auto v = vec3(*tr);
        t.transX = v[0], t.transY = v[1], t.transZ = v[2];
    }
    if (const Json* sc = cfg.find("Scale")) {
        if (sc->isNumber()) {
            t.scaleX = t.scaleY = t.scaleZ = num(*sc);
        } else if (sc->isArray()) {
            const auto v = vec3(*sc);
            t.scaleX = v[0], t.scaleY = v[1], t.scaleZ = v[2];
        } else if (sc->isObject()) {
            const auto s = vec3(sc->at("Scale")), k = vec3(sc->at("Skew")), h = vec3(sc->at("Shift"));
            t.scaleX = s[0], t.scaleY = s[1], t.scaleZ = s[2];
            t.skewXY = k[0], t.skewXZ = k[1], t.skewYZ = k[2];
            t.shiftX = h[0], t.shiftY = h[1], t.shiftZ = h[2];
        } else {
            fail("invalid transform Scale" + sc->where());
        }
    }
    if (const Json* rot = cfg.find("Rotation")) {
        const auto& a = rot->array();
        if (a.size() == 3) {  // Euler angles: pitch about x, yaw about y, roll about z, half angles
            const float hp = 0.5f * num(a[0]), hy = 0.5f * num(a[1]), hr = 0.5f * num(a[2]);
            const float cp = std::cos(hp), sp = std::sin(hp), cy = std::cos(hy), sy = std::sin(hy);
            const float cr = std::cos(hr), sr = std::sin(hr);
            t.quatW = cr * cp * cy + sr * sp * sy;
            t.quatX = cr * sp * cy + sr * cp * sy;
            t.quatY = cr * cp * sy - sr * sp * cy;
            t.quatZ = sr * cp * cy - cr * sp * sy;
        } else if (a.size() == 4) {
            t.quatW = num(a[0]), t.quatX = num(a[1]), t.quatY = num(a[2]), t.quatZ = num(a[3]);
        } else {
            fail("invalid transform Rotation" + rot->where());
        }
    }
    return t;
}

class SceneBuilder {
public:
    SceneBuilder(HIPTracer& tracer, fs::path base) : mTracer(tracer), mBase(std::move(base)) {}
    std::map<std::string, std::shared_ptr<Node>> named;              // nodes with an ID
    std::map<const Node*, std::shared_ptr<Sensor>> sensors;          // sensor leaves

    // buildScene / buildSceneImpl (PiperRenderer.cpp:154-214)
    std::shared_ptr<Node> build(const Json& cfg) {
        std::shared_ptr<Node> node;
        const std::string& type = cfg.at("Type").string();
        if (type == "Instance") {
            if (cfg.find("ClassID")) {
                const auto obj = object(cfg);
                node = mTracer.buildNode(obj);
                if (auto s = std::dynamic_pointer_cast<Sensor>(obj)) sensors[node.get()] = s;
            } else
                node = mTracer.buildNode(mTracer.buildGSMInstance(geometry(cfg.at("Geometry")), surface(cfg.at("Surface")), nullptr));
        } else if (type == "Group") {
            std::vector<std::pair<TransformInfo, std::shared_ptr<Node>>> children;
            for (const Json& child : cfg.at("Children").array()) {
                TransformInfo info;
                if (const Json* tr = child.find("Transform")) {
                    if (const Json* offset = tr->find("Offset")) {  // keyframes (PiperRenderer.cpp:170-187)
                        info.offset = num(*offset);
                        info.step = num(tr->at("Step"));
                        for (const Json& key : tr->at("KeyFrames").array())
                            info.transforms.push_back({count(key.at("Key")), parseTransform(key.at("Transform"))});
                        if (info.transforms.size() < 2) fail("Invalid dynamic transform" + tr->where());
                        for (size_t k = 1; k < info.transforms.size(); ++k)
                            if (info.transforms[k].first <= info.transforms[k - 1].first)
                                fail("Invalid dynamic transform" + tr->where());
                    } else {
                        info.transforms.push_back({0u, parseTransform(*tr)});
                    }
                }
                children.push_back({info, build(child)});
            }
            node = mTracer.buildNode(children);
        } else {
            fail("invalid scene node type \"" + type + "\"" + cfg.where());
        }
        if (const Json* id = cfg.find("ID")) named[id->string()] = node;
        return node;
    }

private:
    HIPTracer& mTracer;
    fs::path mBase;

    std::shared_ptr<Geometry> meshGeometry(const std::vector<float>& verts, const std::vector<uint32_t>& index,
                                           const std::vector<float>& uv = {}) {
        if (index.empty()) fail("a mesh without triangles");
        const size_t vBytes = verts.size() * 4, iOff = (vBytes + 15) / 16 * 16;
        const size_t uvOff = (iOff + index.size() * 4 + 15) / 16 * 16;
        const size_t size = uv.empty() ? iOff + index.size() * 4 : uvOff + uv.size() * 4;
        auto buffer = mTracer.getAccelerator().createBuffer(size, 16, [&](std::byte* p) {
            std::memcpy(p, verts.data(), vBytes);
            std::memcpy(p + iOff, index.data(), index.size() * 4);
            if (!uv.empty()) std::memcpy(p + uvOff, uv.data(), uv.size() * 4);
        });
        GeometryDesc gd;
        gd.desc = TriangleIndexedGeometryDesc{static_cast<uint32_t>(verts.size() / 3), static_cast<uint32_t>(index.size() / 3),
                                              buffer, size, 0, iOff};
        if (!uv.empty()) gd.desc.texCoords = uvOff;
        return std::make_shared<Geometry>(gd);
    }
    struct PlanePrim {
        std::array<float, 3> origin, u, v;
    };
    static std::vector<PlanePrim> planes(const Json& cfg) {  // BasicShapes Plane (BasicShapes.cpp:43-70)
        std::vector<PlanePrim> out;
        for (const Json& p : cfg.at("Primitives").array()) out.push_back({vec3(p.at("Origin")), vec3(p.at("U")), vec3(p.at("V"))});
        if (out.empty()) fail("a Plane without primitives" + cfg.where());
        return out;
    }
    std::shared_ptr<Geometry> geometry(const Json& cfg) {
        const auto [m, c] = classOf(cfg);
        if (m == "Piper.BuiltinComponent.TriangleMesh" && c == "TriangleMesh") {
            fs::path p = cfg.at("Path").string();
            if (p.is_relative()) p = mBase / p;
            const auto mesh = PiperFrontend::loadMesh(p.string());
            return meshGeometry(mesh.verts, mesh.index, mesh.uv);
        }
        if (m == "Piper.BuiltinComponent.BasicShapes" && c == "Plane") {
            // a plane as a surface: two triangles whose corners carry the parallelogram coordinates
            // (0,0), (1,0), (1,1), (0,1), so the interpolated texture coordinate of a hit is its
            // (u, v) in the plane (calcTexCoord, BasicShapes/Kernel.cpp:30-56 and :109)
            std::vector<float> verts, uv;
            std::vector<uint32_t> index;
            for (const auto& p : planes(cfg)) {
                const uint32_t b = static_cast<uint32_t>(verts.size() / 3);
                for (int corner = 0; corner < 4; ++corner) {
                    const float a = (corner == 1 || corner == 2) ? 1.0f : 0.0f, d = corner >= 2 ? 1.0f : 0.0f;
                    for (int k = 0; k < 3; ++k) verts.push_back(p.origin[k] + a * p.u[k] + d * p.v[k]);
                    uv.push_back(a), uv.push_back(d);
                }
                index.insert(index.end(), {b, b + 1, b + 2, b, b + 2, b + 3});
            }
            return meshGeometry(verts, index, uv);
        }
        fail("unsupported geometry " + cfg.at("ClassID").string() + cfg.where());
    }
    // Surface/Classic (Classic.cpp:207-237): each input is kept as its texture config; the tracer
    // loads it when the pipeline materialises the surface (generateTexture checks the channels).
    std::shared_ptr<Surface> surface(const Json& cfg) const {
        const auto [m, c] = classOf(cfg);
        if (m != "Piper.BuiltinComponent.ClassicSurface") fail("unsupported surface " + cfg.at("ClassID").string() + cfg.where());
        auto in = [&](std::initializer_list<const char*> keys) {
            std::vector<Json> v;
            for (const char* k : keys) v.push_back(cfg.at(k));
            return v;
        };
        using K = Surface::Kind;
        if (c == "Matte") return std::make_shared<Surface>(K::Matte, in({"Diffuse", "Roughness"}));
        if (c == "Plastic") return std::make_shared<Surface>(K::Plastic, in({"Diffuse", "Specular", "Roughness"}));
        if (c == "Mirror") return std::make_shared<Surface>(K::Mirror, in({"Reflection"}));
        if (c == "Glass")
            return std::make_shared<Surface>(K::Glass, in({"Reflection", "Transmission", "RoughnessX", "RoughnessY"}));
        if (c == "Substrate")
            return std::make_shared<Surface>(K::Substrate, in({"Diffuse", "Specular", "RoughnessX", "RoughnessY"}));
        if (c == "BlackBody") return std::make_shared<Surface>(K::BlackBody, std::vector<Json>{});
        // the GGX conductor this path adds (not a reference class): F0, Roughness (alpha = roughness^2)
        if (c == "Conductor") return std::make_shared<Surface>(K::Conductor, in({"F0", "Roughness"}));
        fail("unsupported surface class " + c + cfg.where());
    }
    static std::shared_ptr<Object> object(const Json& cfg) {  // an Instance with a ClassID: a light or a sensor
        const auto [m, c] = classOf(cfg);
        if (m == "Piper.BuiltinComponent.AreaLight" && c == "Diffuse") {  // AreaLight.cpp:17-20
            const Json& g = cfg.at("Geometry");
            requireClass(g, "Piper.BuiltinComponent.BasicShapes", "Plane");
            std::vector<Light::PlanePrimitive> prims;
            for (const auto& p : planes(g)) prims.push_back({p.origin, p.u, p.v});
            return Light::area(std::move(prims), vec3(cfg.at("Radiance")));
        }
        if (m == "Piper.BuiltinComponent.BasicLight" && c == "Point") return Light::point(vec3(cfg.at("Intensity")));
        if (m == "Piper.BuiltinComponent.BasicLight" && c == "SampledEnvironment")
            return Light::environment(vec3(cfg.at("Radiance")));
        if (m == "Piper.BuiltinComponent.PerspectiveCamera" && c == "Sensor") {  // PerspectiveCamera.cpp:35-46
            auto s = std::make_shared<Sensor>();
            s->lookAt = vec3(cfg.at("LookAt"));
            const auto size = vec2(cfg.at("SensorSize"));  // millimetres
            s->size = {size[0] * 1e-3f, size[1] * 1e-3f};
            s->upRef = vec3(cfg.at("Up"));
            s->focalLength = num(cfg.at("FocalLength")) * 1e-3f;
            s->apertureRadius = s->focalLength / (2.0f * num(cfg.at("FStop")));
            return s;
        }
        fail("unsupported scene object " + cfg.at("ClassID").string() + cfg.where());
    }
};

struct Action {  // ActionInfo (PiperRenderer.cpp:228-249)
    std::string activeSensor;
    float timeBegin, shutterOpen, shutterClose;
    uint32_t frameCount;
    FitMode fitMode;
};
std::vector<Action> parseActions(const Json& cfg) {
    std::vector<Action> out;
    for (const Json& a : cfg.array()) {
        const std::string& fit = a.at("FitMode").string();
        if (fit != "Fill" && fit != "OverScan") fail("Invalid FitMode \"" + fit + "\"" + a.where());
        out.push_back({a.at("ActiveSensor").string(), num(a.at("TimeBegin")), num(a.at("ShutterOpenTime")),
                       num(a.at("ShutterCloseTime")), count(a.at("FrameCount")),
                       fit == "Fill" ? FitMode::Fill : FitMode::OverScan});
    }
    return out;
}

struct Globals {
    std::shared_ptr<FixedSampler> driver;
    Integrator integrator;
    LightSampler lightSampler;
    std::shared_ptr<Sampler> sampler;
};
Globals parseGlobals(const Json& scene) {
    Globals g;
    const Json& rd = scene.at("RenderDriver");
    requireClass(rd, "Piper.BuiltinComponent.FixedSampler", "Driver");
    g.driver = std::make_shared<FixedSampler>();
    const Json& flt = rd.at("Filter");  // SimpleFilter (FixedSampler.cpp:28-45)
    requireClass(flt, "Piper.BuiltinComponent.FixedSampler", "SimpleFilter");
    g.driver->filter.filterType = flt.at("FilterType").string();
    if (g.driver->filter.filterType == "gaussian") g.driver->filter.alpha = num(flt.at("Alpha"));
    const Json& integ = scene.at("Integrator");
    requireClass(integ, "Piper.BuiltinComponent.PathTracing", "Integrator");
    g.integrator.maxDepth = count(integ.at("MaxTraceDepth"));
    g.integrator.minDepthForRussianRoulette = count(integ.at("MinRussianRouletteDepth"));
    requireClass(scene.at("LightSampler"), "Piper.BuiltinComponent.BasicLight", "UniformLightSampler");
    const Json& smp = scene.at("Sampler");
    requireClass(smp, "Piper.BuiltinComponent.ScrambledSobolSampler", "Sampler");
    g.sampler = std::make_shared<Sampler>();
    g.sampler->samplesPerPixel = count(smp.at("SamplesPerPixel"));
    if (const Json* s = smp.find("Scramble")) (void)s->boolean();  // the reference's kernels never scramble
    return g;
}

template <typename T>
void put(std::ofstream& f, const T* p, size_t n) {
    f.write(reinterpret_cast<const char*>(p), static_cast<std::streamsize>(n * sizeof(T)));
}

// Version 1: the layout below; version 2 (scenes with moving instances) always writes the point
// light count and a texture flag, then the moving instances with their keys for the action's
// first frame.
void exportAction(const fs::path& out, const SceneExport& e, const Node* sensorNode, const Globals& g, uint32_t width,
                  uint32_t height, FitMode fit, float frameBegin, float frameEnd) {
    const pg_camera* cam = nullptr;
    for (size_t i = 0; i < e.sensors.size(); ++i)
        if (e.sensors[i].first == sensorNode) cam = &e.cameras[i];
    if (!cam) fail("the active sensor is not in the scene");
    const uint32_t nv = static_cast<uint32_t>(e.verts.size() / 3), nt = static_cast<uint32_t>(e.index.size() / 3);
    const uint32_t no = static_cast<uint32_t>(e.objectMaterial.size()), nm = static_cast<uint32_t>(e.materials.size());
    const uint32_t nq = static_cast<uint32_t>(e.quads.size());
    const uint32_t head[12] = {nv, nt, no, nm, nq, width, height, g.sampler->samplesPerPixel, g.integrator.maxDepth,
                               g.integrator.minDepthForRussianRoulette, 0u, fit == FitMode::Fill ? 0u : 1u};
    std::ofstream f(out, std::ios::binary);
    if (!f) fail("cannot write " + out.string());
    f.write("PGSC", 4);
    const bool moving = !e.instanceFirst.empty();
    bool multiPrimitive = false;  // some area light has several primitives: their counts end the file
    for (uint32_t c : e.lightQuads) multiPrimitive = multiPrimitive || c != 1;
    const bool anisotropic = e.anisotropic();  // RoughnessY per material and its texture end the file
    const uint32_t version = (moving ? 2u : 1u) | (multiPrimitive ? 4u : 0u) | (anisotropic ? 8u : 0u);
    put(f, &version, 1);
    put(f, head, 12);
    put(f, e.verts.data(), e.verts.size());
    put(f, e.index.data(), e.index.size());
    put(f, e.triObject.data(), e.triObject.size());
    put(f, e.objectMaterial.data(), e.objectMaterial.size());
    for (const pg_material& m : e.materials) {
        const float row[8] = {static_cast<float>(m.type), m.albedo[0], m.albedo[1], m.albedo[2], m.alpha,
                              m.reserved[0], m.reserved[1], m.reserved[2]};
        put(f, row, 8);
    }
    for (const pg_quad& q : e.quads) {
        put(f, q.origin, 3), put(f, q.u, 3), put(f, q.v, 3), put(f, q.radiance, 3);
    }
    const float camera[14] = {cam->position[0], cam->position[1], cam->position[2], cam->look_at[0], cam->look_at[1],
                              cam->look_at[2], cam->up[0], cam->up[1], cam->up[2], cam->sensor_size[0],
                              cam->sensor_size[1], cam->focal_length, cam->aperture_radius, 0.0f};
    put(f, camera, 14);
    put(f, e.env.data(), 3);
    const TransformSRT id = identitySRT();
    for (uint32_t o = 0; o < no; ++o) put(f, reinterpret_cast<const float*>(&id), 16);
    const uint32_t np = static_cast<uint32_t>(e.points.size());
    // with moving instances or several-primitive lights the tail is explicit: point lights, a
    // texture flag, the textures, the instances, the light counts
    const bool explicitTail = moving || multiPrimitive || anisotropic;
    if (np || !e.textures.empty() || explicitTail) {  // optional tail: point lights, then textures
        put(f, &np, 1);
        for (const pg_point_light& p : e.points) put(f, p.position, 3), put(f, p.intensity, 3);
    }
    if (explicitTail) {
        const uint32_t hasTextures = e.textures.empty() ? 0u : 1u;
        put(f, &hasTextures, 1);
    }
    if (!e.textures.empty()) {
        // texcoords (float2 per vertex, zeros when no mesh had any), the textures (type, channels,
        // black, white, scale, width, height, wrap, then the texels), the bindings per material
        std::vector<float> uv = e.texcoords;
        uv.resize(2 * static_cast<size_t>(nv), 0.0f);
        put(f, uv.data(), uv.size());
        const uint32_t nx = static_cast<uint32_t>(e.textures.size());
        put(f, &nx, 1);
        for (size_t i = 0; i < e.textures.size(); ++i) {
            const pg_texture& d = e.textures[i];
            const uint32_t head4[2] = {d.type, d.channels};
            put(f, head4, 2), put(f, d.black, 4), put(f, d.white, 4), put(f, &d.scale, 1);
            const uint32_t dims[3] = {d.width, d.height, d.wrap};
            put(f, dims, 3);
            put(f, e.textureTexels[i].data(), e.textureTexels[i].size());
        }
        for (const pg_material_textures& b : e.materialTextures) put(f, &b.albedo, 1), put(f, &b.reserved, 1), put(f, &b.alpha, 1);
    }
    if (moving) {  // moving instances: first triangles, then per instance its frame keys
        const uint32_t ni = static_cast<uint32_t>(e.instanceFirst.size());
        put(f, &ni, 1);
        put(f, e.instanceFirst.data(), ni);
        for (const auto& m : e.motions) {
            const pg_motion pm = motionForInterval(*m.keys, m.parent, frameBegin, frameEnd);
            const uint32_t head2[2] = {pm.num_keys, pm.visible};
            put(f, head2, 2), put(f, &pm.t0, 1), put(f, &pm.t1, 1);
            for (uint32_t k = 0; k < pm.num_keys; ++k) put(f, pm.keys[k], 16);
            put(f, pm.parent, 12);
        }
    }
    if (multiPrimitive) {  // primitives (consecutive quads) per area light
        const uint32_t nl = static_cast<uint32_t>(e.lightQuads.size());
        put(f, &nl, 1);
        put(f, e.lightQuads.data(), nl);
    }
    if (anisotropic) {  // RoughnessY per material, then its texture (PG_TEX_NONE: the constant)
        put(f, e.roughnessY.data(), nm);
        put(f, e.roughnessYTextures.data(), nm);
    }
    if (!f) fail("cannot write " + out.string());
}

int render(const fs::path& optionsPath, const std::string& exportDir) {
    const Json opt = Json::parse(readFile(optionsPath));
    const auto size = opt.at("ImageSize").array();
    if (size.size() != 2) fail("Invalid ImageSize.");
    const uint32_t width = count(size[0]), height = count(size[1]);
    fs::path scenePath = opt.at("SceneDesc").string();
    if (scenePath.is_relative()) scenePath = optionsPath.parent_path() / scenePath;
    const Json scene = Json::parse(readFile(scenePath));
    const Globals g = parseGlobals(scene);
    const auto actions = parseActions(scene.at("Actions"));
    for (const Json& obj : scene.at("Objects").array()) (void)obj.at("ID").string();  // named objects (unused)

    const bool exporting = !exportDir.empty();
    auto accelerator = exporting ? HIPAccelerator::hostOnly() : std::make_shared<HIPAccelerator>(0);
    HIPTracer tracer(accelerator);
    tracer.setResourceRoot(scenePath.parent_path().string());  // image paths are relative to the scene file
    SceneBuilder builder(tracer, scenePath.parent_path());
    const auto root = builder.build(scene.at("Scene"));
    auto sensorOf = [&](const Action& a) {
        const auto it = builder.named.find(a.activeSensor);
        if (it == builder.named.end())
            fail("Failed to find a node with sensor named \"" + a.activeSensor + "\".");
        return it->second;
    };
    if (exporting) {
        fs::create_directories(exportDir);
        const SceneExport e = exportScene(tracer, root);
        for (size_t i = 0; i < actions.size(); ++i) {
            const auto node = sensorOf(actions[i]);
            if (!builder.sensors.count(node.get()))
                fail("ActiveSensor \"" + actions[i].activeSensor + "\" is not a sensor");
            exportAction(fs::path(exportDir) / ("Action" + std::to_string(i) + ".pgsc"), e, node.get(), g, width, height,
                         actions[i].fitMode, actions[i].timeBegin, actions[i].timeBegin + actions[i].shutterOpen);
        }
        std::printf("exported %zu action(s) to %s\n", actions.size(), exportDir.c_str());
        return 0;
    }
    auto pipeline = tracer.buildPipeline(root, const_cast<Integrator&>(g.integrator), *g.driver,
                                         const_cast<LightSampler&>(g.lightSampler), g.sampler);
    const fs::path output = opt.at("OutputDir").string();
    for (size_t i = 0; i < actions.size(); ++i) {
        const Action& a = actions[i];
        const auto launcher = pipeline->prepare(sensorOf(a), width, height, a.fitMode);
        const fs::path dir = output / ("Action" + std::to_string(i));
        fs::create_directories(dir);
        for (uint32_t k = 0; k < a.frameCount; ++k) {
            const float begin = a.timeBegin + static_cast<float>(k) * (a.shutterOpen + a.shutterClose);
            launcher->updateTimeInterval(begin, begin + a.shutterOpen);
            const auto frame = g.driver->renderFrame(tracer, *launcher);
            const fs::path file = dir / ("Frame" + std::to_string(k) + ".exr");
            const size_t bad = PiperFrontend::writeExrRgb(file.string(), width, height, frame);
            if (bad) std::fprintf(stderr, "warning: %zu invalid pixels in %s\n", bad, file.string().c_str());
            std::printf("wrote %s\n", file.string().c_str());
        }
    }
    std::printf("%s\n", pipeline->generateStatisticsReport().c_str());
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        if (argc == 6 && std::string(argv[1]) == "--write-exr") {
            const uint32_t w = static_cast<uint32_t>(std::stoul(argv[3])), h = static_cast<uint32_t>(std::stoul(argv[4]));
            const std::string raw = readFile(argv[2]);
            if (raw.size() != static_cast<size_t>(w) * h * 12) fail("raw RGB size does not match WIDTH x HEIGHT");
            std::vector<std::array<float, 3>> rgb(static_cast<size_t>(w) * h);
            std::memcpy(rgb.data(), raw.data(), raw.size());
            const size_t bad = PiperFrontend::writeExrRgb(argv[5], w, h, rgb);
            std::printf("%zu invalid pixels\n", bad);
            return 0;
        }
        if (argc == 2) return render(argv[1], "");
        if (argc == 4 && std::string(argv[2]) == "--export") return render(argv[1], argv[3]);
        std::fprintf(stderr, "usage: piper_render OPTIONS.json [--export DIR]\n"
                             "       piper_render --write-exr RGB.f32 WIDTH HEIGHT OUT.exr\n");
        return 2;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}

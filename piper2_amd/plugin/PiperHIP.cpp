// HIPAccelerator / HIPTracer / the built-in components / FixedSampler / the module entry points,
// over the C ABI (declarations and the reference interfaces they implement: PiperHIP.hpp).
#include "PiperHIP.hpp"

#include <algorithm>
#include <cmath>
#include <deque>
#include <filesystem>
#include <new>
#include <sstream>

#include "../frontend/Image.hpp"

namespace Piper {
namespace {

[[noreturn]] void raise(const std::string& what) { throw PiperException(what); }

void check(int code, const char* what) {
    if (code != PG_OK) raise(std::string(what) + ": " + pg_last_error());
}

using Mat34 = std::array<float, 12>;  // row-major 3x4, local -> world

Mat34 identity34() { return {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0}; }

// a * b (apply b, then a)
Mat34 compose(const Mat34& a, const Mat34& b) {
    Mat34 r{};
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 4; ++j) {
            float s = a[4 * i] * b[j] + a[4 * i + 1] * b[4 + j] + a[4 * i + 2] * b[8 + j];
            if (j == 3) s += a[4 * i + 3];
            r[4 * i + j] = s;
        }
    }
    return r;
}

// translate * rotate(unit quaternion) * (scale / skew / shift), the decomposition TransformSRT
// stores (Kernel/Transform.hpp:390-393, Embree's quaternion decomposition).
Mat34 fromSRT(const TransformSRT& t) {
    const Mat34 s = {t.scaleX, t.skewXY, t.skewXZ, t.shiftX, 0.0f, t.scaleY, t.skewYZ, t.shiftY, 0.0f, 0.0f, t.scaleZ, t.shiftZ};
    const float n = std::sqrt(t.quatW * t.quatW + t.quatX * t.quatX + t.quatY * t.quatY + t.quatZ * t.quatZ);
    if (!(n > 0.0f)) raise("TransformSRT: zero quaternion");
    const float w = t.quatW / n, x = t.quatX / n, y = t.quatY / n, z = t.quatZ / n;
    const Mat34 r = {1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y), t.transX,
                     2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x), t.transY,
                     2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y), t.transZ};
    return compose(r, s);
}

bool isIdentity(const TransformSRT& t) {
    const TransformSRT i = identitySRT();
    return std::memcmp(&t, &i, sizeof(t)) == 0;
}

std::array<float, 3> applyPoint(const Mat34& m, const float* p) {
    std::array<float, 3> r{};
    for (int i = 0; i < 3; ++i) r[i] = m[4 * i] * p[0] + m[4 * i + 1] * p[1] + m[4 * i + 2] * p[2] + m[4 * i + 3];
    return r;
}
std::array<float, 3> applyVector(const Mat34& m, const float* v) {
    std::array<float, 3> r{};
    for (int i = 0; i < 3; ++i) r[i] = m[4 * i] * v[0] + m[4 * i + 1] * v[1] + m[4 * i + 2] * v[2];
    return r;
}

template <typename T>
T unpack(const SBTPayload& p, const char* what) {
    if (p.size() != sizeof(T)) raise(std::string(what) + ": payload of " + std::to_string(p.size()) + " bytes, expected " +
                                     std::to_string(sizeof(T)));
    T v;
    std::memcpy(&v, p.data(), sizeof(T));
    return v;
}

// ---- resources ----
class HostBuffer final : public Resource {
public:
    std::vector<std::byte> bytes;
};

class HIPTiledOutput final : public TiledOutput {
public:
    HIPTiledOutput(pg_device* dev, size_t size) : mDevice(dev), mSize(size) {}
    ~HIPTiledOutput() override {
        if (mFilm) pg_film_destroy(mFilm);
    }
    // The film takes its resolution from the first launch that writes it (size = w * h * RGBW).
    pg_film* bind(uint32_t w, uint32_t h) {
        if (static_cast<size_t>(w) * h * 16 != mSize) raise("TiledOutput size does not match the film resolution");
        if (!mFilm) {
            check(pg_film_create(mDevice, w, h, &mFilm), "pg_film_create");
            mW = w, mH = h;
        } else if (w != mW || h != mH) {
            raise("TiledOutput bound to another film resolution");
        }
        return mFilm;
    }
    [[nodiscard]] Future<std::vector<std::byte>> download() const override {
        FutureVoid dirty = mDirty;
        pg_film* film = mFilm;
        const size_t size = mSize;
        return Future<std::vector<std::byte>>([dirty, film, size] {
            if (dirty.valid()) dirty.getSync();
            std::vector<std::byte> out(size);
            if (film) check(pg_film_download(film, reinterpret_cast<float*>(out.data()), nullptr), "pg_film_download");
            return out;
        });
    }
    void markDirty(FutureVoid future) override { mDirty = std::move(future); }

private:
    pg_device* mDevice;
    size_t mSize;
    pg_film* mFilm = nullptr;
    uint32_t mW = 0, mH = 0;
    FutureVoid mDirty;
};

// ---- accelerator internals ----
class HIPContext final : public Context {
public:
    explicit HIPContext(pg_device* dev) : mDevice(dev) {}
    [[nodiscard]] std::unique_lock<std::recursive_mutex> makeCurrent() override {
        return std::unique_lock<std::recursive_mutex>(mMutex);
    }
    // the library's device handle; its calls run on the one stream it owns
    [[nodiscard]] ContextHandle getHandle() const noexcept override { return reinterpret_cast<ContextHandle>(mDevice); }
    [[nodiscard]] CommandQueueHandle select() override { return reinterpret_cast<CommandQueueHandle>(mDevice); }

private:
    pg_device* mDevice;
    std::recursive_mutex mMutex;
};

class NativeKernel final : public Kernel {
public:
    NativeKernelFunc entry = nullptr;
    std::vector<const void*> dynamicSymbols;
};

// ---- the built-in programs ----
struct HIPProgram final : RTProgram {
    std::string symbol;
};

// One marker per built-in symbol: its address is the symbol's entry in the native table (the
// device implements these inside its wavefront kernels; the host decodes their payloads).
struct BuiltinMarker {
    const char* name;
};
const std::vector<BuiltinMarker>& builtinMarkers() {
    static const std::vector<BuiltinMarker> markers = [] {
        std::vector<BuiltinMarker> m;
        for (const char* s :
             {"rayGen", "sobolStart", "sobolGenerate", "trace", "accumulate", "box", "triangle", "gaussian",
              "matteInit", "matteSample", "matteEvaluate", "mattePdf", "conductorInit", "conductorSample",
              "conductorEvaluate", "conductorPdf", "plasticInit", "plasticSample", "plasticEvaluate", "plasticPdf",
              "mirrorInit", "mirrorSample", "blackBodyInit", "blackBodySample", "blackBodyEvaluate", "blackBodyPdf",
              "glassInit", "glassSample", "glassEvaluate", "glassPdf", "substrateInit", "substrateSample",
              "substrateEvaluate", "substratePdf", "diffuseInit", "diffuseSample", "diffuseEvaluate", "diffusePdf",
              "planeSample", "uniformSelect", "pointInit", "pointSample", "deltaEvaluate", "deltaPdf", "SEInit",
              "SESample", "SEEvaluate", "SEPdf", "constantTexture", "checkBoard", "sampleTexture"})
            m.push_back({s});
        return m;
    }();
    return markers;
}

class HIPTraceLauncher;
void hipRenderKernel(const NativeLaunch& launch);

// ---- textures (Texture/ProceduralTexture, Texture/SampledTexture) ----
class ConstantTexture final : public Texture {
public:
    explicit ConstantTexture(const ConstantData& d) : mData(d) {}
    [[nodiscard]] uint32_t channel() const noexcept override { return mData.channel; }
    [[nodiscard]] TextureProgram materialize(const MaterializeContext& ctx) const override {
        return {packSBTPayload(mData), ctx.tracer.buildProgram(hipBuiltinProgram(), "constantTexture")};
    }

private:
    ConstantData mData;
};
class CheckBoardTexture final : public Texture {
public:
    explicit CheckBoardTexture(const CheckBoardData& d) : mData(d) {}
    [[nodiscard]] uint32_t channel() const noexcept override { return mData.channel; }
    [[nodiscard]] TextureProgram materialize(const MaterializeContext& ctx) const override {
        return {packSBTPayload(mData), ctx.tracer.buildProgram(hipBuiltinProgram(), "checkBoard")};
    }

private:
    CheckBoardData mData;
};
class ImageTexture final : public Texture {  // SampledTexture.cpp:34-80
public:
    ImageTexture(PiperFrontend::DecodedImage img, TextureWrap wrap) : mImage(std::move(img)), mWrap(wrap) {}
    [[nodiscard]] uint32_t channel() const noexcept override { return mImage.channels; }
    [[nodiscard]] TextureProgram materialize(const MaterializeContext& ctx) const override {
        auto texel = ctx.accelerator.createBuffer(mImage.texels.size(), 16, [this](std::byte* p) {
            std::memcpy(p, mImage.texels.data(), mImage.texels.size());
        });
        SampledTextureData d{mWrap, mImage.width, mImage.height, mImage.width * mImage.channels, mImage.channels,
                             ctx.registerResource(texel), nullptr};
        return {packSBTPayload(d), ctx.tracer.buildProgram(hipBuiltinProgram(), "sampleTexture")};
    }

private:
    PiperFrontend::DecodedImage mImage;
    TextureWrap mWrap;
};

// ---- scene graph ----
struct MeshAccel final : AccelerationStructure {
    std::vector<float> verts;     // float3 per vertex, with GeometryDesc::transform applied
    std::vector<uint32_t> index;  // uint3 per triangle
    std::vector<float> uv;        // float2 per vertex, or empty
};
struct HIPGSMInstance final : GSMInstance {
    std::shared_ptr<MeshAccel> mesh;
    std::shared_ptr<Surface> surface;
};
struct LeafNode final : Node {
    std::shared_ptr<Object> object;  // HIPGSMInstance, Light or Sensor
};
struct GroupNode final : Node {
    struct Child {
        Mat34 m;                                    // static transform (identity for a keyframed child)
        std::shared_ptr<const TransformInfo> keys;  // keyframed transform (EmbreeBranchNode), or null
        std::shared_ptr<Node> node;
    };
    std::vector<Child> children;
};

// The scene graph collected for materialisation (Embree.cpp:1096-1106 collect()): instances,
// lights and sensors with their local-to-world transforms; a TraversalHandle points at one of
// `transforms`.
struct Collected {
    struct Inst {
        const HIPGSMInstance* inst;
        Mat34 toWorld;  // into world space, or into its moving instance's space
        int mover;      // index into movers, -1 for static geometry
    };
    struct Mover {  // a keyframed Group child: its subtree moves as one instance
        std::shared_ptr<const TransformInfo> keys;
        Mat34 parent;  // static transform of the Group it sits in
    };
    std::vector<Mover> movers;
    struct Lit {
        std::shared_ptr<Light> light;
        TraversalHandle traversal;
    };
    struct Cam {
        const Node* node;
        const Sensor* sensor;
        TraversalHandle traversal;
    };
    std::vector<Inst> instances;
    std::vector<Lit> lights;
    std::vector<Cam> sensors;
    std::deque<Mat34> transforms;
    TraversalHandle handle(const Mat34& m) {
        transforms.push_back(m);
        return reinterpret_cast<TraversalHandle>(&transforms.back());
    }
    static const Mat34& transformOf(TraversalHandle h) {
        if (!h) raise("payload without a traversal handle");
        return *reinterpret_cast<const Mat34*>(h);
    }
};

// A keyframed child starts a moving instance (the reference's instance node with a time-sampled
// transform, Embree.cpp:777-835): its subtree is flattened into the child's own space under the
// static transform of the Group it sits in.
void collect(const std::shared_ptr<Node>& node, const Mat34& toWorld, Collected& c, int mover = -1) {
    if (auto* g = dynamic_cast<const GroupNode*>(node.get())) {
        for (const auto& ch : g->children) {
            if (ch.keys) {
                if (mover >= 0) raise("buildPipeline: keyframed transforms nested in a keyframed transform are not supported");
                c.movers.push_back({ch.keys, toWorld});
                collect(ch.node, identity34(), c, static_cast<int>(c.movers.size() - 1));
            } else {
                collect(ch.node, compose(toWorld, ch.m), c, mover);
            }
        }
        return;
    }
    auto* leaf = dynamic_cast<const LeafNode*>(node.get());
    if (!leaf) raise("buildPipeline: node was not built by this tracer");
    if (auto* inst = dynamic_cast<const HIPGSMInstance*>(leaf->object.get())) {
        c.instances.push_back({inst, toWorld, mover});
    } else if (mover >= 0) {
        raise("buildPipeline: lights and sensors under a keyframed transform are not supported");
    } else if (auto light = std::dynamic_pointer_cast<Light>(leaf->object)) {
        c.lights.push_back({light, c.handle(toWorld)});
    } else if (auto* sensor = dynamic_cast<const Sensor*>(leaf->object.get())) {
        c.sensors.push_back({node.get(), sensor, c.handle(toWorld)});
    } else {
        raise("buildPipeline: unsupported scene object");
    }
}

// What the components registered while materialising (Embree.cpp:1068-1094): calls as
// (symbol, payload) with handle = index + 1, and resources by index.
struct Registry {
    struct CallRec {
        std::string symbol;
        SBTPayload payload;
    };
    std::vector<CallRec> calls;
    std::vector<std::shared_ptr<Resource>> resources;

    MaterializeContext context(Tracer& tracer) {
        return MaterializeContext{
            tracer, tracer.getAccelerator(),
            [this](std::shared_ptr<Resource> r) {
                resources.push_back(std::move(r));
                return static_cast<uint32_t>(resources.size() - 1);
            },
            [this](const std::shared_ptr<RTProgram>& program, const SBTPayload& payload) {
                return add(program, payload);
            },
            [this, &tracer](const Config& desc, uint32_t channel) {
                const auto texture = tracer.generateTexture(desc, channel);
                const auto prog = texture->materialize(context(tracer));
                return add(prog.sample, prog.payload);
            }};
    }
    CallHandle add(const std::shared_ptr<RTProgram>& program, const SBTPayload& payload) {
        auto* p = dynamic_cast<const HIPProgram*>(program.get());
        if (!p) raise("registerCall: the program was not built by HIPTracer");
        calls.push_back({p->symbol, payload});
        return reinterpret_cast<CallHandle>(static_cast<uintptr_t>(calls.size()));
    }
    const CallRec& call(CallHandle h, const char* what) const {
        const uintptr_t i = reinterpret_cast<uintptr_t>(h);
        if (i == 0 || i > calls.size()) raise(std::string(what) + ": call handle not registered by this pipeline");
        return calls[i - 1];
    }
    const std::vector<std::byte>& buffer(uint32_t index, const char* what) const {
        if (index >= resources.size()) raise(std::string(what) + ": resource index out of range");
        auto* b = dynamic_cast<const HostBuffer*>(resources[index].get());
        if (!b) raise(std::string(what) + ": resource is not a buffer of this accelerator");
        return b->bytes;
    }
};

std::string symbolOf(const std::shared_ptr<RTProgram>& p) {
    auto* h = dynamic_cast<const HIPProgram*>(p.get());
    if (!h) raise("program was not built by HIPTracer");
    return h->symbol;
}
void expectSymbol(const std::shared_ptr<RTProgram>& p, const char* want, const char* what) {
    if (symbolOf(p) != want) raise(std::string(what) + ": expected program " + want + ", got " + symbolOf(p));
}

// A texture input decoded from its call: a constant, or a texture of the export.
struct Input {
    bool textured = false;
    std::array<float, 4> value{};
    pg_texture tex{};
    std::vector<uint8_t> texels;
};
Input decodeTexture(const Registry& reg, Call<TextureSampleFunc> call, uint32_t channels, const char* what) {
    const auto& c = reg.call(call.handle, what);
    Input in;
    if (c.symbol == "constantTexture") {
        const auto d = unpack<ConstantData>(c.payload, what);
        if (d.channel != channels) raise(std::string(what) + ": constant of " + std::to_string(d.channel) + " channels");
        std::copy(d.value, d.value + 4, in.value.begin());
        return in;
    }
    in.textured = true;
    if (c.symbol == "checkBoard") {
        const auto d = unpack<CheckBoardData>(c.payload, what);
        in.tex.type = PG_TEX_CHECKERBOARD, in.tex.channels = d.channel, in.tex.scale = d.scale;
        std::memcpy(in.tex.black, d.black, 16), std::memcpy(in.tex.white, d.white, 16);
    } else if (c.symbol == "sampleTexture") {
        const auto d = unpack<SampledTextureData>(c.payload, what);
        const auto& bytes = reg.buffer(d.texel, what);
        const size_t n = static_cast<size_t>(d.width) * d.height * d.channel;
        if (d.stride != d.width * d.channel || bytes.size() < n) raise(std::string(what) + ": texel buffer too small");
        in.tex.type = PG_TEX_IMAGE, in.tex.channels = d.channel, in.tex.width = d.width, in.tex.height = d.height;
        in.tex.wrap = d.wrap == TextureWrap::Mirror ? PG_WRAP_MIRROR : PG_WRAP_REPEAT;
        in.texels.resize(n);
        std::memcpy(in.texels.data(), bytes.data(), n);
    } else {
        raise(std::string(what) + ": " + c.symbol + " is not a texture program");
    }
    if (in.tex.channels != channels) raise(std::string(what) + ": texture of " + std::to_string(in.tex.channels) + " channels");
    return in;
}

// Builds the flat scene the C ABI takes from the decoded payloads of every component.
class SceneDecoder {
public:
    SceneDecoder(Tracer& tracer, Registry& reg) : mTracer(tracer), mReg(reg) {}

    uint32_t material(const Surface& s) {
        for (size_t i = 0; i < mSurfaces.size(); ++i)
            if (mSurfaces[i] == &s) return static_cast<uint32_t>(i);
        const SurfaceProgram prog = s.materialize(mReg.context(mTracer));
        pg_material m{};
        Input a, b, r, ry;  // colour, second colour, roughness (glass, substrate: RoughnessX), RoughnessY
        const char* what = "surface";
        switch (s.kind) {
            case Surface::Kind::Matte: {
                expectSymbol(prog.init, "matteInit", what);
                const auto d = unpack<MatteData>(prog.payload, "MatteData");
                a = decodeTexture(mReg, d.diffuseTexture, 4, "Matte Diffuse");
                r = decodeTexture(mReg, d.roughnessTexture, 1, "Matte Roughness");
                if (!r.textured && !(r.value[0] >= 0.0f && r.value[0] <= 1.0f)) raise("matte roughness must be in [0, 1]");
                m.type = PG_MAT_LAMBERT;
                break;
            }
            case Surface::Kind::Conductor: {
                expectSymbol(prog.init, "conductorInit", what);
                const auto d = unpack<ConductorData>(prog.payload, "ConductorData");
                a = decodeTexture(mReg, d.f0Texture, 4, "Conductor F0");
                r = decodeTexture(mReg, d.roughnessTexture, 1, "Conductor Roughness");
                if (d.roughnessIsAlpha && r.textured) raise("conductor: a roughness texture must hold roughness, not alpha");
                if (!r.textured) {
                    if (!(r.value[0] > 0.0f)) raise("conductor roughness must be positive");
                    if (!d.roughnessIsAlpha) r.value[0] = r.value[0] * r.value[0];
                }
                m.type = PG_MAT_CONDUCTOR;
                break;
            }
            case Surface::Kind::Plastic: {
                expectSymbol(prog.init, "plasticInit", what);
                const auto d = unpack<PlasticData>(prog.payload, "PlasticData");
                a = decodeTexture(mReg, d.diffuseTexture, 4, "Plastic Diffuse");
                b = decodeTexture(mReg, d.specularTexture, 4, "Plastic Specular");
                r = decodeTexture(mReg, d.roughnessTexture, 1, "Plastic Roughness");
                if (!r.textured && !std::isfinite(r.value[0])) raise("plastic roughness must be finite");
                m.type = PG_MAT_PLASTIC;
                break;
            }
            case Surface::Kind::Mirror: {
                expectSymbol(prog.init, "mirrorInit", what);
                const auto d = unpack<MirrorData>(prog.payload, "MirrorData");
                a = decodeTexture(mReg, d.reflectionTexture, 4, "Mirror Reflection");
                m.type = PG_MAT_MIRROR;
                break;
            }
            case Surface::Kind::Glass:
            case Surface::Kind::Substrate: {
                const bool glass = s.kind == Surface::Kind::Glass;
                expectSymbol(prog.init, glass ? "glassInit" : "substrateInit", what);
                const auto d = unpack<GlassData>(prog.payload, glass ? "GlassData" : "SubstrateData");  // same layout
                a = decodeTexture(mReg, d.reflection, 4, glass ? "Glass Reflection" : "Substrate Diffuse");
                b = decodeTexture(mReg, d.transmission, 4, glass ? "Glass Transmission" : "Substrate Specular");
                r = decodeTexture(mReg, d.roughnessX, 1, "RoughnessX");
                ry = decodeTexture(mReg, d.roughnessY, 1, "RoughnessY");
                for (const Input* in : {&r, &ry}) {
                    if (!in->textured && !std::isfinite(in->value[0])) raise("roughness must be finite");
                    if (glass && !in->textured && in->value[0] < 0.0f)
                        raise("glass roughness must be finite and not negative");
                }
                m.type = glass ? PG_MAT_GLASS : PG_MAT_SUBSTRATE;
                break;
            }
            case Surface::Kind::BlackBody:  // no lobe: a path that reaches it ends (blackBodySample)
                expectSymbol(prog.init, "blackBodyInit", what);
                m.type = PG_MAT_LAMBERT;
                break;
        }
        if (!a.textured) std::copy(a.value.begin(), a.value.begin() + 3, m.albedo);
        if (!b.textured) std::copy(b.value.begin(), b.value.begin() + 3, m.reserved);
        m.alpha = r.textured ? 0.0f : r.value[0];
        out.materials.push_back(m);
        out.materialTextures.push_back({textureIndex(a), textureIndex(b), textureIndex(r)});
        // RoughnessY: glass and substrate only; the others repeat their one roughness
        const bool two = s.kind == Surface::Kind::Glass || s.kind == Surface::Kind::Substrate;
        out.roughnessY.push_back(two && !ry.textured ? ry.value[0] : m.alpha);
        out.roughnessYTextures.push_back(two ? textureIndex(ry) : PG_TEX_NONE);
        mSurfaces.push_back(&s);
        return static_cast<uint32_t>(out.materials.size() - 1);
    }

    void light(const Light& l, TraversalHandle traversal) {
        const LightProgram prog = l.materialize(traversal, mReg.context(mTracer));
        const std::string init = symbolOf(prog.init);
        if (init == "SEInit") {
            const auto d = unpack<SampledEnvironmentData>(prog.payload, "SampledEnvironmentData");
            if (mEnvironments++) raise("Only one infinite light is supported.");  // Embree.cpp:1179-1183
            out.env = {d.texture.x, d.texture.y, d.texture.z};
        } else if (init == "pointInit") {  // pointInit: the node's origin in world space
            const auto d = unpack<PointLightData>(prog.payload, "PointLightData");
            const auto& m = Collected::transformOf(d.traversal);
            pg_point_light p{};
            p.position[0] = m[3], p.position[1] = m[7], p.position[2] = m[11];
            p.intensity[0] = d.intensity.x, p.intensity[1] = d.intensity.y, p.intensity[2] = d.intensity.z;
            out.points.push_back(p);
        } else if (init == "diffuseInit") {
            const auto d = unpack<DiffuseAreaLightData>(prog.payload, "DiffuseAreaLightData");
            const auto& shape = mReg.call(d.sample.handle, "AreaLight sample");
            if (shape.symbol != "planeSample") raise("AreaLight: only BasicShapes Plane geometry is supported");
            const auto cdf = unpack<CDFData>(shape.payload, "CDFData");
            if (cdf.size == 0) raise("CDFData: a Plane without primitives");
            const auto& bytes = mReg.buffer(cdf.primitives, "CDFData primitives");
            if (bytes.size() < cdf.size * sizeof(PerPlaneData)) raise("CDFData: primitive buffer too small");
            // every primitive of the Plane, in order; the library rebuilds the area CDF from them as
            // the Plane constructor does (BasicShapes.cpp:59-94)
            const Mat34& m = Collected::transformOf(cdf.traversal);
            for (uint32_t i = 0; i < cdf.size; ++i) {
                PerPlaneData p;
                std::memcpy(&p, bytes.data() + i * sizeof(PerPlaneData), sizeof(p));
                const auto o = applyPoint(m, &p.origin.x), u = applyVector(m, &p.u.x), v = applyVector(m, &p.v.x);
                pg_quad q{};
                for (int k = 0; k < 3; ++k) q.origin[k] = o[k], q.u[k] = u[k], q.v[k] = v[k];
                q.radiance[0] = d.radiance.x, q.radiance[1] = d.radiance.y, q.radiance[2] = d.radiance.z;
                out.quads.push_back(q);
            }
            out.lightQuads.push_back(cdf.size);
        } else {
            raise("light: " + init + " is not a light program");
        }
    }

    pg_camera camera(const Sensor& s, TraversalHandle traversal) {
        const SensorProgram prog = s.materialize(traversal, mReg.context(mTracer));
        expectSymbol(prog.rayGen, "rayGen", "sensor");
        const auto d = unpack<PCData>(prog.payload, "PCData");
        const Mat34& m = Collected::transformOf(d.traversal);
        pg_camera c{};
        c.position[0] = m[3], c.position[1] = m[7], c.position[2] = m[11];  // the node's translation
        c.look_at[0] = d.lookAt.x, c.look_at[1] = d.lookAt.y, c.look_at[2] = d.lookAt.z;
        c.up[0] = d.upRef.x, c.up[1] = d.upRef.y, c.up[2] = d.upRef.z;
        c.sensor_size[0] = d.size[0], c.sensor_size[1] = d.size[1];
        c.focal_length = d.focalLength;
        c.aperture_radius = d.apertureRadius;
        return c;
    }

    SceneExport out;

private:
    uint32_t textureIndex(const Input& in) {
        if (!in.textured) return PG_TEX_NONE;
        for (size_t i = 0; i < out.textures.size(); ++i) {
            pg_texture t = out.textures[i];
            t.texels = nullptr;
            if (std::memcmp(&t, &in.tex, sizeof(pg_texture)) == 0 && out.textureTexels[i] == in.texels)
                return static_cast<uint32_t>(i);
        }
        out.textures.push_back(in.tex);
        out.textureTexels.push_back(in.texels);
        return static_cast<uint32_t>(out.textures.size() - 1);
    }

    Tracer& mTracer;
    Registry& mReg;
    std::vector<const Surface*> mSurfaces;
    uint32_t mEnvironments = 0;
};

// Materialises and decodes a scene graph: instances (world-space triangles, one material each),
// lights in the pipeline's order (the environment first, Embree.cpp:1175-1188) and sensors.
struct DecodedScene {
    SceneExport e;
    std::vector<std::shared_ptr<Light>> lights;  // pipeline order
};
DecodedScene decodeScene(Tracer& tracer, Registry& reg, const std::shared_ptr<Node>& scene) {
    Collected c;
    collect(scene, identity34(), c);
    SceneDecoder dec(tracer, reg);
    SceneExport& f = dec.out;
    // static triangles first, then each moving instance's (pg_scene_create_instanced's layout)
    std::stable_sort(c.instances.begin(), c.instances.end(),
                     [](const Collected::Inst& a, const Collected::Inst& b) { return a.mover < b.mover; });
    int lastMover = -1;
    for (const auto& [inst, toWorld, mover] : c.instances) {
        if (mover >= 0 && mover != lastMover) {  // a moving instance's triangles start here
            f.instanceFirst.push_back(static_cast<uint32_t>(f.index.size() / 3));
            f.motions.push_back({c.movers[mover].keys, c.movers[mover].parent});
            lastMover = mover;
        }
        const uint32_t base = static_cast<uint32_t>(f.verts.size() / 3);
        for (size_t v = 0; v < inst->mesh->verts.size(); v += 3) {
            const auto p = applyPoint(toWorld, &inst->mesh->verts[v]);
            f.verts.insert(f.verts.end(), p.begin(), p.end());
        }
        if (!inst->mesh->uv.empty() || !f.texcoords.empty()) {
            f.texcoords.resize(2 * static_cast<size_t>(base), 0.0f);
            if (inst->mesh->uv.empty()) f.texcoords.resize(f.verts.size() / 3 * 2, 0.0f);
            else f.texcoords.insert(f.texcoords.end(), inst->mesh->uv.begin(), inst->mesh->uv.end());
        }
        const uint32_t obj = static_cast<uint32_t>(f.objectMaterial.size());
        f.objectMaterial.push_back(dec.material(*inst->surface));
        for (uint32_t i : inst->mesh->index) f.index.push_back(base + i);
        f.triObject.insert(f.triObject.end(), inst->mesh->index.size() / 3, obj);
    }
    DecodedScene out;
    auto env = std::stable_partition(c.lights.begin(), c.lights.end(),
                                     [](const Collected::Lit& l) { return l.light->kind == Light::Kind::Environment; });
    (void)env;
    for (const auto& l : c.lights) {
        dec.light(*l.light, l.traversal);
        out.lights.push_back(l.light);
    }
    for (const auto& s : c.sensors) {
        f.sensors.push_back({s.node, Collected::transformOf(s.traversal)});
        f.cameras.push_back(dec.camera(*s.sensor, s.traversal));
    }
    if (f.textures.empty()) {
        f.texcoords.clear();
        f.materialTextures.clear();
    }
    for (size_t i = 0; i < f.textures.size(); ++i)
        f.textures[i].texels = f.textureTexels[i].empty() ? nullptr : f.textureTexels[i].data();
    out.e = std::move(f);
    return out;
}

class HIPTraceLauncher final : public TraceLauncher {
public:
    struct Tile {
        RenderRECT rect;
        std::shared_ptr<Resource> film;
    };
    struct Shared {  // pending tiles, flushed by the first future that is polled
        HIPAccelerator* accelerator;
        std::shared_ptr<Kernel> kernel;
        pg_scene* scene;
        pg_render_params base;
        std::vector<Tile> pending;
        std::vector<pg_stats> stats;
        std::string error;  // of the last updateTimeInterval
    };

    HIPTraceLauncher(HIPAccelerator& acc, std::shared_ptr<Kernel> kernel, std::shared_ptr<pg_scene> scene, uint32_t w,
                     uint32_t h, RenderRECT rect, const pg_render_params& base)
        : mScene(std::move(scene)), mW(w), mH(h), mRect(rect), mShared(std::make_shared<Shared>()) {
        mShared->accelerator = &acc;
        mShared->kernel = std::move(kernel);
        mShared->scene = mScene.get();
        mShared->base = base;
    }
    [[nodiscard]] std::pair<uint32_t, uint32_t> getFilmResolution() const noexcept override { return {mW, mH}; }
    [[nodiscard]] RenderRECT getRenderRECT() const noexcept override { return mRect; }
    // Sets every moving instance's keys for the frame (the ray time is Sobol dimension 2 within
    // it); a frame the reference cannot render is reported by the next launch.
    void updateTimeInterval(float begin, float end) noexcept override {
        mTime = {begin, end};
        mShared->error.clear();
        try {
            for (size_t k = 0; k < mMotions.size(); ++k) {
                const pg_motion m = motionForInterval(*mMotions[k].keys, mMotions[k].parent, begin, end);
                check(pg_scene_set_motion(mScene.get(), static_cast<uint32_t>(k), &m), "pg_scene_set_motion");
            }
        } catch (const std::exception& e) {
            mShared->error = e.what();
        }
    }
    void setMotions(std::vector<SceneExport::Motion> motions) { mMotions = std::move(motions); }

    [[nodiscard]] FutureVoid launch(const RenderRECT& rect, const std::function<SBTPayload(uint32_t)>& launchData,
                                    const std::vector<std::shared_ptr<Resource>>& resources) override {
        if (!mShared->error.empty()) raise(mShared->error);
        const auto ld = unpack<LaunchData>(launchData(0), "launch: LaunchData");
        if (ld.w != mW || ld.h != mH) raise("launch: LaunchData resolution differs from the film");
        if (ld.rgbw >= resources.size()) raise("launch: LaunchData names a missing resource");
        auto* film = dynamic_cast<HIPTiledOutput*>(resources[ld.rgbw].get());
        if (!film) raise("launch: the output was not created by HIPAccelerator");
        if (rect.width == 0 || rect.height == 0 || rect.left + rect.width > mW || rect.top + rect.height > mH)
            raise("launch: tile outside the film");
        film->bind(mW, mH);
        mShared->pending.push_back({rect, resources[ld.rgbw]});
        std::shared_ptr<Shared> shared = mShared;
        std::shared_ptr<pg_scene> keep = mScene;
        return FutureVoid([shared, keep] {
            flush(*shared);
            return Unit{};
        });
    }
    std::vector<pg_stats> stats() const { return mShared->stats; }

private:
    // Tiles of one film that exactly cover their bounding rect render as one launch over that rect
    // (the same film as one pg_render of it); otherwise each tile launches on its own. A launch is
    // the pipeline's kernel through Accelerator::launchKernel over Dim3{rect w, rect h, spp}.
    static void flush(Shared& s) {
        while (!s.pending.empty()) {
            const std::shared_ptr<Resource> film = s.pending.front().film;
            std::vector<RenderRECT> tiles;
            std::vector<Tile> rest;
            for (const Tile& t : s.pending) (t.film == film ? tiles.push_back(t.rect) : rest.push_back(t));
            s.pending.swap(rest);
            uint32_t x0 = UINT32_MAX, y0 = UINT32_MAX, x1 = 0, y1 = 0;
            for (const RenderRECT& r : tiles)
                x0 = std::min(x0, r.left), y0 = std::min(y0, r.top), x1 = std::max(x1, r.left + r.width),
                y1 = std::max(y1, r.top + r.height);
            std::vector<uint8_t> cover(static_cast<size_t>(x1 - x0) * (y1 - y0), 0);
            bool exact = true;
            for (const RenderRECT& r : tiles)
                for (uint32_t y = r.top; y < r.top + r.height && exact; ++y)
                    for (uint32_t x = r.left; x < r.left + r.width; ++x) {
                        uint8_t& c = cover[static_cast<size_t>(y - y0) * (x1 - x0) + (x - x0)];
                        if (c) {
                            exact = false;
                            break;
                        }
                        c = 1;
                    }
            exact = exact && std::all_of(cover.begin(), cover.end(), [](uint8_t c) { return c != 0; });
            if (exact) tiles = {RenderRECT{x0, y0, x1 - x0, y1 - y0}};
            auto lut = s.accelerator->createResourceLUT({film});
            for (const RenderRECT& r : tiles) {
                pg_render_params p = s.base;
                p.rect[0] = r.left, p.rect[1] = r.top, p.rect[2] = r.width, p.rect[3] = r.height;
                s.stats.push_back(pg_stats{});
                s.accelerator
                    ->launchKernel(Dim3{r.width, r.height, p.spp_end - p.spp_begin}, s.kernel, lut,
                                   reinterpret_cast<uint64_t>(s.scene), p,
                                   reinterpret_cast<uint64_t>(&s.stats.back()))
                    .getSync();
            }
        }
    }

    std::shared_ptr<pg_scene> mScene;
    uint32_t mW, mH;
    RenderRECT mRect;
    std::pair<float, float> mTime{0.0f, 0.0f};
    std::vector<SceneExport::Motion> mMotions;
    std::shared_ptr<Shared> mShared;
};

// The pipeline's kernel (entry "piperHIPRender"): args = (pg_scene* as uint64, pg_render_params,
// pg_stats* as uint64), resources = {the film}; renders the params' rect x sample range.
void hipRenderKernel(const NativeLaunch& l) {
    auto* scene = reinterpret_cast<pg_scene*>(l.args.get<uint64_t>(0));
    const auto p = l.args.get<pg_render_params>(1);
    auto* stats = reinterpret_cast<pg_stats*>(l.args.get<uint64_t>(2));
    if (l.resources.size() != 1) raise("piperHIPRender: expects one resource, the film");
    auto* film = dynamic_cast<HIPTiledOutput*>(l.resources[0].get());
    if (!film) raise("piperHIPRender: the resource is not a tiled output of this accelerator");
    if (l.size.x != p.rect[2] || l.size.y != p.rect[3] || l.size.z != p.spp_end - p.spp_begin)
        raise("piperHIPRender: launch size differs from the render rect and sample range");
    auto lock = l.context.makeCurrent();
    check(pg_render(scene, &p, film->bind(p.width, p.height), stats), "pg_render");
}

class HIPPipeline final : public Pipeline {
public:
    HIPPipeline(HIPAccelerator& acc, DecodedScene scene, pg_render_params base, std::shared_ptr<Sampler> sampler)
        : mAcc(acc), mScene(std::move(scene)), mBase(base), mSampler(std::move(sampler)) {
        // Embree.cpp:1278-1284 compiles piperMain with the components' static bindings; this
        // pipeline links the HIP built-ins and enters through its render kernel
        mKernel = mAcc.compileKernel({hipBuiltinProgram()}, {{"piperMain", "piperHIPRender"}}, {}, "piperMain");
    }

    [[nodiscard]] std::string generateStatisticsReport() const override {
        uint64_t paths = 0, ext = 0, shadow = 0, mis = 0, calls = 0;
        double ms = 0.0;
        for (const auto& l : mLaunchers)
            if (auto sp = l.lock())
                for (const pg_stats& s : sp->stats())
                    paths += s.paths, ext += s.extension_rays, shadow += s.shadow_rays, mis += s.mis_rays,
                        ms += s.ms_total, ++calls;
        std::ostringstream o;
        o << "HIP pipeline: " << calls << " render calls, " << paths << " paths, " << ext << " extension rays, "
          << shadow << " shadow rays, " << mis << " MIS rays, " << ms << " ms";
        if (ms > 0.0) o << " (" << static_cast<double>(paths) / ms / 1e3 << " Mpaths/s)";
        return o.str();
    }

    // Pipeline::prepare: the scene goes to the device with the camera of `sensor` (its node's
    // translation is the camera position, PerspectiveCamera.cpp:39-45), the sampler's payload for
    // the film (SobolData, ScrambledSobolSampler.cpp:66-72), then the film rect and NDC transform
    // follow calcRenderRECT.
    [[nodiscard]] std::shared_ptr<TraceLauncher> prepare(const std::shared_ptr<Node>& sensor, uint32_t width,
                                                         uint32_t height, FitMode fitMode) override {
        const SceneExport& f = mScene.e;
        const pg_camera* cam = nullptr;
        for (size_t i = 0; i < f.sensors.size(); ++i)
            if (f.sensors[i].first == sensor.get()) cam = &f.cameras[i];
        if (!cam) raise("prepare: the sensor node is not part of the pipeline's scene");
        const SamplerAttributes sa = mSampler->generatePayload(width, height);
        const auto sobol = unpack<SobolData>(sa.payload, "SobolData");
        uint32_t m = 0;
        while ((1u << m) < std::max(width, height)) ++m;
        if (sobol.log2Resolution != m || sobol.resolution != (1u << m))
            raise("prepare: the sampler's resolution differs from the film's");
        pg_scene_desc d{};
        d.vertices = f.verts.data();
        d.num_vertices = static_cast<uint32_t>(f.verts.size() / 3);
        d.indices = f.index.data();
        d.num_triangles = static_cast<uint32_t>(f.index.size() / 3);
        d.tri_object = f.triObject.data();
        d.object_material = f.objectMaterial.data();
        d.num_objects = static_cast<uint32_t>(f.objectMaterial.size());
        d.materials = f.materials.data();
        d.num_materials = static_cast<uint32_t>(f.materials.size());
        d.quads = f.quads.data();
        d.num_quads = static_cast<uint32_t>(f.quads.size());
        d.area_light_quads = f.lightQuads.empty() ? nullptr : f.lightQuads.data();
        d.num_area_lights = static_cast<uint32_t>(f.lightQuads.size());
        if (f.anisotropic()) {
            d.roughness_y = f.roughnessY.data();
            d.roughness_y_textures = f.roughnessYTextures.data();
        }
        d.point_lights = f.points.empty() ? nullptr : f.points.data();
        d.num_point_lights = static_cast<uint32_t>(f.points.size());
        if (!f.textures.empty()) {
            d.textures = f.textures.data();
            d.num_textures = static_cast<uint32_t>(f.textures.size());
            d.material_textures = f.materialTextures.data();
            d.texcoords = f.texcoords.empty() ? nullptr : f.texcoords.data();
        }
        d.camera = *cam;
        for (int k = 0; k < 3; ++k) d.env_radiance[k] = f.env[k];
        if (d.num_materials == 0) {  // a scene of lights only still needs one material
            static const pg_material none{PG_MAT_LAMBERT, {0.0f, 0.0f, 0.0f}, 0.0f, {0.0f, 0.0f, 0.0f}};
            d.materials = &none;
            d.num_materials = 1;
        }
        pg_scene* raw = nullptr;
        if (f.instanceFirst.empty())
            check(pg_scene_create(mAcc.device(), &d, &raw), "pg_scene_create");
        else
            check(pg_scene_create_instanced(mAcc.device(), &d, f.instanceFirst.data(),
                                            static_cast<uint32_t>(f.instanceFirst.size()), &raw),
                  "pg_scene_create_instanced");
        std::shared_ptr<pg_scene> scene(raw, pg_scene_destroy);
        const auto [ndc, rect] = calcRenderRECT(width, height, cam->sensor_size[0] / cam->sensor_size[1], fitMode);
        pg_render_params p = mBase;
        p.width = width, p.height = height, p.spp = sa.samplesPerPixel;
        p.spp_begin = 0, p.spp_end = sa.samplesPerPixel;
        p.ndc[0] = ndc.ox, p.ndc[1] = ndc.oy, p.ndc[2] = ndc.sx, p.ndc[3] = ndc.sy;
        auto launcher = std::make_shared<HIPTraceLauncher>(mAcc, mKernel, scene, width, height, rect, p);
        launcher->setMotions(f.motions);
        mLaunchers.push_back(launcher);
        return launcher;
    }

private:
    HIPAccelerator& mAcc;
    DecodedScene mScene;
    pg_render_params mBase;
    std::shared_ptr<Sampler> mSampler;
    std::shared_ptr<Kernel> mKernel;
    std::vector<std::weak_ptr<HIPTraceLauncher>> mLaunchers;
};

}  // namespace

// ---- free functions ----
std::pair<SensorNDCAffineTransform, RenderRECT> calcRenderRECT(uint32_t width, uint32_t height,
                                                                float deviceAspectRatio, FitMode fitMode) {
    RenderRECT rect{};
    SensorNDCAffineTransform t{};
    const float imageAspectRatio = static_cast<float>(width) / static_cast<float>(height);
    const float iiar = 1.0f / imageAspectRatio;
    const float idar = 1.0f / deviceAspectRatio;
    if (fitMode == FitMode::Fill) {
        rect = {0, 0, width, height};
        if (imageAspectRatio > deviceAspectRatio)
            t = {0.0f, (idar - iiar) * 0.5f * deviceAspectRatio, 1.0f, iiar * deviceAspectRatio};
        else
            t = {(deviceAspectRatio - imageAspectRatio) * 0.5f * idar, 0.0f, imageAspectRatio * idar, 1.0f};
    } else if (imageAspectRatio > deviceAspectRatio) {
        t = {-(imageAspectRatio - deviceAspectRatio) * 0.5f * idar, 0.0f, imageAspectRatio * idar, 1.0f};
        rect.left = static_cast<uint32_t>(
            std::floor(std::max(0.0f, static_cast<float>(width) * (imageAspectRatio - deviceAspectRatio) * 0.5f * iiar)));
        rect.top = 0;
        rect.width = std::min(width, static_cast<uint32_t>(std::ceil(static_cast<float>(width) *
                                                                      (imageAspectRatio + deviceAspectRatio) * 0.5f * iiar)));
        rect.height = height;
        rect.width -= rect.left;
    } else {
        t = {0.0f, -(iiar - idar) * 0.5f * deviceAspectRatio, 1.0f, deviceAspectRatio * iiar};
        rect.left = 0;
        rect.top = static_cast<uint32_t>(
            std::floor(std::max(0.0f, static_cast<float>(height) * (iiar - idar) * 0.5f * imageAspectRatio)));
        rect.width = width;
        rect.height = static_cast<uint32_t>(std::ceil(static_cast<float>(height) * (iiar + idar) * 0.5f * imageAspectRatio));
        rect.height -= rect.top;
    }
    return {t, rect};
}

TransformSRT identitySRT() { return {1, 1, 1, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0}; }

LinkableProgram makeNativeLinkable(const std::vector<std::pair<std::string, const void*>>& symbols, uint64_t uid) {
    LinkableProgram p;
    p.format = "Native";
    p.UID = uid;
    auto put = [&](const void* src, size_t n) {
        const auto* b = static_cast<const std::byte*>(src);
        p.exchange.insert(p.exchange.end(), b, b + n);
    };
    put("Native", 6);
    for (const auto& [name, address] : symbols) {
        put(name.c_str(), name.size() + 1);
        const uint64_t a = reinterpret_cast<uint64_t>(address);
        put(&a, 8);
    }
    return p;
}

std::unordered_map<std::string, const void*> parseNativeLinkable(const LinkableProgram& program) {
    if (program.format != "Native") raise("Unsupported linkable format " + program.format);
    const auto& d = program.exchange;
    if (d.size() < 6 || std::memcmp(d.data(), "Native", 6) != 0) raise("Bad header in native linkable file.");
    std::unordered_map<std::string, const void*> out;
    size_t i = 6;
    while (i < d.size()) {
        const size_t b = i;
        while (i < d.size() && d[i] != std::byte{0}) ++i;
        if (i + 9 > d.size()) raise("The native linkable file is corrupted.");
        std::string name(reinterpret_cast<const char*>(d.data() + b), i - b);
        uint64_t a;
        std::memcpy(&a, d.data() + i + 1, 8);
        out[std::move(name)] = reinterpret_cast<const void*>(a);
        i += 9;
    }
    return out;
}

const LinkableProgram& hipBuiltinProgram() {
    static const LinkableProgram program = [] {
        std::vector<std::pair<std::string, const void*>> symbols;
        for (const BuiltinMarker& m : builtinMarkers()) symbols.push_back({m.name, &m});
        symbols.push_back({"piperHIPRender", reinterpret_cast<const void*>(&hipRenderKernel)});
        return makeNativeLinkable(symbols, 0x4849504275696C74ull);  // "HIPBuilt"
    }();
    return program;
}

const std::vector<std::string>& hipBuiltinSymbols() {
    static const std::vector<std::string> s = [] {
        std::vector<std::string> v;
        for (const BuiltinMarker& m : builtinMarkers()) v.push_back(m.name);
        return v;
    }();
    return s;
}

Config constantTextureConfig(const std::vector<float>& value) {
    std::vector<Config> v;
    for (float x : value) v.push_back(Config::makeNumber(x));
    return Config::makeObject({{"ClassID", Config::makeString("Piper.BuiltinComponent.ProceduralTexture.ConstantTexture")},
                               {"Value", Config::makeArray(std::move(v))}});
}
Config colourConfig(std::array<float, 3> rgb) { return constantTextureConfig({rgb[0], rgb[1], rgb[2], 1.0f}); }
Config scalarConfig(float v) { return constantTextureConfig({v}); }
Config checkerboardConfig(const std::vector<float>& black, const std::vector<float>& white, float scale) {
    std::vector<Config> b, w;
    for (float x : black) b.push_back(Config::makeNumber(x));
    for (float x : white) w.push_back(Config::makeNumber(x));
    return Config::makeObject({{"ClassID", Config::makeString("Piper.BuiltinComponent.ProceduralTexture.CheckBoard")},
                               {"Black", Config::makeArray(std::move(b))},
                               {"White", Config::makeArray(std::move(w))},
                               {"Scale", Config::makeNumber(scale)}});
}

// ---- components ----
Surface::Surface(Kind k, std::vector<Config> in) : kind(k), inputs(std::move(in)) {
    static const size_t counts[] = {2, 2, 3, 1, 4, 4, 0};
    if (inputs.size() != counts[static_cast<int>(kind)]) raise("Surface: wrong number of inputs for its class");
}
std::shared_ptr<Surface> Surface::matte(std::array<float, 3> reflectance, float roughness) {
    if (!(roughness >= 0.0f && roughness <= 1.0f)) raise("matte roughness must be in [0, 1]");
    return std::make_shared<Surface>(Kind::Matte, std::vector<Config>{colourConfig(reflectance), scalarConfig(roughness)});
}
std::shared_ptr<Surface> Surface::conductor(std::array<float, 3> f0, float roughness) {
    if (!(roughness > 0.0f)) raise("conductor roughness must be positive");
    return std::make_shared<Surface>(Kind::Conductor, std::vector<Config>{colourConfig(f0), scalarConfig(roughness)});
}
std::shared_ptr<Surface> Surface::conductorAlpha(std::array<float, 3> f0, float alpha) {
    if (!(alpha > 0.0f)) raise("conductor alpha must be positive");
    auto s = std::make_shared<Surface>(Kind::Conductor, std::vector<Config>{colourConfig(f0), scalarConfig(alpha)});
    s->roughnessIsAlpha = true;
    return s;
}
std::shared_ptr<Surface> Surface::plastic(std::array<float, 3> diffuse, std::array<float, 3> specular, float roughness) {
    if (!std::isfinite(roughness)) raise("plastic roughness must be finite");
    return std::make_shared<Surface>(
        Kind::Plastic, std::vector<Config>{colourConfig(diffuse), colourConfig(specular), scalarConfig(roughness)});
}
std::shared_ptr<Surface> Surface::mirror(std::array<float, 3> reflection) {
    return std::make_shared<Surface>(Kind::Mirror, std::vector<Config>{colourConfig(reflection)});
}
std::shared_ptr<Surface> Surface::glass(std::array<float, 3> reflection, std::array<float, 3> transmission,
                                        float roughnessX, float roughnessY) {
    for (float v : {roughnessX, roughnessY})
        if (!(v >= 0.0f) || !std::isfinite(v)) raise("glass roughness must be finite and not negative");
    return std::make_shared<Surface>(Kind::Glass, std::vector<Config>{colourConfig(reflection), colourConfig(transmission),
                                                                      scalarConfig(roughnessX), scalarConfig(roughnessY)});
}
std::shared_ptr<Surface> Surface::substrate(std::array<float, 3> diffuse, std::array<float, 3> specular,
                                            float roughnessX, float roughnessY) {
    if (!std::isfinite(roughnessX) || !std::isfinite(roughnessY)) raise("substrate roughness must be finite");
    return std::make_shared<Surface>(Kind::Substrate, std::vector<Config>{colourConfig(diffuse), colourConfig(specular),
                                                                          scalarConfig(roughnessX), scalarConfig(roughnessY)});
}

// Classic.cpp:79-200: the class's four programs and its payload of texture calls (colours are
// loaded with 4 channels, roughness with 1).
SurfaceProgram Surface::materialize(const MaterializeContext& ctx) const {
    static const char* names[] = {"matte", "conductor", "plastic", "mirror", "glass", "substrate", "blackBody"};
    const std::string n = names[static_cast<int>(kind)];
    const LinkableProgram& linkable = hipBuiltinProgram();
    SurfaceProgram res;
    res.init = ctx.tracer.buildProgram(linkable, n + "Init");
    res.sample = ctx.tracer.buildProgram(linkable, n + "Sample");
    const bool delta = kind == Kind::Mirror;  // the mirror evaluates as a black body (Classic.cpp:163-166)
    res.evaluate = ctx.tracer.buildProgram(linkable, delta ? "blackBodyEvaluate" : n + "Evaluate");
    res.pdf = ctx.tracer.buildProgram(linkable, delta ? "blackBodyPdf" : n + "Pdf");
    auto tex = [&](size_t i, uint32_t channels) { return Call<TextureSampleFunc>{ctx.loadTexture(inputs[i], channels)}; };
    switch (kind) {
        case Kind::Matte: res.payload = packSBTPayload(MatteData{tex(0, 4), tex(1, 1)}); break;
        case Kind::Conductor:
            res.payload = packSBTPayload(ConductorData{tex(0, 4), tex(1, 1), roughnessIsAlpha ? 1u : 0u});
            break;
        case Kind::Plastic: res.payload = packSBTPayload(PlasticData{tex(0, 4), tex(1, 4), tex(2, 1)}); break;
        case Kind::Mirror: res.payload = packSBTPayload(MirrorData{tex(0, 4)}); break;
        case Kind::Glass: res.payload = packSBTPayload(GlassData{tex(0, 4), tex(1, 4), tex(2, 1), tex(3, 1)}); break;
        case Kind::Substrate:
            res.payload = packSBTPayload(SubstrateData{tex(0, 4), tex(1, 4), tex(2, 1), tex(3, 1)});
            break;
        case Kind::BlackBody: break;
    }
    return res;
}

std::shared_ptr<Light> Light::areaPlane(std::array<float, 3> origin, std::array<float, 3> u, std::array<float, 3> v,
                                        std::array<float, 3> radiance) {
    return area({{origin, u, v}}, radiance);
}
std::shared_ptr<Light> Light::area(std::vector<PlanePrimitive> primitives, std::array<float, 3> radiance) {
    if (primitives.empty()) raise("a Plane without primitives");
    auto l = std::make_shared<Light>();
    l->kind = Kind::Area, l->primitives = std::move(primitives), l->radiance = radiance;
    return l;
}
std::shared_ptr<Light> Light::point(std::array<float, 3> intensity) {
    auto l = std::make_shared<Light>();
    l->kind = Kind::Point, l->radiance = intensity;
    return l;
}
std::shared_ptr<Light> Light::environment(std::array<float, 3> radiance) {
    auto l = std::make_shared<Light>();
    l->kind = Kind::Environment, l->radiance = radiance;
    return l;
}

// AreaLight.cpp:44-56 with the Plane's sampled-geometry program (BasicShapes.cpp:61-95, 135-170):
// per primitive origin, U, V, unit normal U x V, unit tangent U and the axis of U x V's largest
// component; the area CDF and PDF over the primitives.
LightProgram Light::materialize(TraversalHandle traversal, const MaterializeContext& ctx) const {
    const LinkableProgram& linkable = hipBuiltinProgram();
    LightProgram res;
    const Vec3 rad{radiance[0], radiance[1], radiance[2]};
    if (kind == Kind::Environment) {
        res.init = ctx.tracer.buildProgram(linkable, "SEInit");
        res.sample = ctx.tracer.buildProgram(linkable, "SESample");
        res.evaluate = ctx.tracer.buildProgram(linkable, "SEEvaluate");
        res.pdf = ctx.tracer.buildProgram(linkable, "SEPdf");
        res.payload = packSBTPayload(SampledEnvironmentData{traversal, rad});
        return res;
    }
    if (kind == Kind::Point) {
        res.init = ctx.tracer.buildProgram(linkable, "pointInit");
        res.sample = ctx.tracer.buildProgram(linkable, "pointSample");
        res.evaluate = ctx.tracer.buildProgram(linkable, "deltaEvaluate");
        res.pdf = ctx.tracer.buildProgram(linkable, "deltaPdf");
        res.payload = packSBTPayload(PointLightData{traversal, rad});
        return res;
    }
    res.init = ctx.tracer.buildProgram(linkable, "diffuseInit");
    res.sample = ctx.tracer.buildProgram(linkable, "diffuseSample");
    res.evaluate = ctx.tracer.buildProgram(linkable, "diffuseEvaluate");
    res.pdf = ctx.tracer.buildProgram(linkable, "diffusePdf");
    std::vector<PerPlaneData> planes;
    std::vector<float> cdf, pdf;
    float total = 0.0f;
    for (const auto& p : primitives) {
        const Vec3 u{p.u[0], p.u[1], p.u[2]}, v{p.v[0], p.v[1], p.v[2]};
        const Vec3 det{u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x};
        const float area = std::sqrt(det.x * det.x + det.y * det.y + det.z * det.z);
        const float ul = std::sqrt(u.x * u.x + u.y * u.y + u.z * u.z);
        const float a0 = std::fabs(det.x), a1 = std::fabs(det.y), a2 = std::fabs(det.z);
        const uint32_t sel = (a0 >= a1 && a0 >= a2) ? 0u : ((a1 >= a0 && a1 >= a2) ? 1u : 2u);
        planes.push_back({{p.origin[0], p.origin[1], p.origin[2]}, u, v, {det.x / area, det.y / area, det.z / area},
                          {u.x / ul, u.y / ul, u.z / ul}, sel});
        total += area;
        cdf.push_back(total);
        pdf.push_back(area);
    }
    for (float& c : cdf) c /= total;
    for (float& q : pdf) q /= total;
    auto buffer = [&](const void* src, size_t n) {
        return ctx.accelerator.createBuffer(n, 16, [src, n](std::byte* dst) { std::memcpy(dst, src, n); });
    };
    SampledGeometryProgram shape;
    shape.sample = ctx.tracer.buildProgram(linkable, "planeSample");
    shape.payload = packSBTPayload(CDFData{ctx.registerResource(buffer(planes.data(), planes.size() * sizeof(PerPlaneData))),
                                           traversal, ctx.registerResource(buffer(cdf.data(), cdf.size() * 4)),
                                           ctx.registerResource(buffer(pdf.data(), pdf.size() * 4)), 1.0f / total,
                                           static_cast<uint32_t>(planes.size())});
    res.payload = packSBTPayload(DiffuseAreaLightData{rad, {ctx.registerCall(shape.sample, shape.payload)}, total});
    return res;
}

SensorProgram Sensor::materialize(TraversalHandle traversal, const MaterializeContext& ctx) const {
    SensorProgram res;  // PerspectiveCamera.cpp:51-61
    res.rayGen = ctx.tracer.buildProgram(hipBuiltinProgram(), "rayGen");
    res.payload = packSBTPayload(PCData{traversal,
                                        {lookAt[0], lookAt[1], lookAt[2]},
                                        {size[0], size[1]},
                                        {upRef[0], upRef[1], upRef[2]},
                                        focalLength,
                                        apertureRadius});
    return res;
}

IntegratorProgram Integrator::materialize(const MaterializeContext& ctx) const {
    IntegratorProgram res;  // PathTracing.cpp:42-56 (no profiler: the statistics handles are null)
    res.trace = ctx.tracer.buildProgram(hipBuiltinProgram(), "trace");
    res.payload = packSBTPayload(PathTracingData{maxDepth, minDepthForRussianRoulette, nullptr, nullptr, nullptr});
    return res;
}

void LightSampler::preprocess(const std::vector<std::shared_ptr<Light>>& lights) {
    if (lights.empty()) raise("Need environment");  // BasicLight.cpp:90-92
    deltaFlags.clear();
    for (const auto& l : lights) deltaFlags.push_back(l->isDelta() ? 1 : 0);
}
LightSamplerProgram LightSampler::materialize(const MaterializeContext& ctx) const {
    LightSamplerProgram res;  // BasicLight.cpp:97-111
    res.select = ctx.tracer.buildProgram(hipBuiltinProgram(), "uniformSelect");
    const uint32_t n = static_cast<uint32_t>(deltaFlags.size());
    res.payload.resize(4 + n);
    std::memcpy(res.payload.data(), &n, 4);
    std::memcpy(res.payload.data() + 4, deltaFlags.data(), n);
    return res;
}

SamplerProgram Sampler::materialize(const MaterializeContext& ctx) const {
    return {ctx.tracer.buildProgram(hipBuiltinProgram(), "sobolStart"),
            ctx.tracer.buildProgram(hipBuiltinProgram(), "sobolGenerate")};
}
SamplerAttributes Sampler::generatePayload(uint32_t sampleWidth, uint32_t sampleHeight) const {
    uint32_t log2 = 0;  // ScrambledSobolSampler.cpp:66-72
    while ((1u << log2) < std::max(sampleWidth, sampleHeight)) ++log2;
    uint32_t spp = 1;
    while (spp < samplesPerPixel) spp <<= 1;  // ScrambledSobolSampler.cpp:43-50
    return {packSBTPayload(SobolData{1u << log2, log2}), 1024u, spp};
}

FilterProgram SimpleFilter::materialize(const MaterializeContext& ctx) const {
    FilterProgram res;  // FixedSampler.cpp:36-53: the program is named by FilterType
    res.weight = ctx.tracer.buildProgram(hipBuiltinProgram(), filterType);
    GaussianFilterData g{0.0f, 0.0f};
    if (filterType == "gaussian") g = {-alpha, std::exp(-alpha)};
    res.payload = packSBTPayload(g);
    return res;
}
RenderDriverProgram FixedSampler::materialize(const MaterializeContext& ctx) const {
    RenderDriverProgram res;  // FixedSampler.cpp:135-143
    res.accumulate = ctx.tracer.buildProgram(hipBuiltinProgram(), "accumulate");
    const FilterProgram f = filter.materialize(ctx);
    res.payload = packSBTPayload(RDData{{ctx.registerCall(f.weight, f.payload)}});
    return res;
}

// ---- HIPAccelerator ----
HIPAccelerator::HIPAccelerator(int ordinal) {
    check(pg_device_create(ordinal, &mDevice), "pg_device_create");
    mContext = std::make_unique<HIPContext>(mDevice);
    mContexts = {mContext.get()};
}
std::shared_ptr<HIPAccelerator> HIPAccelerator::hostOnly() {
    std::shared_ptr<HIPAccelerator> a(new HIPAccelerator());
    a->mContext = std::make_unique<HIPContext>(nullptr);  // host kernels only: its handle is null
    a->mContexts = {a->mContext.get()};
    return a;
}
HIPAccelerator::~HIPAccelerator() {
    if (mDevice) pg_device_destroy(mDevice);
}
const char* HIPAccelerator::getNativePlatform() const noexcept {
    return mDevice ? pg_device_platform(mDevice) : "host (no device)";
}
std::vector<std::string> HIPAccelerator::getSupportedLinkableFormat() const { return {"Native"}; }

std::shared_ptr<Kernel> HIPAccelerator::compileKernel(const std::vector<LinkableProgram>& linkable,
                                                      std::unordered_map<std::string, std::string> staticRedirectedSymbols,
                                                      std::vector<std::string> dynamicSymbols, std::string entryFunction) {
    std::unordered_map<std::string, const void*> symbols;
    for (const LinkableProgram& p : linkable) {
        auto table = parseNativeLinkable(p);
        symbols.insert(table.begin(), table.end());
    }
    auto resolve = [&](std::string name) {
        const auto r = staticRedirectedSymbols.find(name);
        if (r != staticRedirectedSymbols.end()) name = r->second;
        const auto it = symbols.find(name);
        if (it == symbols.end()) raise("compileKernel: undefined symbol " + name);
        return it->second;
    };
    auto k = std::make_shared<NativeKernel>();
    const void* entry = resolve(entryFunction);
    for (const BuiltinMarker& m : builtinMarkers())
        if (entry == &m) raise("compileKernel: " + entryFunction + " is a device built-in, not a kernel");
    k->entry = reinterpret_cast<NativeKernelFunc>(const_cast<void*>(entry));
    for (const std::string& s : dynamicSymbols) k->dynamicSymbols.push_back(resolve(s));
    return k;
}

FutureVoid HIPAccelerator::launchKernelImpl(const Dim3& size, std::shared_ptr<Kernel> kernel,
                                            std::shared_ptr<ResourceLookUpTable> root, ArgumentPackage args) {
    auto k = std::dynamic_pointer_cast<NativeKernel>(kernel);
    if (!k) raise("launchKernel: the kernel was not compiled by HIPAccelerator");
    if (!root) root = createResourceLUT({});
    Context* ctx = mContext.get();
    return FutureVoid([k, root, args = std::move(args), size, ctx] {
        k->entry(NativeLaunch{size, *ctx, root->resources, args, k->dynamicSymbols});
        return Unit{};
    });
}

std::shared_ptr<Resource> HIPAccelerator::createBuffer(size_t size, size_t alignment,
                                                       std::function<void(std::byte*)> data) {
    if (alignment == 0 || (alignment & (alignment - 1)) != 0) raise("createBuffer: alignment must be a power of two");
    auto b = std::make_shared<HostBuffer>();
    b->bytes.resize(size);
    if (data) data(b->bytes.data());
    return b;
}
std::shared_ptr<TiledOutput> HIPAccelerator::createTiledOutput(size_t size, size_t alignment) {
    if (!mDevice) raise("createTiledOutput: this accelerator has no device");
    if (size % 16 != 0 || alignment > 16) raise("createTiledOutput: expects an RGBW float buffer");
    return std::make_shared<HIPTiledOutput>(mDevice, size);
}
std::shared_ptr<ResourceLookUpTable> HIPAccelerator::createResourceLUT(std::vector<std::shared_ptr<Resource>> resources) {
    for (const auto& r : resources)
        if (!r) raise("createResourceLUT: null resource");
    return std::make_shared<ResourceLookUpTable>(std::move(resources));
}
const std::vector<Context*>& HIPAccelerator::enumerateContexts() const { return mContexts; }

// ---- HIPTracer ----
HIPTracer::HIPTracer(std::shared_ptr<HIPAccelerator> accelerator) : mAccelerator(std::move(accelerator)) {
    if (!mAccelerator) raise("HIPTracer needs a HIPAccelerator");
}

std::shared_ptr<AccelerationStructure> HIPTracer::buildAcceleration(const GeometryDesc& desc) {
    const TriangleIndexedGeometryDesc& g = desc.desc;
    auto* buf = dynamic_cast<const HostBuffer*>(g.buffer.get());
    if (!buf) raise("buildAcceleration: the buffer was not created by HIPAccelerator");
    const size_t vEnd = g.vertices + static_cast<size_t>(g.vertCount) * 12;
    const size_t iEnd = g.index + static_cast<size_t>(g.triCount) * 12;
    if (vEnd > g.bufferSize || iEnd > g.bufferSize || g.bufferSize > buf->bytes.size())
        raise("buildAcceleration: vertex or index range outside the buffer");
    auto mesh = std::make_shared<MeshAccel>();
    mesh->verts.resize(static_cast<size_t>(g.vertCount) * 3);
    mesh->index.resize(static_cast<size_t>(g.triCount) * 3);
    std::memcpy(mesh->verts.data(), buf->bytes.data() + g.vertices, static_cast<size_t>(g.vertCount) * 12);
    std::memcpy(mesh->index.data(), buf->bytes.data() + g.index, static_cast<size_t>(g.triCount) * 12);
    for (uint32_t i : mesh->index)
        if (i >= g.vertCount) raise("buildAcceleration: triangle index out of range");
    if (g.texCoords != SIZE_MAX) {
        if (g.texCoords + static_cast<size_t>(g.vertCount) * 8 > g.bufferSize)
            raise("buildAcceleration: texture coordinates outside the buffer");
        mesh->uv.resize(static_cast<size_t>(g.vertCount) * 2);
        std::memcpy(mesh->uv.data(), buf->bytes.data() + g.texCoords, static_cast<size_t>(g.vertCount) * 8);
    }
    if (desc.transform) {
        for (size_t v = 0; v < mesh->verts.size(); v += 3) {
            const auto p = applyPoint(*desc.transform, &mesh->verts[v]);
            std::copy(p.begin(), p.end(), mesh->verts.begin() + static_cast<long>(v));
        }
    }
    return mesh;
}

std::shared_ptr<GSMInstance> HIPTracer::buildGSMInstance(std::shared_ptr<Geometry> geometry,
                                                         std::shared_ptr<Surface> surface,
                                                         std::shared_ptr<Medium> medium) {
    if (!geometry || !surface) raise("buildGSMInstance: geometry and surface are required");
    if (medium) raise("buildGSMInstance: participating media are not supported by the HIP tracer");
    auto inst = std::make_shared<HIPGSMInstance>();
    inst->mesh = std::static_pointer_cast<MeshAccel>(buildAcceleration(geometry->desc()));
    inst->surface = std::move(surface);
    return inst;
}

std::shared_ptr<Node> HIPTracer::buildNode(const std::shared_ptr<Object>& object) {
    if (!dynamic_cast<const HIPGSMInstance*>(object.get()) && !dynamic_cast<const Light*>(object.get()) &&
        !dynamic_cast<const Sensor*>(object.get()))
        raise("buildNode: expects a GSMInstance, a Light or a Sensor");
    auto n = std::make_shared<LeafNode>();
    n->object = object;
    return n;
}

// EmbreeBranchNode (Embree.cpp:784-808): a static child transform is one key of index 0 with
// step <= 0; a keyframed one (step > 0) moves its subtree as one instance at each ray's time.
std::shared_ptr<Node> HIPTracer::buildNode(const std::vector<std::pair<TransformInfo, std::shared_ptr<Node>>>& children) {
    auto g = std::make_shared<GroupNode>();
    for (const auto& [info, child] : children) {
        if (!child) raise("buildNode: null child");
        if (info.step > 0.0f) {
            if (info.transforms.empty()) raise("buildNode: a keyframed transform without keys");
            for (size_t k = 1; k < info.transforms.size(); ++k)
                if (info.transforms[k].first <= info.transforms[k - 1].first)
                    raise("buildNode: keyframe indices must increase");
            for (const auto& key : info.transforms) (void)fromSRT(key.second);  // rejects a zero quaternion
            g->children.push_back({identity34(), std::make_shared<const TransformInfo>(info), child});
            continue;
        }
        Mat34 m = identity34();
        if (!info.transforms.empty()) {
            if (info.transforms.size() != 1 || info.transforms.front().first != 0u) raise("Unrecognized transform.");
            if (!isIdentity(info.transforms.front().second)) m = fromSRT(info.transforms.front().second);
        }
        g->children.push_back({m, nullptr, child});
    }
    return g;
}

pg_motion motionForInterval(const TransformInfo& info, const std::array<float, 12>& parent, float begin, float end) {
    pg_motion out{};
    std::memcpy(out.parent, parent.data(), sizeof(out.parent));
    const auto& keys = info.transforms;
    if (keys.empty()) raise("updateTimeInterval: a keyframed transform without keys");
    auto timeOf = [&](const std::pair<uint32_t, TransformSRT>& k) { return info.offset + info.step * static_cast<float>(k.first); };
    auto put = [&](uint32_t slot, const TransformSRT& t) { std::memcpy(out.keys[slot], &t, sizeof(t)); };
    out.num_keys = 1;
    put(0, keys.front().second);
    if (timeOf(keys.back()) < begin || timeOf(keys.front()) > end) {  // rtcDisableGeometry (:718-724)
        out.visible = 0;
        return out;
    }
    out.visible = 1;
    // both ends looked up with `begin` (:726-734): the last key at or before it and the first after it
    auto after = std::upper_bound(keys.begin(), keys.end(), begin,
                                  [&](float t, const std::pair<uint32_t, TransformSRT>& k) { return t < timeOf(k); });
    auto first = after, last = after;
    if (first != keys.begin()) --first;
    if (last == keys.end()) --last;
    uint32_t step = 0;  // gcd of the index steps from first to last (:742-749)
    for (auto it = first; it != last; ++it) {
        uint32_t a = step, b = std::next(it)->first - it->first;
        while (b) a %= b, std::swap(a, b);
        step = a;
    }
    if (step == 0 || !(end > begin))
        raise("updateTimeInterval: the frame [" + std::to_string(begin) + ", " + std::to_string(end) +
              "] opens outside the keyframes' time span (before the first key or at or after the last) or is empty; "
              "the reference divides by zero there");
    const uint32_t count = (last->first - first->first) / step + 1;
    if (count > PG_MAX_MOTION_KEYS) raise("updateTimeInterval: more than 8 time steps in one frame");
    out.num_keys = count;
    out.t0 = (timeOf(*first) - begin) / (end - begin);
    out.t1 = (timeOf(*last) - begin) / (end - begin);
    put(0, first->second);
    auto cur = first, nxt = std::next(first);
    for (uint32_t i = 1; i < count; ++i) {  // in-between steps: lerp of the neighbouring keys (:762-771)
        const uint32_t key = first->first + i * step;
        if (key == nxt->first) {
            put(i, nxt->second);
            cur = nxt, ++nxt;
        } else {
            const float u = static_cast<float>(key - cur->first) / static_cast<float>(nxt->first - cur->first);
            const float* x = reinterpret_cast<const float*>(&cur->second);
            const float* y = reinterpret_cast<const float*>(&nxt->second);
            for (int c = 0; c < 16; ++c) out.keys[i][c] = x[c] * (1.0f - u) + y[c] * u;
        }
    }
    return out;
}

std::shared_ptr<RTProgram> HIPTracer::buildProgram(const LinkableProgram& linkable, std::string symbol) {
    const auto formats = mAccelerator->getSupportedLinkableFormat();
    if (std::find(formats.begin(), formats.end(), linkable.format) == formats.end())
        raise("buildProgram: unsupported linkable format " + linkable.format);
    const auto table = parseNativeLinkable(linkable);
    if (!table.count(symbol)) raise("buildProgram: the program does not provide " + symbol);
    const auto& known = hipBuiltinSymbols();
    if (std::find(known.begin(), known.end(), symbol) == known.end())
        raise("buildProgram: " + symbol + " has no native HIP implementation");
    auto p = std::make_shared<HIPProgram>();
    p->symbol = std::move(symbol);
    return p;
}

std::shared_ptr<Texture> HIPTracer::generateTexture(const Config& desc, uint32_t channel) const {
    if (!desc.isObject()) raise("texture: expected a texture object" + desc.where());
    std::shared_ptr<Texture> res;
    auto channels = [&](const Config& values) {
        const auto& a = values.array();
        if (a.size() != 1 && a.size() != 2 && a.size() != 4)
            raise("Unsupported channel " + std::to_string(a.size()) + desc.where());
        return static_cast<uint32_t>(a.size());
    };
    if (const Config* id = desc.find("ClassID")) {
        const std::string& cls = id->string();
        if (cls == "Piper.BuiltinComponent.ProceduralTexture.ConstantTexture") {  // ProceduralTexture.cpp:26-57
            ConstantData d{};
            const Config& v = desc.at("Value");
            d.channel = channels(v);
            for (uint32_t k = 0; k < d.channel; ++k) d.value[k] = static_cast<float>(v.array()[k].number());
            res = std::make_shared<ConstantTexture>(d);
        } else if (cls == "Piper.BuiltinComponent.ProceduralTexture.CheckBoard") {  // ProceduralTexture.cpp:59-96
            CheckBoardData d{};
            d.scale = static_cast<float>(desc.at("Scale").number());
            const Config &b = desc.at("Black"), &w = desc.at("White");
            d.channel = channels(b);
            if (w.array().size() != b.array().size()) raise("Channel of black and white must be identical." + desc.where());
            for (uint32_t k = 0; k < d.channel; ++k)
                d.black[k] = static_cast<float>(b.array()[k].number()), d.white[k] = static_cast<float>(w.array()[k].number());
            res = std::make_shared<CheckBoardTexture>(d);
        } else {
            raise("unsupported texture class " + cls + desc.where());
        }
    } else {  // a sampled image (Embree.cpp:1340-1350)
        const std::string& wrap = desc.at("WrapMode").string();
        if (wrap != "Repeat" && wrap != "Mirror") raise("Unrecognized wrap mode " + wrap + desc.where());
        const Config& img = desc.at("Image");
        if (img.at("ClassID").string() != "Piper.BuiltinComponent.SampledTexture.Image")
            raise("Image: expected Piper.BuiltinComponent.SampledTexture.Image" + img.where());
        std::filesystem::path path = img.at("Path").string();
        if (path.is_relative() && !mResourceRoot.empty()) path = std::filesystem::path(mResourceRoot) / path;
        PiperFrontend::DecodedImage decoded;
        try {
            decoded = PiperFrontend::loadImage(path.string());
        } catch (const std::runtime_error& e) {
            raise(e.what());
        }
        res = std::make_shared<ImageTexture>(std::move(decoded), wrap == "Repeat" ? TextureWrap::Repeat : TextureWrap::Mirror);
    }
    if (res->channel() != channel)  // Embree.cpp:1423-1432
        raise("Mismatched channel. Expect " + std::to_string(channel) + ", but get " + std::to_string(res->channel()) +
              desc.where());
    return res;
}

std::unique_ptr<Pipeline> HIPTracer::buildPipeline(const std::shared_ptr<Node>& scene, Integrator& integrator,
                                                   RenderDriver& renderDriver, LightSampler& lightSampler,
                                                   std::shared_ptr<Sampler> sampler) {
    if (!scene || !sampler) raise("buildPipeline: scene and sampler are required");
    if (!mAccelerator->device()) raise("buildPipeline: the accelerator has no device");
    Registry reg;
    const MaterializeContext ctx = reg.context(*this);
    DecodedScene decoded = decodeScene(*this, reg, scene);
    pg_render_params base{};
    {  // the render driver and its filter (RDData -> the filter's call, FixedSampler.cpp:135-143)
        const RenderDriverProgram rd = renderDriver.materialize(ctx);
        expectSymbol(rd.accumulate, "accumulate", "render driver");
        const auto d = unpack<RDData>(rd.payload, "RDData");
        const auto& f = reg.call(d.filter.handle, "RDData filter");
        const auto g = unpack<GaussianFilterData>(f.payload, "GaussianFilterData");
        if (f.symbol == "box") base.filter = PG_FILTER_BOX;
        else if (f.symbol == "triangle") base.filter = PG_FILTER_TRIANGLE;
        else if (f.symbol == "gaussian") base.filter = PG_FILTER_GAUSSIAN, base.filter_alpha = -g.negAlpha;
        else raise("buildPipeline: " + f.symbol + " is not a filter program");
        if (base.filter == PG_FILTER_GAUSSIAN && !(base.filter_alpha > 0.0f))
            raise("buildPipeline: gaussian filter needs Alpha > 0");
    }
    {  // lights are selected uniformly; the sampler's payload flags the delta lights (BasicLight.cpp:97-111)
        lightSampler.preprocess(decoded.lights.empty() ? std::vector<std::shared_ptr<Light>>{Light::environment({0, 0, 0})}
                                                       : decoded.lights);
        const LightSamplerProgram ls = lightSampler.materialize(ctx);
        expectSymbol(ls.select, "uniformSelect", "light sampler");
        uint32_t n = 0;
        if (ls.payload.size() < 4) raise("light sampler: payload too short");
        std::memcpy(&n, ls.payload.data(), 4);
        if (ls.payload.size() != 4 + static_cast<size_t>(n)) raise("light sampler: payload size differs from its count");
        for (uint32_t i = 0; i < n && !decoded.lights.empty(); ++i)
            if ((ls.payload[4 + i] != std::byte{0}) != decoded.lights[i]->isDelta())
                raise("light sampler: delta flags differ from the scene's lights");
    }
    {  // the integrator (Data, PathTracing/Shared.hpp:22-28)
        const IntegratorProgram ip = integrator.materialize(ctx);
        expectSymbol(ip.trace, "trace", "integrator");
        const auto d = unpack<PathTracingData>(ip.payload, "PathTracing Data");
        base.max_depth = d.maxDepth, base.min_rr = d.minDepthForRussianRoulette;
        base.flags = integrator.compatFlags | PG_NO_STAGE_TIMING;
    }
    {
        const SamplerProgram sp = sampler->materialize(ctx);
        expectSymbol(sp.start, "sobolStart", "sampler");
        expectSymbol(sp.generate, "sobolGenerate", "sampler");
    }
    return std::make_unique<HIPPipeline>(*mAccelerator, std::move(decoded), base, std::move(sampler));
}

SceneExport exportScene(Tracer& tracer, const std::shared_ptr<Node>& scene) {
    if (!scene) raise("exportScene: null scene");
    Registry reg;
    DecodedScene d = decodeScene(tracer, reg, scene);
    SceneExport e = std::move(d.e);
    for (size_t i = 0; i < e.textures.size(); ++i)  // texels moved with their vectors
        e.textures[i].texels = e.textureTexels[i].empty() ? nullptr : e.textureTexels[i].data();
    return e;
}

size_t HIPTracer::getAlignmentRequirement(AlignmentRequirement requirement) const noexcept {
    (void)requirement;
    return 16;  // Embree.cpp:1437-1448 asks for 16 as well
}
Accelerator& HIPTracer::getAccelerator() const noexcept { return *mAccelerator; }

// ---- FixedSampler ----
std::vector<std::array<float, 3>> FixedSampler::renderFrame(Tracer& tracer, TraceLauncher& launcher) {
    const auto [width, height] = launcher.getFilmResolution();
    auto buffer = tracer.getAccelerator().createTiledOutput(static_cast<size_t>(width) * height * 16, 16);
    const RenderRECT rect = launcher.getRenderRECT();
    constexpr uint32_t tileSize = 8;
    const uint32_t blockX = (rect.width + tileSize - 1) / tileSize;
    const uint32_t blockY = (rect.height + tileSize - 1) / tileSize;
    const std::function<SBTPayload(uint32_t)> launchData = [w = width, h = height](uint32_t) {
        return packSBTPayload(LaunchData{0, w, h});
    };
    const std::vector<std::shared_ptr<Resource>> resources{buffer};
    std::vector<FutureVoid> tiles;
    for (uint32_t bx = 0; bx < blockX; ++bx)
        for (uint32_t by = 0; by < blockY; ++by) {
            const uint32_t left = rect.left + bx * tileSize, top = rect.top + by * tileSize;
            const RenderRECT tile{left, top, std::min(tileSize, rect.width - bx * tileSize),
                                  std::min(tileSize, rect.height - by * tileSize)};
            tiles.push_back(launcher.launch(tile, launchData, resources));
        }
    buffer->markDirty(FutureVoid([tiles] {
        for (const auto& t : tiles) t.getSync();
        return Unit{};
    }));
    for (const auto& t : tiles) t.ready();
    const std::vector<std::byte> bytes = buffer->download().getSync();
    lastFilm.resize(bytes.size() / 4);
    std::memcpy(lastFilm.data(), bytes.data(), bytes.size());
    std::vector<std::array<float, 3>> res(static_cast<size_t>(width) * height);
    for (size_t i = 0; i < res.size(); ++i) {
        const float* p = &lastFilm[4 * i];
        if (p[3] != 0.0f) res[i] = {p[0] / p[3], p[1] / p[3], p[2] / p[3]};
        else res[i] = {0.0f, 0.0f, 0.0f};
    }
    return res;
}

// ---- module ----
namespace {
class HIPModule final : public Module {
public:
    explicit HIPModule(std::string path) : mPath(std::move(path)) {}
    // Module::newInstance (Module.hpp:26-32; the Parallel and Embree modules answer "Accelerator"
    // and "Tracer", Parallel.cpp:526-535, Embree.cpp:1466-1474)
    [[nodiscard]] std::shared_ptr<Object> newInstance(std::string_view classID, const Config& config) override {
        if (classID == "Accelerator") return accelerator(config);
        if (classID == "Tracer") {
            const Config* acc = config.isObject() ? config.find("Accelerator") : nullptr;
            return std::make_shared<HIPTracer>(accelerator(acc ? *acc : Config::makeObject({})));
        }
        raise("module " + mPath + " has no class " + std::string(classID));
    }

private:
    static std::shared_ptr<HIPAccelerator> accelerator(const Config& config) {
        if (config.isObject()) {
            if (const Config* h = config.find("HostOnly"); h && h->boolean()) return HIPAccelerator::hostOnly();
            if (const Config* d = config.find("Device")) return std::make_shared<HIPAccelerator>(static_cast<int>(d->number()));
        }
        return std::make_shared<HIPAccelerator>(0);
    }
    std::string mPath;
};
}  // namespace

}  // namespace Piper

extern "C" Piper::Module* piperInitModule(Piper::PiperContext& context, Piper::Allocator& allocator, const char* path) {
    (void)context;
    void* mem = allocator.alloc(sizeof(Piper::HIPModule), alignof(Piper::HIPModule));
    if (!mem) throw std::bad_alloc();
    try {
        return new (mem) Piper::HIPModule(path ? path : "");
    } catch (...) {
        allocator.free(mem);
        throw;
    }
}
extern "C" const char* piperGetProtocol() { return "GCC@STD@0"; }

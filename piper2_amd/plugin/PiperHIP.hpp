// The MI355X component pair behind Piper2's plugin interfaces:
//   HIPAccelerator : Accelerator (Interface/Infrastructure/Accelerator.hpp:130-161)
//   HIPTracer      : Tracer      (Interface/BuiltinComponent/Tracer.hpp:140-159)
// with the Pipeline and TraceLauncher (Tracer.hpp:96-116) it returns, the module entry points
// that load the pair (piperInitModule / piperGetProtocol, Interface/Infrastructure/Module.hpp:54-70),
// and the built-in components a scene is made of (surfaces, textures, lights, sensor, integrator,
// sampler, light sampler, the FixedSampler render driver and its filter), each of which
// materialises the reference's own SBT payload layouts through a MaterializeContext
// (Tracer.hpp:121-133) exactly as its reference counterpart does.
//
// Piper's own runtime (EASTL containers, SharedPtr, PiperContext, scheduler) is not part of this
// build, so the interfaces are restated with std types; each method keeps the reference's name,
// arguments and meaning. Configs are JSON values (the reference's Config is a JSON tree). Errors:
// the reference hands them to ErrorHandler::raiseException, which ends the program; here they
// throw PiperException carrying the message, so the host application decides.
//
// Everything below calls the C ABI of include/piper_hip.h; nothing here touches the device directly.
#pragma once
#include <array>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <optional>
#include <stdexcept>
#include <string>
#include <string_view>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/piper_hip.h"
#include "../frontend/Json.hpp"

namespace Piper {

class PiperException : public std::runtime_error {
public:
    explicit PiperException(const std::string& msg) : std::runtime_error(msg) {}
};

using Config = PiperFrontend::Json;  // Interface/Infrastructure/Config.hpp: a JSON tree

// ---- value types (restated) ----
struct Dim3 final {  // Kernel/DeviceRuntime.hpp:22-24
    uint32_t x, y, z;
};
struct RenderRECT final {  // Tracer.hpp:87-89
    uint32_t left, top, width, height;
};
struct SensorNDCAffineTransform final {  // Kernel/Protocol.hpp:113-115
    float ox, oy, sx, sy;
};
enum class FitMode : uint32_t { Fill, OverScan };  // Interface/BuiltinComponent/Sensor.hpp:36
// Sensor.hpp:38-74
std::pair<SensorNDCAffineTransform, RenderRECT> calcRenderRECT(uint32_t width, uint32_t height,
                                                                float deviceAspectRatio, FitMode fitMode);

// Kernel/Transform.hpp:390-393: local -> world = translate * rotate(quat) * (scale, skew, shift).
struct alignas(16) TransformSRT final {
    float scaleX, scaleY, scaleZ, skewXY, skewXZ, skewYZ, shiftX, shiftY, shiftZ, quatW, quatX, quatY, quatZ, transX,
        transY, transZ;
};
TransformSRT identitySRT();
// Tracer.hpp:80-84. Static: step <= 0 and at most one key, index 0 (EmbreeBranchNode,
// Embree.cpp:790-801). Keyframed (step > 0): key k sits at time offset + step * index.
struct TransformInfo final {
    float offset = 0.0f, step = 0.0f;
    std::vector<std::pair<uint32_t, TransformSRT>> transforms;  // local to world
};
// The transform of a keyframed instance for the frame [begin, end], as updateTransformSRT
// (Embree.cpp:709-773) hands it to Embree: the keys around `begin` (the reference looks up both
// ends of the key range with `begin`), their times mapped to the shutter as (t - begin) / (end -
// begin); invisible when the keys' span misses the frame. parent: the static transform of the
// Group the instance sits in. Raises where the reference divides by zero (a shutter that opens
// before the first key or at or after the last, or end == begin).
pg_motion motionForInterval(const TransformInfo& info, const std::array<float, 12>& parent, float begin, float end);

using SBTPayload = std::vector<std::byte>;  // Tracer.hpp:86-91
template <typename T>
SBTPayload packSBTPayload(const T& data) {
    static_assert(std::is_trivially_copyable_v<T>);
    SBTPayload p(sizeof(T));
    std::memcpy(p.data(), &data, sizeof(T));
    return p;
}

// Handles held by payloads (Kernel/Protocol.hpp:128-134): pointer-sized, so every payload below has
// the reference's 64-bit layout. A pipeline gives them meaning: a CallHandle names a registered
// call, a TraversalHandle a scene node's transform.
struct TraversalHandleReserved;
using TraversalHandle = const TraversalHandleReserved*;
struct CallHandleReserved;
using CallHandle = const CallHandleReserved*;
struct StatisticsHandleReserved;
using StatisticsHandle = const StatisticsHandleReserved*;
template <typename Protocol>
struct Call final {  // Kernel/Protocol.hpp:307-316
    CallHandle handle;
};
struct TextureSampleFunc;
struct GeometrySampleFunc;
struct FilterFunc;
enum class TextureWrap : uint32_t { Repeat, Mirror };  // Kernel/Protocol.hpp:227

// ---- SBT payloads, byte for byte as the reference components pack them ----
struct Vec3 final {
    float x, y, z;
};
struct PCData final {  // Sensor/PerspectiveCamera/Shared.hpp:22-29
    TraversalHandle traversal;
    Vec3 lookAt;
    float size[2];
    Vec3 upRef;
    float focalLength;
    float apertureRadius;
};
struct SobolData final {  // Sampler/ScrambledSobolSampler/Shared.hpp:24-27
    uint32_t resolution;
    uint32_t log2Resolution;
};
struct PathTracingData final {  // Integrator/PathTracing/Shared.hpp:22-28 (Data)
    uint32_t maxDepth;
    uint32_t minDepthForRussianRoulette;
    StatisticsHandle profileDepth, profilePathTime, profileValidRay;
};
struct MatteData final {  // Surface/Classic/Shared.hpp:51-54
    Call<TextureSampleFunc> diffuseTexture, roughnessTexture;
};
struct PlasticData final {  // Classic/Shared.hpp:55-59
    Call<TextureSampleFunc> diffuseTexture, specularTexture, roughnessTexture;
};
struct MirrorData final {  // Classic/Shared.hpp:60-62
    Call<TextureSampleFunc> reflectionTexture;
};
struct GlassData final {  // Classic/Shared.hpp:63-68
    Call<TextureSampleFunc> reflection, transmission, roughnessX, roughnessY;
};
struct SubstrateData final {  // Classic/Shared.hpp:69-74
    Call<TextureSampleFunc> diffuse, specular, roughnessX, roughnessY;
};
// The GGX conductor (Schlick F0) that this path adds; the reference has no conductor.
// roughnessIsAlpha: 1 when the roughness input already holds GGX alpha, 0 for alpha = roughness^2.
struct ConductorData final {
    Call<TextureSampleFunc> f0Texture, roughnessTexture;
    uint32_t roughnessIsAlpha;
};
struct ConstantData final {  // Texture/ProceduralTexture/Shared.hpp:22-25
    uint32_t channel;
    float value[4];
};
struct CheckBoardData final {  // ProceduralTexture/Shared.hpp:26-30
    float scale;
    uint32_t channel;
    float black[4], white[4];
};
struct SampledTextureData final {  // Texture/SampledTexture/Shared.hpp:22-30 (Data)
    TextureWrap wrap;
    uint32_t width, height, stride, channel;
    uint32_t texel;  // resource index of the texels
    StatisticsHandle profileSample;
};
struct DiffuseAreaLightData final {  // Light/AreaLight/Shared.hpp:22-26
    Vec3 radiance;
    Call<GeometrySampleFunc> sample;
    float area;
};
struct PerPlaneData final {  // Geometry/BasicShapes/Shared.hpp:22-29
    Vec3 origin, u, v, normal, tangent;
    uint32_t maxDetComp;
};
struct PlaneData final {  // BasicShapes/Shared.hpp:30-32
    uint32_t primitives;  // resource index of the PerPlaneData array
};
struct CDFData final {  // BasicShapes/Shared.hpp:33-40
    uint32_t primitives;  // resource index of the PerPlaneData array
    TraversalHandle traversal;
    uint32_t cdf, pdf;    // resource indices of the area CDF and PDF (floats)
    float inverseArea;
    uint32_t size;
};
struct SampledEnvironmentData final {  // Light/BasicLight/Shared.hpp:22-25
    TraversalHandle traversal;
    Vec3 texture;  // constant radiance
};
struct PointLightData final {  // BasicLight/Shared.hpp:26-29
    TraversalHandle traversal;
    Vec3 intensity;
};
struct RDData final {  // RenderDriver/FixedSampler/Shared.hpp:22-24
    Call<FilterFunc> filter;
};
struct LaunchData final {  // FixedSampler/Shared.hpp:25-28
    uint32_t rgbw;  // index of the film among the launch's resources
    uint32_t w, h;
};
struct GaussianFilterData final {  // FixedSampler/Shared.hpp:29-32
    float negAlpha;
    float sub;
};
static_assert(sizeof(PCData) == 48 && sizeof(PathTracingData) == 32 && sizeof(MatteData) == 16 &&
              sizeof(GlassData) == 32 && sizeof(ConstantData) == 20 && sizeof(CheckBoardData) == 40 &&
              sizeof(SampledTextureData) == 32 && sizeof(DiffuseAreaLightData) == 32 && sizeof(PerPlaneData) == 64 &&
              sizeof(CDFData) == 32 && sizeof(SampledEnvironmentData) == 24 && sizeof(PointLightData) == 24 &&
              sizeof(RDData) == 8 && sizeof(LaunchData) == 12,
              "payload layouts follow the reference's 64-bit structs");

// ---- Future (Interface/Infrastructure/Concurrency.hpp): ready() polls, getSync() waits ----
// The work behind a future runs when it is first polled or awaited.
template <typename T>
class Future final {
    struct State {
        std::function<T()> run;
        std::optional<T> value;
    };
    std::shared_ptr<State> mState;

public:
    Future() = default;
    explicit Future(std::function<T()> run) : mState(std::make_shared<State>()) { mState->run = std::move(run); }
    bool valid() const noexcept { return mState != nullptr; }
    bool ready() const {
        if (!mState->value) mState->value = mState->run();
        return true;
    }
    const T& getSync() const {
        ready();
        return *mState->value;
    }
};
struct Unit final {};
using FutureVoid = Future<Unit>;

class Object {
public:
    virtual ~Object() = default;
};

// ---- the accelerator (Accelerator.hpp:30-161) ----
struct ContextHandleReserved;
using ContextHandle = ContextHandleReserved*;  // manager of resources and command queues
struct CommandQueueHandleReserved;
using CommandQueueHandle = CommandQueueHandleReserved*;
class Context : public Object {  // Accelerator.hpp:37-43
public:
    [[nodiscard]] virtual std::unique_lock<std::recursive_mutex> makeCurrent() = 0;
    [[nodiscard]] virtual ContextHandle getHandle() const noexcept = 0;
    [[nodiscard]] virtual CommandQueueHandle select() = 0;
};
class Resource : public Object {};
class TiledOutput : public Resource {  // Accelerator.hpp:80-87
public:
    [[nodiscard]] virtual Future<std::vector<std::byte>> download() const = 0;
    virtual void markDirty(FutureVoid future) = 0;
};
class Kernel : public Resource {};
class ResourceLookUpTable : public Resource {
public:
    explicit ResourceLookUpTable(std::vector<std::shared_ptr<Resource>> r) : resources(std::move(r)) {}
    std::vector<std::shared_ptr<Resource>> resources;
};
// Trivially copyable kernel arguments packed back to back with their offsets (Accelerator.hpp:96-123).
class ArgumentPackage final {
    void append() {}
    template <typename T, typename... Args>
    void append(const T& first, const Args&... args) {
        static_assert(std::is_trivially_copyable_v<T>);
        offset.push_back({static_cast<uint32_t>(data.size()), static_cast<uint32_t>(sizeof(T))});
        const auto* b = reinterpret_cast<const std::byte*>(&first);
        data.insert(data.end(), b, b + sizeof(T));
        append(args...);
    }

public:
    std::vector<std::byte> data;
    std::vector<std::pair<uint32_t, uint32_t>> offset;
    template <typename... Args>
    explicit ArgumentPackage(const Args&... args) {
        append(args...);
    }
    // argument `index` as T (its size must match)
    template <typename T>
    T get(uint32_t index) const {
        if (index >= offset.size() || offset[index].second != sizeof(T))
            throw PiperException("kernel argument " + std::to_string(index) + " is missing or of another size");
        T v;
        std::memcpy(&v, data.data() + offset[index].first, sizeof(T));
        return v;
    }
};

// A program the tracer and the accelerator link (Tracer.hpp:38-45 LinkableProgram): `exchange`
// holds the program in `format`. This accelerator links the "Native" format only: the bytes
// "Native", then per symbol its name, a NUL and a 64-bit address (Parallel.cpp:219-241 parses the
// same layout).
struct LinkableProgram final {
    std::vector<std::byte> exchange;
    std::string format;
    uint64_t UID = 0;
};
LinkableProgram makeNativeLinkable(const std::vector<std::pair<std::string, const void*>>& symbols, uint64_t uid);
std::unordered_map<std::string, const void*> parseNativeLinkable(const LinkableProgram& program);

// What a native kernel of this accelerator receives at launch: it runs on the host and enqueues
// device work through the C ABI on the context's device.
struct NativeLaunch final {
    Dim3 size;
    Context& context;
    const std::vector<std::shared_ptr<Resource>>& resources;  // the launch's resource look-up table
    const ArgumentPackage& args;
    const std::vector<const void*>& dynamicSymbols;  // addresses of compileKernel's dynamicSymbols
};
using NativeKernelFunc = void (*)(const NativeLaunch&);

class Accelerator : public Object {  // Accelerator.hpp:130-161
    [[nodiscard]] virtual FutureVoid launchKernelImpl(const Dim3& size, std::shared_ptr<Kernel> kernel,
                                                      std::shared_ptr<ResourceLookUpTable> root, ArgumentPackage args) = 0;

public:
    [[nodiscard]] virtual const char* getNativePlatform() const noexcept = 0;
    [[nodiscard]] virtual std::vector<std::string> getSupportedLinkableFormat() const = 0;
    [[nodiscard]] virtual std::shared_ptr<Kernel> compileKernel(const std::vector<LinkableProgram>& linkable,
                                                                std::unordered_map<std::string, std::string> staticRedirectedSymbols,
                                                                std::vector<std::string> dynamicSymbols,
                                                                std::string entryFunction) = 0;
    template <typename... Args>
    [[nodiscard]] FutureVoid launchKernel(const Dim3& size, std::shared_ptr<Kernel> kernel,
                                          std::shared_ptr<ResourceLookUpTable> root, const Args&... args) {
        return launchKernelImpl(size, std::move(kernel), std::move(root), ArgumentPackage(args...));
    }
    [[nodiscard]] virtual std::shared_ptr<Resource> createBuffer(size_t size, size_t alignment,
                                                                 std::function<void(std::byte*)> data) = 0;
    [[nodiscard]] virtual std::shared_ptr<TiledOutput> createTiledOutput(size_t size, size_t alignment) = 0;
    [[nodiscard]] virtual std::shared_ptr<ResourceLookUpTable> createResourceLUT(
        std::vector<std::shared_ptr<Resource>> resources) = 0;
    [[nodiscard]] virtual const std::vector<Context*>& enumerateContexts() const = 0;
};

// ---- scene objects and programs (Tracer.hpp:33-84, Interface/BuiltinComponent/*.hpp) ----
class AccelerationStructure : public Object {};
class Node : public Object {};
class RTProgram : public Object {};
class GSMInstance : public Object {};
class Tracer;

struct TriangleIndexedGeometryDesc final {  // Tracer.hpp:51-63
    uint32_t vertCount, triCount;
    std::shared_ptr<Resource> buffer;
    size_t bufferSize;
    size_t vertices;  // byte offset of float3 vertices
    size_t index;     // byte offset of uint3 indices
    size_t texCoords = SIZE_MAX;  // optional byte offset of float2 texture coordinates (one per vertex)
    size_t normal = SIZE_MAX, tangent = SIZE_MAX;  // optional, unused by the path
};
struct GeometryDesc final {  // Tracer.hpp:70-73; transform: row-major 3x4 local -> world
    std::optional<std::array<float, 12>> transform;
    TriangleIndexedGeometryDesc desc;
};

using CallSiteRegister = std::function<CallHandle(const std::shared_ptr<RTProgram>&, const SBTPayload&)>;
using TextureLoader = std::function<CallHandle(const Config&, uint32_t)>;
using ResourceRegister = std::function<uint32_t(std::shared_ptr<Resource>)>;
struct MaterializeContext final {  // Tracer.hpp:121-133 (cache manager and profiler are not restated)
    Tracer& tracer;
    Accelerator& accelerator;
    const ResourceRegister registerResource;
    const CallSiteRegister registerCall;
    const TextureLoader loadTexture;
};

struct SurfaceProgram final {  // Interface/BuiltinComponent/Surface.hpp:22-28
    SBTPayload payload;
    std::shared_ptr<RTProgram> init, sample, evaluate, pdf;
};
struct TextureProgram final {  // Texture.hpp:22-25
    SBTPayload payload;
    std::shared_ptr<RTProgram> sample;
};
struct SampledGeometryProgram final {  // Geometry.hpp:29-32
    SBTPayload payload;
    std::shared_ptr<RTProgram> sample;
};
struct LightProgram final {  // Light.hpp:34-40
    std::shared_ptr<RTProgram> init, sample, evaluate, pdf;
    SBTPayload payload;
};
struct LightSamplerProgram final {  // Light.hpp:22-25
    std::shared_ptr<RTProgram> select;
    SBTPayload payload;
};
struct SensorProgram final {  // Sensor.hpp:22-25
    SBTPayload payload;
    std::shared_ptr<RTProgram> rayGen;
};
struct IntegratorProgram final {  // Integrator.hpp:22-25
    SBTPayload payload;
    std::shared_ptr<RTProgram> trace;
};
struct SamplerProgram final {  // Sampler.hpp:22-25
    std::shared_ptr<RTProgram> start, generate;
};
struct SamplerAttributes final {  // Sampler.hpp:27-31
    SBTPayload payload;
    uint32_t maxDimension;
    uint32_t samplesPerPixel;
};
struct FilterProgram final {  // RenderDriver.hpp:23-26
    SBTPayload payload;
    std::shared_ptr<RTProgram> weight;
};
struct RenderDriverProgram final {  // RenderDriver.hpp:34-37
    SBTPayload payload;
    std::shared_ptr<RTProgram> accumulate;
};

// The built-in programs every component of this build links against: the HIP accelerator's
// "Native" table of the per-sample and per-bounce functions its kernels implement.
const LinkableProgram& hipBuiltinProgram();
const std::vector<std::string>& hipBuiltinSymbols();

// ---- components ----
// A triangle mesh (Geometry/TriangleMesh): the tracer builds its acceleration structure.
class Geometry : public Object {
public:
    explicit Geometry(GeometryDesc desc) : mDesc(std::move(desc)) {}
    const GeometryDesc& desc() const noexcept { return mDesc; }

private:
    GeometryDesc mDesc;
};
// Texture (Texture.hpp:29-35): made by Tracer::generateTexture from a config.
class Texture : public Object {
public:
    [[nodiscard]] virtual uint32_t channel() const noexcept = 0;
    [[nodiscard]] virtual TextureProgram materialize(const MaterializeContext& ctx) const = 0;
};
// Texture configs (the reference's own JSON): ConstantTexture Value, CheckBoard Black/White/Scale.
Config constantTextureConfig(const std::vector<float>& value);
Config colourConfig(std::array<float, 3> rgb);  // ConstantTexture (r, g, b, 1)
Config scalarConfig(float v);                   // ConstantTexture (v)
Config checkerboardConfig(const std::vector<float>& black, const std::vector<float>& white, float scale);

// Surface/Classic: each input is a texture config, loaded with the channel count the class asks
// for (Classic.cpp:79-200): colours 4, roughness 1.
class Surface : public Object {
public:
    enum class Kind { Matte, Conductor, Plastic, Mirror, Glass, Substrate, BlackBody };
    // Inputs in payload order. Matte: Diffuse, Roughness (0 Lambert, > 0 Oren-Nayar).
    // Conductor (GGX, Schlick F0; added by this build): F0, Roughness (alpha = roughness^2).
    // Plastic: Diffuse, Specular, Roughness. Mirror: Reflection. Glass: Reflection, Transmission,
    // RoughnessX, RoughnessY. Substrate: Diffuse, Specular, RoughnessX, RoughnessY. BlackBody: none.
    Surface(Kind kind, std::vector<Config> inputs);
    static std::shared_ptr<Surface> matte(std::array<float, 3> reflectance, float roughness = 0.0f);
    static std::shared_ptr<Surface> conductor(std::array<float, 3> f0, float roughness);
    // a conductor whose roughness input already holds GGX alpha
    static std::shared_ptr<Surface> conductorAlpha(std::array<float, 3> f0, float alpha);
    static std::shared_ptr<Surface> plastic(std::array<float, 3> diffuse, std::array<float, 3> specular, float roughness);
    static std::shared_ptr<Surface> mirror(std::array<float, 3> reflection);
    static std::shared_ptr<Surface> glass(std::array<float, 3> reflection, std::array<float, 3> transmission,
                                          float roughnessX = 0.0f, float roughnessY = 0.0f);
    static std::shared_ptr<Surface> substrate(std::array<float, 3> diffuse, std::array<float, 3> specular,
                                              float roughnessX, float roughnessY);
    [[nodiscard]] SurfaceProgram materialize(const MaterializeContext& ctx) const;
    Kind kind;
    std::vector<Config> inputs;
    bool roughnessIsAlpha = false;  // conductorAlpha
};
class Medium : public Object {};
// Lights (Light.hpp:49-58): AreaLight "Diffuse" over a BasicShapes Plane (one parallelogram per
// primitive), the environment (BasicLight SampledEnvironment with constant radiance) and the point
// light (BasicLight Point: Intensity, placed at its node's origin).
class Light : public Object {
public:
    enum class Kind { Area, Environment, Point };
    struct PlanePrimitive {
        std::array<float, 3> origin, u, v;
    };
    static std::shared_ptr<Light> areaPlane(std::array<float, 3> origin, std::array<float, 3> u,
                                            std::array<float, 3> v, std::array<float, 3> radiance);
    static std::shared_ptr<Light> area(std::vector<PlanePrimitive> primitives, std::array<float, 3> radiance);
    static std::shared_ptr<Light> environment(std::array<float, 3> radiance);
    static std::shared_ptr<Light> point(std::array<float, 3> intensity);
    [[nodiscard]] LightProgram materialize(TraversalHandle traversal, const MaterializeContext& ctx) const;
    [[nodiscard]] bool isDelta() const noexcept { return kind == Kind::Point; }
    Kind kind = Kind::Area;
    std::vector<PlanePrimitive> primitives;
    std::array<float, 3> radiance{};  // the point light's intensity
};
// Sensor/PerspectiveCamera (config PerspectiveCamera.cpp:35-46). The camera sits at its node's
// translation.
class Sensor : public Object {
public:
    std::array<float, 3> lookAt{}, upRef{0.0f, 1.0f, 0.0f};
    std::array<float, 2> size{0.036f, 0.024f};  // metres
    float focalLength = 0.035f;
    float apertureRadius = 0.0f;
    float getAspectRatio() const noexcept { return size[0] / size[1]; }
    [[nodiscard]] SensorProgram materialize(TraversalHandle traversal, const MaterializeContext& ctx) const;
};
// Integrator/PathTracing (PathTracing.cpp:30-56).
class Integrator : public Object {
public:
    uint32_t maxDepth = 8, minDepthForRussianRoulette = 3;
    uint32_t compatFlags = 0;  // PG_COMPAT_*: reproduce the reference's documented deviations (not a payload field)
    [[nodiscard]] IntegratorProgram materialize(const MaterializeContext& ctx) const;
};
// Light/BasicLight UniformLightSampler (BasicLight.cpp:82-112): preprocess records which lights
// are delta lights; the payload is {uint32 count, bool delta[count]}.
class LightSampler : public Object {
public:
    void preprocess(const std::vector<std::shared_ptr<Light>>& lights);
    [[nodiscard]] LightSamplerProgram materialize(const MaterializeContext& ctx) const;
    std::vector<uint8_t> deltaFlags;
};
// Sampler/ScrambledSobolSampler (ScrambledSobolSampler.cpp:37-75).
class Sampler : public Object {
public:
    uint32_t samplesPerPixel = 16;  // rounded up to a power of two
    [[nodiscard]] SamplerProgram materialize(const MaterializeContext& ctx) const;
    [[nodiscard]] SamplerAttributes generatePayload(uint32_t sampleWidth, uint32_t sampleHeight) const;
};
class TraceLauncher;
class RenderDriver : public Object {  // Interface/BuiltinComponent/RenderDriver.hpp:38-43
public:
    // RGB radiance per film pixel (rgb / w), row-major
    [[nodiscard]] virtual std::vector<std::array<float, 3>> renderFrame(Tracer& tracer, TraceLauncher& launcher) = 0;
    [[nodiscard]] virtual RenderDriverProgram materialize(const MaterializeContext& ctx) const = 0;
};

class TraceLauncher {  // Tracer.hpp:96-107
public:
    virtual ~TraceLauncher() = default;
    [[nodiscard]] virtual std::pair<uint32_t, uint32_t> getFilmResolution() const noexcept = 0;
    [[nodiscard]] virtual RenderRECT getRenderRECT() const noexcept = 0;
    virtual void updateTimeInterval(float begin, float end) noexcept = 0;
    [[nodiscard]] virtual FutureVoid launch(const RenderRECT& rect, const std::function<SBTPayload(uint32_t)>& launchData,
                                            const std::vector<std::shared_ptr<Resource>>& resources) = 0;
};

class Pipeline {  // Tracer.hpp:109-116
public:
    virtual ~Pipeline() = default;
    [[nodiscard]] virtual std::string generateStatisticsReport() const = 0;
    [[nodiscard]] virtual std::shared_ptr<TraceLauncher> prepare(const std::shared_ptr<Node>& sensor, uint32_t width,
                                                                 uint32_t height, FitMode fitMode) = 0;
};

enum class AlignmentRequirement { VertexBuffer, IndexBuffer, TextureCoordsBuffer, BoundsBuffer };  // Tracer.hpp:135

class Tracer : public Object {  // Tracer.hpp:140-159
public:
    [[nodiscard]] virtual std::shared_ptr<AccelerationStructure> buildAcceleration(const GeometryDesc& desc) = 0;
    [[nodiscard]] virtual std::shared_ptr<GSMInstance> buildGSMInstance(std::shared_ptr<Geometry> geometry,
                                                                        std::shared_ptr<Surface> surface,
                                                                        std::shared_ptr<Medium> medium) = 0;
    // leaf node of one object: a GSMInstance, a Light or a Sensor
    [[nodiscard]] virtual std::shared_ptr<Node> buildNode(const std::shared_ptr<Object>& object) = 0;
    [[nodiscard]] virtual std::shared_ptr<Node> buildNode(
        const std::vector<std::pair<TransformInfo, std::shared_ptr<Node>>>& children) = 0;
    [[nodiscard]] virtual std::shared_ptr<RTProgram> buildProgram(const LinkableProgram& linkable, std::string symbol) = 0;
    [[nodiscard]] virtual std::unique_ptr<Pipeline> buildPipeline(const std::shared_ptr<Node>& scene, Integrator& integrator,
                                                                  RenderDriver& renderDriver, LightSampler& lightSampler,
                                                                  std::shared_ptr<Sampler> sampler) = 0;
    [[nodiscard]] virtual std::shared_ptr<Texture> generateTexture(const Config& textureDesc, uint32_t channel) const = 0;
    [[nodiscard]] virtual size_t getAlignmentRequirement(AlignmentRequirement requirement) const noexcept = 0;
    [[nodiscard]] virtual Accelerator& getAccelerator() const noexcept = 0;
};

// ---- the MI355X pair ----
class HIPAccelerator final : public Accelerator {
    [[nodiscard]] FutureVoid launchKernelImpl(const Dim3& size, std::shared_ptr<Kernel> kernel,
                                              std::shared_ptr<ResourceLookUpTable> root, ArgumentPackage args) override;

public:
    explicit HIPAccelerator(int ordinal);  // pg_device_create
    // An accelerator without a device: scene graphs build (buffers are host memory) and export,
    // and native kernels launch on its one context (null handle), but pipelines and tiled outputs
    // raise.
    static std::shared_ptr<HIPAccelerator> hostOnly();
    ~HIPAccelerator() override;
    HIPAccelerator(const HIPAccelerator&) = delete;
    HIPAccelerator& operator=(const HIPAccelerator&) = delete;
    [[nodiscard]] const char* getNativePlatform() const noexcept override;
    [[nodiscard]] std::vector<std::string> getSupportedLinkableFormat() const override;  // {"Native"}
    // Links "Native" programs only (an "LLVM IR" module is rejected): the entry function, after
    // staticRedirectedSymbols, must be one of their symbols and names a NativeKernelFunc; every
    // dynamic symbol must resolve too.
    [[nodiscard]] std::shared_ptr<Kernel> compileKernel(const std::vector<LinkableProgram>& linkable,
                                                        std::unordered_map<std::string, std::string> staticRedirectedSymbols,
                                                        std::vector<std::string> dynamicSymbols,
                                                        std::string entryFunction) override;
    [[nodiscard]] std::shared_ptr<Resource> createBuffer(size_t size, size_t alignment,
                                                         std::function<void(std::byte*)> data) override;
    [[nodiscard]] std::shared_ptr<TiledOutput> createTiledOutput(size_t size, size_t alignment) override;
    [[nodiscard]] std::shared_ptr<ResourceLookUpTable> createResourceLUT(
        std::vector<std::shared_ptr<Resource>> resources) override;
    // one context per accelerator: the device handle (null for a host-only accelerator)
    [[nodiscard]] const std::vector<Context*>& enumerateContexts() const override;
    pg_device* device() const noexcept { return mDevice; }

private:
    HIPAccelerator() = default;
    pg_device* mDevice = nullptr;
    std::unique_ptr<Context> mContext;
    std::vector<Context*> mContexts;
};

class HIPTracer final : public Tracer {
public:
    explicit HIPTracer(std::shared_ptr<HIPAccelerator> accelerator);
    [[nodiscard]] std::shared_ptr<AccelerationStructure> buildAcceleration(const GeometryDesc& desc) override;
    [[nodiscard]] std::shared_ptr<GSMInstance> buildGSMInstance(std::shared_ptr<Geometry> geometry,
                                                                std::shared_ptr<Surface> surface,
                                                                std::shared_ptr<Medium> medium) override;
    [[nodiscard]] std::shared_ptr<Node> buildNode(const std::shared_ptr<Object>& object) override;
    [[nodiscard]] std::shared_ptr<Node> buildNode(
        const std::vector<std::pair<TransformInfo, std::shared_ptr<Node>>>& children) override;
    // Accepts a "Native" program that provides `symbol`, which must be one of the HIP built-ins.
    [[nodiscard]] std::shared_ptr<RTProgram> buildProgram(const LinkableProgram& linkable, std::string symbol) override;
    [[nodiscard]] std::unique_ptr<Pipeline> buildPipeline(const std::shared_ptr<Node>& scene, Integrator& integrator,
                                                          RenderDriver& renderDriver, LightSampler& lightSampler,
                                                          std::shared_ptr<Sampler> sampler) override;
    // Embree.cpp:1335-1351, 1423-1432: a ConstantTexture or CheckBoard (ProceduralTexture), or a
    // sampled image {WrapMode, Image: {Path}}; its channel count must equal `channel`. Relative
    // image paths resolve against setResourceRoot's folder.
    [[nodiscard]] std::shared_ptr<Texture> generateTexture(const Config& textureDesc, uint32_t channel) const override;
    [[nodiscard]] size_t getAlignmentRequirement(AlignmentRequirement requirement) const noexcept override;
    [[nodiscard]] Accelerator& getAccelerator() const noexcept override;
    void setResourceRoot(std::string folder) { mResourceRoot = std::move(folder); }

private:
    std::shared_ptr<HIPAccelerator> mAccelerator;
    std::string mResourceRoot;
};

// The FixedSampler's filter (SimpleFilter, RenderDriver/FixedSampler/FixedSampler.cpp:28-54):
// FilterType "box", "triangle" or "gaussian" (Alpha > 0), weights as FixedSampler/Kernel.cpp:55-70.
struct SimpleFilter {
    std::string filterType = "box";
    float alpha = 0.0f;
    [[nodiscard]] FilterProgram materialize(const MaterializeContext& ctx) const;
};

// RenderDriver/FixedSampler: 8x8 tiles, one launch each, then rgb / w per pixel
// (FixedSampler.cpp:68-143).
class FixedSampler final : public RenderDriver {
public:
    SimpleFilter filter;
    [[nodiscard]] std::vector<std::array<float, 3>> renderFrame(Tracer& tracer, TraceLauncher& launcher) override;
    [[nodiscard]] RenderDriverProgram materialize(const MaterializeContext& ctx) const override;
    std::vector<float> lastFilm;  // RGBW of the last frame, for inspection
};

// What buildPipeline uploads for a scene graph, decoded from the components' payloads:
// world-space triangles grouped by instance, one material per instance, the lights (environment,
// quads, points) and every sensor with its node's local-to-world transform (row-major 3x4) and
// its camera (PCData). No device is needed.
struct SceneExport {
    std::vector<float> verts;
    std::vector<uint32_t> index, triObject, objectMaterial;
    std::vector<pg_material> materials;
    std::vector<pg_quad> quads;
    std::vector<uint32_t> lightQuads;  // primitives (consecutive quads) of each area light
    std::vector<pg_point_light> points;
    std::array<float, 3> env{};
    std::vector<std::pair<const Node*, std::array<float, 12>>> sensors;
    std::vector<pg_camera> cameras;                      // per sensor, position from its transform
    std::vector<float> texcoords;                        // float2 per vertex, or empty
    std::vector<pg_texture> textures;                    // texels point into textureTexels
    std::vector<std::vector<uint8_t>> textureTexels;
    std::vector<pg_material_textures> materialTextures;  // per material (empty without textures)
    // per material: RoughnessY and its texture (glass, substrate; the others repeat their roughness)
    std::vector<float> roughnessY;
    std::vector<uint32_t> roughnessYTextures;
    [[nodiscard]] bool anisotropic() const {
        for (size_t i = 0; i < materials.size(); ++i)
            if (roughnessY[i] != materials[i].alpha || roughnessYTextures[i] != PG_TEX_NONE) return true;
        return false;
    }
    // moving instances: instance k's triangles start at instanceFirst[k] (static ones come first)
    struct Motion {
        std::shared_ptr<const TransformInfo> keys;
        std::array<float, 12> parent;  // static transform of the Group it sits in
    };
    std::vector<uint32_t> instanceFirst;
    std::vector<Motion> motions;
};
SceneExport exportScene(Tracer& tracer, const std::shared_ptr<Node>& scene);

// ---- module (Interface/Infrastructure/Module.hpp) ----
class PiperContext {  // the host's context; this module needs nothing from it
public:
    virtual ~PiperContext() = default;
};
class Allocator {  // Interface/Infrastructure/Allocator.hpp: the host's allocator
public:
    virtual ~Allocator() = default;
    virtual void* alloc(size_t size, size_t align) = 0;
    virtual void free(void* ptr) = 0;
};
class Module : public Object {  // Module.hpp:26-32
public:
    [[nodiscard]] virtual std::shared_ptr<Object> newInstance(std::string_view classID, const Config& config) = 0;
};

}  // namespace Piper

// Module entry points (PIPER_INIT_MODULE_IMPL, Module.hpp:54-70). piperInitModule builds the
// module in memory from `allocator` (the host destroys it and frees that memory); its classes are
// "Accelerator" (config: optional "Device" ordinal, or "HostOnly": true) and "Tracer" (config:
// optional "Accelerator" config). piperGetProtocol returns "GCC@STD@0" (ABI, STL, interface
// version), the protocol a GCC host with std containers checks.
extern "C" Piper::Module* piperInitModule(Piper::PiperContext& context, Piper::Allocator& allocator, const char* path);
extern "C" const char* piperGetProtocol();

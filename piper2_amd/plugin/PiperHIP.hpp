// The MI355X component pair behind Piper2's plugin interfaces:
//   HIPAccelerator : Accelerator (Interface/Infrastructure/Accelerator.hpp:130-161)
//   HIPTracer      : Tracer      (Interface/BuiltinComponent/Tracer.hpp:140-159)
// with the Pipeline and TraceLauncher (Tracer.hpp:96-116) it returns, and the FixedSampler render
// driver (RenderDriver/FixedSampler/FixedSampler.cpp:60-133) that drives a frame through them.
//
// Piper's own runtime (EASTL-based containers, SharedPtr, PiperContext, module loader, scheduler)
// is not part of this build, so the interfaces the path touches are restated here with std types.
// Each method keeps the reference's name, arguments and meaning. Errors: the reference hands them
// to ErrorHandler::raiseException, which ends the program; here they throw PiperException carrying
// the C ABI's message, so the host application decides.
//
// Everything below calls the C ABI of include/piper_hip.h; nothing here touches the device directly.
#pragma once
#include <array>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/piper_hip.h"

namespace Piper {

class PiperException : public std::runtime_error {
public:
    explicit PiperException(const std::string& msg) : std::runtime_error(msg) {}
};

// ---- value types (restated) ----
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
struct TransformInfo final {  // Tracer.hpp:80-84 (keyframes; only static transforms are supported)
    float offset = 0.0f, step = 0.0f;
    std::vector<std::pair<uint32_t, TransformSRT>> transforms;  // local to world
};

using SBTPayload = std::vector<std::byte>;  // Tracer.hpp:86-91
template <typename T>
SBTPayload packSBTPayload(const T& data) {
    SBTPayload p(sizeof(T));
    std::memcpy(p.data(), &data, sizeof(T));
    return p;
}

// Payload of one FixedSampler launch (RenderDriver/FixedSampler/Shared.hpp:25-28).
struct LaunchData final {
    uint32_t rgbw;  // index of the film among the launch's resources
    uint32_t w, h;
};

// A program the tracer can link: the reference hands LLVM-IR modules to the JIT; here a program
// names the built-in kernels it provides.
struct LinkableProgram final {
    std::string format;
    std::vector<std::string> symbols;
};

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

// ---- resources ----
class Resource {
public:
    virtual ~Resource() = default;
};
class TiledOutput : public Resource {  // Accelerator.hpp:80-87
public:
    [[nodiscard]] virtual Future<std::vector<std::byte>> download() const = 0;
    virtual void markDirty(FutureVoid future) = 0;
};

class Accelerator {  // Accelerator.hpp:130-161 (the JIT entry points are not part of this path)
public:
    virtual ~Accelerator() = default;
    [[nodiscard]] virtual const char* getNativePlatform() const noexcept = 0;
    [[nodiscard]] virtual std::vector<std::string> getSupportedLinkableFormat() const = 0;
    [[nodiscard]] virtual std::shared_ptr<Resource> createBuffer(size_t size, size_t alignment,
                                                                 std::function<void(std::byte*)> data) = 0;
    [[nodiscard]] virtual std::shared_ptr<TiledOutput> createTiledOutput(size_t size, size_t alignment) = 0;
};

// ---- scene objects (Tracer.hpp:33-63) ----
class Object {
public:
    virtual ~Object() = default;
};
class AccelerationStructure : public Object {};
class Node : public Object {};
class RTProgram : public Object {};
class GSMInstance : public Object {};

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

// Components the tracer materialises. In the reference each is a plugin whose materialize() yields
// programs and an SBT payload; here each carries its payload directly.
class Geometry : public Object {
public:
    explicit Geometry(GeometryDesc desc) : mDesc(std::move(desc)) {}
    const GeometryDesc& desc() const noexcept { return mDesc; }

private:
    GeometryDesc mDesc;
};
// Surface/Classic: Matte (MatteData, Classic/Shared.hpp:51-54 with a constant texture), Plastic
// (PlasticData, Classic/Shared.hpp:56-60 and Classic.cpp:123-147: Diffuse, Specular, Roughness as
// constant textures) and the GGX conductor this path adds (alpha = roughness^2, Schlick F0).
// Texture/ProceduralTexture CheckBoard (Black, White, Scale) and Texture/SampledTexture (an 8-bit
// image with Repeat or Mirror wrap): bound to a surface's colour inputs.
class Texture : public Object {
public:
    static std::shared_ptr<Texture> checkerboard(std::array<float, 4> black, std::array<float, 4> white, float scale,
                                                 uint32_t channels = 4);
    static std::shared_ptr<Texture> image(uint32_t width, uint32_t height, uint32_t channels, uint32_t wrap,
                                          std::vector<uint8_t> texels);
    pg_texture desc{};             // texels points into `texels`
    std::vector<uint8_t> texels;
};
class Surface : public Object {
public:
    enum class Kind { Matte, Conductor, Plastic, Mirror, Glass, Substrate };
    // roughness 0 = Lambert, in (0, 1] Oren-Nayar (matteInit)
    static std::shared_ptr<Surface> matte(std::array<float, 3> reflectance, float roughness = 0.0f);
    static std::shared_ptr<Surface> conductor(std::array<float, 3> f0, float roughness);
    // roughness as the reference's config gives it: the library applies remapRoughness
    static std::shared_ptr<Surface> plastic(std::array<float, 3> diffuse, std::array<float, 3> specular, float roughness);
    // Mirror (Classic.cpp:148-170: Reflection)
    static std::shared_ptr<Surface> mirror(std::array<float, 3> reflection);
    // Glass (Reflection, Transmission, RoughnessX, RoughnessY): smooth with roughness 0, rough
    // (microfacet reflection + transmission) otherwise; isotropic only
    static std::shared_ptr<Surface> glass(std::array<float, 3> reflection, std::array<float, 3> transmission,
                                          float roughnessX = 0.0f, float roughnessY = 0.0f);
    // Substrate (Classic.cpp:172-195: Diffuse, Specular, RoughnessX, RoughnessY): isotropic only
    static std::shared_ptr<Surface> substrate(std::array<float, 3> diffuse, std::array<float, 3> specular,
                                              float roughnessX, float roughnessY);
    Kind kind;
    std::array<float, 3> color;          // matte reflectance, conductor F0, plastic / substrate diffuse,
                                         // mirror / glass reflection
    float alpha;                         // matte roughness, conductor GGX alpha, plastic / substrate roughness
    std::array<float, 3> specular{};     // plastic / substrate specular, glass transmission
    std::shared_ptr<Texture> colorTexture, specularTexture;  // replace color / specular when set
    std::shared_ptr<Texture> roughnessTexture;               // replaces the roughness (channel 0) when set
};
class Medium : public Object {};
// Light/AreaLight with a Geometry/BasicShapes plane (origin, U, V), the environment
// (Light/BasicLight SampledEnvironment with constant radiance) and the point light
// (BasicLight.cpp:55-80: Intensity, placed at its node's origin).
class Light : public Object {
public:
    static std::shared_ptr<Light> areaPlane(std::array<float, 3> origin, std::array<float, 3> u,
                                            std::array<float, 3> v, std::array<float, 3> radiance);
    static std::shared_ptr<Light> environment(std::array<float, 3> radiance);
    static std::shared_ptr<Light> point(std::array<float, 3> intensity);
    bool isEnvironment = false, isPoint = false;
    std::array<float, 3> origin{}, u{}, v{}, radiance{};  // radiance: the point light's intensity
};
// Sensor/PerspectiveCamera (PCData, Sensor/PerspectiveCamera/Shared.hpp:22-29; config
// PerspectiveCamera.cpp:39-45). The camera sits at its node's translation.
class Sensor : public Object {
public:
    std::array<float, 3> lookAt{}, upRef{0.0f, 1.0f, 0.0f};
    std::array<float, 2> size{0.036f, 0.024f};
    float focalLength = 0.035f;
    float apertureRadius = 0.0f;
    float getAspectRatio() const noexcept { return size[0] / size[1]; }
};
// Integrator/PathTracing (Data, Integrator/PathTracing/Shared.hpp:22-29).
class Integrator : public Object {
public:
    uint32_t maxDepth = 8, minDepthForRussianRoulette = 3;
    uint32_t compatFlags = 0;  // PG_COMPAT_*: reproduce the reference's documented deviations
};
class LightSampler : public Object {};  // LightSampler/UniformLightSampler
class Sampler : public Object {         // Sampler/ScrambledSobolSampler
public:
    uint32_t samplesPerPixel = 16;
};
class Tracer;
class TraceLauncher;
class RenderDriver : public Object {  // Interface/BuiltinComponent/RenderDriver.hpp
public:
    // RGB radiance per film pixel (rgb / w), row-major
    [[nodiscard]] virtual std::vector<std::array<float, 3>> renderFrame(Tracer& tracer, TraceLauncher& launcher) = 0;
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

class Tracer {  // Tracer.hpp:140-159
public:
    virtual ~Tracer() = default;
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
    [[nodiscard]] virtual size_t getAlignmentRequirement(AlignmentRequirement requirement) const noexcept = 0;
    [[nodiscard]] virtual Accelerator& getAccelerator() const noexcept = 0;
};

// ---- the MI355X pair ----
class HIPAccelerator final : public Accelerator {
public:
    explicit HIPAccelerator(int ordinal);  // pg_device_create
    // An accelerator without a device: scene graphs build (buffers are host memory) and export,
    // but pipelines and tiled outputs raise.
    static std::shared_ptr<HIPAccelerator> hostOnly();
    ~HIPAccelerator() override;
    HIPAccelerator(const HIPAccelerator&) = delete;
    HIPAccelerator& operator=(const HIPAccelerator&) = delete;
    [[nodiscard]] const char* getNativePlatform() const noexcept override;
    [[nodiscard]] std::vector<std::string> getSupportedLinkableFormat() const override;
    [[nodiscard]] std::shared_ptr<Resource> createBuffer(size_t size, size_t alignment,
                                                         std::function<void(std::byte*)> data) override;
    [[nodiscard]] std::shared_ptr<TiledOutput> createTiledOutput(size_t size, size_t alignment) override;
    pg_device* device() const noexcept { return mDevice; }

private:
    HIPAccelerator() = default;
    pg_device* mDevice = nullptr;
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
    [[nodiscard]] std::shared_ptr<RTProgram> buildProgram(const LinkableProgram& linkable, std::string symbol) override;
    [[nodiscard]] std::unique_ptr<Pipeline> buildPipeline(const std::shared_ptr<Node>& scene, Integrator& integrator,
                                                          RenderDriver& renderDriver, LightSampler& lightSampler,
                                                          std::shared_ptr<Sampler> sampler) override;
    [[nodiscard]] size_t getAlignmentRequirement(AlignmentRequirement requirement) const noexcept override;
    [[nodiscard]] Accelerator& getAccelerator() const noexcept override;

private:
    std::shared_ptr<HIPAccelerator> mAccelerator;
};

// The FixedSampler's filter (SimpleFilter, RenderDriver/FixedSampler/FixedSampler.cpp:28-54):
// FilterType "box", "triangle" or "gaussian" (Alpha > 0), weights as FixedSampler/Kernel.cpp:55-70.
struct SimpleFilter {
    std::string filterType = "box";
    float alpha = 0.0f;
};

// RenderDriver/FixedSampler: 8x8 tiles, one launch each, then rgb / w per pixel
// (FixedSampler.cpp:68-133).
class FixedSampler final : public RenderDriver {
public:
    SimpleFilter filter;
    [[nodiscard]] std::vector<std::array<float, 3>> renderFrame(Tracer& tracer, TraceLauncher& launcher) override;
    std::vector<float> lastFilm;  // RGBW of the last frame, for inspection
};

// The built-in kernel symbols a HIP pipeline provides (buildProgram accepts only these).
const std::vector<std::string>& hipBuiltinSymbols();

// What buildPipeline uploads for a scene graph: world-space triangles grouped by instance, one
// material per instance, the lights (environment, quads, points) and every sensor with its node's
// local-to-world transform (row-major 3x4). No device is needed.
struct SceneExport {
    std::vector<float> verts;
    std::vector<uint32_t> index, triObject, objectMaterial;
    std::vector<pg_material> materials;
    std::vector<pg_quad> quads;
    std::vector<pg_point_light> points;
    std::array<float, 3> env{};
    std::vector<std::pair<const Node*, std::array<float, 12>>> sensors;
    std::vector<float> texcoords;                        // float2 per vertex, or empty
    std::vector<std::shared_ptr<Texture>> textures;
    std::vector<pg_material_textures> materialTextures;  // per material (empty without textures)
};
SceneExport exportScene(const std::shared_ptr<Node>& scene);

}  // namespace Piper

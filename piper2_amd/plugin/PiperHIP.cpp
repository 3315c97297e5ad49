// HIPAccelerator / HIPTracer / FixedSampler over the C ABI (declarations and the reference
// interfaces they implement: PiperHIP.hpp).
#include "PiperHIP.hpp"

#include <algorithm>
#include <cmath>
#include <sstream>

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

// ---- scene objects ----
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
    std::vector<std::pair<Mat34, std::shared_ptr<Node>>> children;
};
struct HIPProgram final : RTProgram {
    std::string symbol;
};

// Flattened scene: world-space triangles grouped by instance, one material per instance.
struct FlatScene {
    std::vector<float> verts;
    std::vector<uint32_t> index, triObject, objectMaterial;
    std::vector<pg_material> materials;
    std::vector<pg_quad> quads;
    std::vector<pg_point_light> points;
    std::vector<float> texcoords;  // float2 per vertex once any mesh has them (zeros elsewhere)
    std::vector<std::shared_ptr<Texture>> textures;
    std::vector<pg_material_textures> materialTextures;
    float env[3] = {0, 0, 0};
    std::vector<std::pair<const Node*, Mat34>> sensors;  // leaf nodes of sensors and their transform
};

uint32_t materialOf(FlatScene& f, const Surface& s) {
    pg_material m{};
    switch (s.kind) {
        case Surface::Kind::Matte: m.type = PG_MAT_LAMBERT; break;
        case Surface::Kind::Conductor: m.type = PG_MAT_CONDUCTOR; break;
        case Surface::Kind::Plastic: m.type = PG_MAT_PLASTIC; break;
        case Surface::Kind::Mirror: m.type = PG_MAT_MIRROR; break;
        case Surface::Kind::Glass: m.type = PG_MAT_GLASS; break;
        case Surface::Kind::Substrate: m.type = PG_MAT_SUBSTRATE; break;
    }
    for (int k = 0; k < 3; ++k) m.albedo[k] = s.color[k];
    m.alpha = s.alpha;
    for (int k = 0; k < 3; ++k) m.reserved[k] = s.specular[k];
    f.materials.push_back(m);
    auto texIndex = [&](const std::shared_ptr<Texture>& t) -> uint32_t {
        if (!t) return PG_TEX_NONE;
        for (size_t i = 0; i < f.textures.size(); ++i)
            if (f.textures[i] == t) return static_cast<uint32_t>(i);
        f.textures.push_back(t);
        return static_cast<uint32_t>(f.textures.size() - 1);
    };
    f.materialTextures.push_back({texIndex(s.colorTexture), texIndex(s.specularTexture), texIndex(s.roughnessTexture)});
    return static_cast<uint32_t>(f.materials.size() - 1);
}

void flatten(const std::shared_ptr<Node>& node, const Mat34& toWorld, FlatScene& f) {
    if (auto* g = dynamic_cast<const GroupNode*>(node.get())) {
        for (const auto& [m, child] : g->children) flatten(child, compose(toWorld, m), f);
        return;
    }
    auto* leaf = dynamic_cast<const LeafNode*>(node.get());
    if (!leaf) raise("buildPipeline: node was not built by this tracer");
    if (auto* inst = dynamic_cast<const HIPGSMInstance*>(leaf->object.get())) {
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
        f.objectMaterial.push_back(materialOf(f, *inst->surface));
        for (uint32_t i : inst->mesh->index) f.index.push_back(base + i);
        f.triObject.insert(f.triObject.end(), inst->mesh->index.size() / 3, obj);
    } else if (auto* light = dynamic_cast<const Light*>(leaf->object.get())) {
        if (light->isEnvironment) {
            for (int k = 0; k < 3; ++k) f.env[k] = light->radiance[k];
        } else if (light->isPoint) {  // pointInit: the node's origin in world space
            const float zero[3] = {0.0f, 0.0f, 0.0f};
            const auto o = applyPoint(toWorld, zero);
            pg_point_light p{};
            for (int k = 0; k < 3; ++k) p.position[k] = o[k], p.intensity[k] = light->radiance[k];
            f.points.push_back(p);
        } else {
            pg_quad q{};
            const auto o = applyPoint(toWorld, light->origin.data());
            const auto u = applyVector(toWorld, light->u.data());
            const auto v = applyVector(toWorld, light->v.data());
            for (int k = 0; k < 3; ++k) q.origin[k] = o[k], q.u[k] = u[k], q.v[k] = v[k], q.radiance[k] = light->radiance[k];
            f.quads.push_back(q);
        }
    } else if (dynamic_cast<const Sensor*>(leaf->object.get())) {
        f.sensors.push_back({node.get(), toWorld});
    } else {
        raise("buildPipeline: unsupported scene object");
    }
}

class HIPTraceLauncher final : public TraceLauncher {
public:
    struct Tile {
        RenderRECT rect;
        HIPTiledOutput* film;
    };
    struct Shared {  // pending tiles, flushed by the first future that is polled
        pg_scene* scene;
        pg_render_params base;
        std::vector<Tile> pending;
        std::vector<pg_stats> stats;
    };

    HIPTraceLauncher(std::shared_ptr<pg_scene> scene, uint32_t w, uint32_t h, RenderRECT rect,
                     SensorNDCAffineTransform ndc, uint32_t spp, const Integrator& integrator, uint32_t filter,
                     float filterAlpha)
        : mScene(std::move(scene)), mW(w), mH(h), mRect(rect), mShared(std::make_shared<Shared>()) {
        pg_render_params& p = mShared->base;
        p = pg_render_params{};
        p.width = w, p.height = h, p.spp = spp;
        p.spp_begin = 0, p.spp_end = spp;
        p.max_depth = integrator.maxDepth, p.min_rr = integrator.minDepthForRussianRoulette;
        p.flags = integrator.compatFlags | PG_NO_STAGE_TIMING;
        p.ndc[0] = ndc.ox, p.ndc[1] = ndc.oy, p.ndc[2] = ndc.sx, p.ndc[3] = ndc.sy;
        p.filter = filter, p.filter_alpha = filterAlpha;
        mShared->scene = mScene.get();
    }
    [[nodiscard]] std::pair<uint32_t, uint32_t> getFilmResolution() const noexcept override { return {mW, mH}; }
    [[nodiscard]] RenderRECT getRenderRECT() const noexcept override { return mRect; }
    // Static scenes: the time interval selects nothing (ray time is still Sobol dimension 2).
    void updateTimeInterval(float begin, float end) noexcept override { mTime = {begin, end}; }

    [[nodiscard]] FutureVoid launch(const RenderRECT& rect, const std::function<SBTPayload(uint32_t)>& launchData,
                                    const std::vector<std::shared_ptr<Resource>>& resources) override {
        const SBTPayload payload = launchData(0);
        if (payload.size() != sizeof(LaunchData)) raise("launch: payload is not a FixedSampler LaunchData");
        LaunchData ld;
        std::memcpy(&ld, payload.data(), sizeof(ld));
        if (ld.w != mW || ld.h != mH) raise("launch: LaunchData resolution differs from the film");
        if (ld.rgbw >= resources.size()) raise("launch: LaunchData names a missing resource");
        auto* film = dynamic_cast<HIPTiledOutput*>(resources[ld.rgbw].get());
        if (!film) raise("launch: the output was not created by HIPAccelerator");
        if (rect.width == 0 || rect.height == 0 || rect.left + rect.width > mW || rect.top + rect.height > mH)
            raise("launch: tile outside the film");
        film->bind(mW, mH);
        mShared->pending.push_back({rect, film});
        std::shared_ptr<Shared> shared = mShared;
        std::shared_ptr<pg_scene> keep = mScene;
        return FutureVoid([shared, keep] {
            flush(*shared);
            return Unit{};
        });
    }
    std::vector<pg_stats> stats() const { return mShared->stats; }

private:
    // Tiles of one film that exactly cover their bounding rect render as one call over that rect
    // (the same film as one pg_render of it); otherwise each tile renders on its own.
    static void flush(Shared& s) {
        while (!s.pending.empty()) {
            HIPTiledOutput* film = s.pending.front().film;
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
            pg_film* f = film->bind(s.base.width, s.base.height);
            for (const RenderRECT& r : tiles) {
                pg_render_params p = s.base;
                p.rect[0] = r.left, p.rect[1] = r.top, p.rect[2] = r.width, p.rect[3] = r.height;
                pg_stats st{};
                check(pg_render(s.scene, &p, f, &st), "pg_render");
                s.stats.push_back(st);
            }
        }
    }

    std::shared_ptr<pg_scene> mScene;
    uint32_t mW, mH;
    RenderRECT mRect;
    std::pair<float, float> mTime{0.0f, 0.0f};
    std::shared_ptr<Shared> mShared;
};

class HIPPipeline final : public Pipeline {
public:
    HIPPipeline(pg_device* dev, FlatScene flat, const Integrator& integrator, std::shared_ptr<Sampler> sampler,
                uint32_t filter, float filterAlpha)
        : mDevice(dev), mFlat(std::move(flat)), mIntegrator(integrator), mSampler(std::move(sampler)), mFilter(filter),
          mFilterAlpha(filterAlpha) {}

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
    // translation is the camera position, PerspectiveCamera.cpp:39-45), then the film rect and
    // NDC transform follow calcRenderRECT.
    [[nodiscard]] std::shared_ptr<TraceLauncher> prepare(const std::shared_ptr<Node>& sensor, uint32_t width,
                                                         uint32_t height, FitMode fitMode) override {
        const Mat34* toWorld = nullptr;
        for (const auto& [n, m] : mFlat.sensors)
            if (n == sensor.get()) toWorld = &m;
        if (!toWorld) raise("prepare: the sensor node is not part of the pipeline's scene");
        const auto* leaf = dynamic_cast<const LeafNode*>(sensor.get());
        const auto& cam = dynamic_cast<const Sensor&>(*leaf->object);
        pg_scene_desc d{};
        d.vertices = mFlat.verts.data();
        d.num_vertices = static_cast<uint32_t>(mFlat.verts.size() / 3);
        d.indices = mFlat.index.data();
        d.num_triangles = static_cast<uint32_t>(mFlat.index.size() / 3);
        d.tri_object = mFlat.triObject.data();
        d.object_material = mFlat.objectMaterial.data();
        d.num_objects = static_cast<uint32_t>(mFlat.objectMaterial.size());
        d.materials = mFlat.materials.data();
        d.num_materials = static_cast<uint32_t>(mFlat.materials.size());
        d.quads = mFlat.quads.data();
        d.num_quads = static_cast<uint32_t>(mFlat.quads.size());
        d.point_lights = mFlat.points.empty() ? nullptr : mFlat.points.data();
        d.num_point_lights = static_cast<uint32_t>(mFlat.points.size());
        std::vector<pg_texture> textures;
        for (const auto& t : mFlat.textures) {
            textures.push_back(t->desc);
            textures.back().texels = t->texels.empty() ? nullptr : t->texels.data();
        }
        if (!textures.empty()) {
            d.textures = textures.data();
            d.num_textures = static_cast<uint32_t>(textures.size());
            d.material_textures = mFlat.materialTextures.data();
            d.texcoords = mFlat.texcoords.empty() ? nullptr : mFlat.texcoords.data();
        }
        for (int k = 0; k < 3; ++k) {
            d.camera.position[k] = (*toWorld)[4 * k + 3];
            d.camera.look_at[k] = cam.lookAt[k];
            d.camera.up[k] = cam.upRef[k];
            d.env_radiance[k] = mFlat.env[k];
        }
        d.camera.sensor_size[0] = cam.size[0], d.camera.sensor_size[1] = cam.size[1];
        d.camera.focal_length = cam.focalLength;
        d.camera.aperture_radius = cam.apertureRadius;
        if (d.num_materials == 0) {  // a scene of lights only still needs one material
            static const pg_material none{PG_MAT_LAMBERT, {0.0f, 0.0f, 0.0f}, 0.0f, {0.0f, 0.0f, 0.0f}};
            d.materials = &none;
            d.num_materials = 1;
        }
        pg_scene* raw = nullptr;
        check(pg_scene_create(mDevice, &d, &raw), "pg_scene_create");
        std::shared_ptr<pg_scene> scene(raw, pg_scene_destroy);
        const auto [ndc, rect] = calcRenderRECT(width, height, cam.getAspectRatio(), fitMode);
        auto launcher = std::make_shared<HIPTraceLauncher>(scene, width, height, rect, ndc, mSampler->samplesPerPixel,
                                                           mIntegrator, mFilter, mFilterAlpha);
        mLaunchers.push_back(launcher);
        return launcher;
    }

private:
    pg_device* mDevice;
    FlatScene mFlat;
    Integrator mIntegrator;
    std::shared_ptr<Sampler> mSampler;
    uint32_t mFilter;
    float mFilterAlpha;
    std::vector<std::weak_ptr<HIPTraceLauncher>> mLaunchers;
};

}  // namespace

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

std::shared_ptr<Surface> Surface::matte(std::array<float, 3> reflectance, float roughness) {
    if (!(roughness >= 0.0f && roughness <= 1.0f)) raise("matte roughness must be in [0, 1]");
    auto s = std::make_shared<Surface>();
    s->kind = Kind::Matte, s->color = reflectance, s->alpha = roughness;
    return s;
}
std::shared_ptr<Surface> Surface::conductor(std::array<float, 3> f0, float roughness) {
    if (!(roughness > 0.0f)) raise("conductor roughness must be positive");
    auto s = std::make_shared<Surface>();
    s->kind = Kind::Conductor, s->color = f0, s->alpha = roughness * roughness;
    return s;
}
std::shared_ptr<Surface> Surface::plastic(std::array<float, 3> diffuse, std::array<float, 3> specular, float roughness) {
    if (!std::isfinite(roughness)) raise("plastic roughness must be finite");
    auto s = std::make_shared<Surface>();
    s->kind = Kind::Plastic, s->color = diffuse, s->alpha = roughness, s->specular = specular;
    return s;
}
std::shared_ptr<Texture> Texture::checkerboard(std::array<float, 4> black, std::array<float, 4> white, float scale,
                                               uint32_t channels) {
    if (channels != 1 && channels != 2 && channels != 4) raise("Texture: channels must be 1, 2 or 4");
    auto t = std::make_shared<Texture>();
    t->desc.type = PG_TEX_CHECKERBOARD, t->desc.channels = channels, t->desc.scale = scale;
    std::memcpy(t->desc.black, black.data(), 16), std::memcpy(t->desc.white, white.data(), 16);
    return t;
}
std::shared_ptr<Texture> Texture::image(uint32_t width, uint32_t height, uint32_t channels, uint32_t wrap,
                                        std::vector<uint8_t> texels) {
    if (channels != 1 && channels != 2 && channels != 4) raise("Texture: channels must be 1, 2 or 4");
    if (wrap != PG_WRAP_REPEAT && wrap != PG_WRAP_MIRROR) raise("Texture: unknown wrap mode");
    if (width == 0 || height == 0 || texels.size() != static_cast<size_t>(width) * height * channels)
        raise("Texture: texel count does not match width x height x channels");
    auto t = std::make_shared<Texture>();
    t->desc.type = PG_TEX_IMAGE, t->desc.channels = channels, t->desc.width = width, t->desc.height = height;
    t->desc.wrap = wrap;
    t->texels = std::move(texels);
    return t;
}
std::shared_ptr<Surface> Surface::mirror(std::array<float, 3> reflection) {
    auto s = std::make_shared<Surface>();
    s->kind = Kind::Mirror, s->color = reflection, s->alpha = 0.0f;
    return s;
}
std::shared_ptr<Surface> Surface::glass(std::array<float, 3> reflection, std::array<float, 3> transmission,
                                        float roughnessX, float roughnessY) {
    if (roughnessX != roughnessY) raise("glass: anisotropic roughness is not supported");
    if (!(roughnessX >= 0.0f) || !std::isfinite(roughnessX)) raise("glass roughness must be finite and not negative");
    auto s = std::make_shared<Surface>();
    s->kind = Kind::Glass, s->color = reflection, s->alpha = roughnessX, s->specular = transmission;
    return s;
}
std::shared_ptr<Surface> Surface::substrate(std::array<float, 3> diffuse, std::array<float, 3> specular,
                                            float roughnessX, float roughnessY) {
    if (roughnessX != roughnessY) raise("substrate: anisotropic roughness is not supported");
    if (!std::isfinite(roughnessX)) raise("substrate roughness must be finite");
    auto s = std::make_shared<Surface>();
    s->kind = Kind::Substrate, s->color = diffuse, s->alpha = roughnessX, s->specular = specular;
    return s;
}
std::shared_ptr<Light> Light::areaPlane(std::array<float, 3> origin, std::array<float, 3> u, std::array<float, 3> v,
                                        std::array<float, 3> radiance) {
    auto l = std::make_shared<Light>();
    l->origin = origin, l->u = u, l->v = v, l->radiance = radiance;
    return l;
}
std::shared_ptr<Light> Light::point(std::array<float, 3> intensity) {
    auto l = std::make_shared<Light>();
    l->isPoint = true, l->radiance = intensity;
    return l;
}
std::shared_ptr<Light> Light::environment(std::array<float, 3> radiance) {
    auto l = std::make_shared<Light>();
    l->isEnvironment = true, l->radiance = radiance;
    return l;
}

const std::vector<std::string>& hipBuiltinSymbols() {
    // the per-sample and per-bounce functions the wavefront kernels implement natively
    static const std::vector<std::string> s = {
        "rayGen",      "sobolStart", "sobolGenerate", "trace",         "accumulate", "box",
        "matteInit",   "conductorInit", "plasticInit", "plasticSample", "plasticEvaluate", "plasticPdf",
        "mirrorInit",  "mirrorSample", "blackBodyEvaluate", "blackBodyPdf", "glassInit", "glassSample",
        "glassEvaluate", "glassPdf", "substrateInit", "substrateSample", "substrateEvaluate", "substratePdf",
        "diffuseSample", "diffuseEvaluate", "diffusePdf", "planeSample",
        "uniformSelect", "pointInit", "pointSample", "deltaEvaluate", "deltaPdf", "SEInit",   "SESample",      "SEEvaluate",    "SEPdf",      "constantTexture"};
    return s;
}

// ---- HIPAccelerator ----
HIPAccelerator::HIPAccelerator(int ordinal) { check(pg_device_create(ordinal, &mDevice), "pg_device_create"); }
std::shared_ptr<HIPAccelerator> HIPAccelerator::hostOnly() {
    return std::shared_ptr<HIPAccelerator>(new HIPAccelerator());
}
HIPAccelerator::~HIPAccelerator() {
    if (mDevice) pg_device_destroy(mDevice);
}
const char* HIPAccelerator::getNativePlatform() const noexcept {
    return mDevice ? pg_device_platform(mDevice) : "host (no device)";
}
std::vector<std::string> HIPAccelerator::getSupportedLinkableFormat() const { return {"HIP-gfx950-builtin"}; }
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

std::shared_ptr<Node> HIPTracer::buildNode(const std::vector<std::pair<TransformInfo, std::shared_ptr<Node>>>& children) {
    auto g = std::make_shared<GroupNode>();
    for (const auto& [info, child] : children) {
        if (!child) raise("buildNode: null child");
        if (info.transforms.size() > 1)
            raise("buildNode: time-sampled transforms (motion blur) are not supported by the HIP tracer");
        Mat34 m = identity34();
        if (!info.transforms.empty() && !isIdentity(info.transforms.front().second))
            m = fromSRT(info.transforms.front().second);
        g->children.push_back({m, child});
    }
    return g;
}

std::shared_ptr<RTProgram> HIPTracer::buildProgram(const LinkableProgram& linkable, std::string symbol) {
    const auto formats = mAccelerator->getSupportedLinkableFormat();
    if (std::find(formats.begin(), formats.end(), linkable.format) == formats.end())
        raise("buildProgram: unsupported linkable format " + linkable.format);
    if (std::find(linkable.symbols.begin(), linkable.symbols.end(), symbol) == linkable.symbols.end())
        raise("buildProgram: the program does not provide " + symbol);
    const auto& known = hipBuiltinSymbols();
    if (std::find(known.begin(), known.end(), symbol) == known.end())
        raise("buildProgram: " + symbol + " has no native HIP implementation");
    auto p = std::make_shared<HIPProgram>();
    p->symbol = std::move(symbol);
    return p;
}

std::unique_ptr<Pipeline> HIPTracer::buildPipeline(const std::shared_ptr<Node>& scene, Integrator& integrator,
                                                   RenderDriver& renderDriver, LightSampler& lightSampler,
                                                   std::shared_ptr<Sampler> sampler) {
    (void)lightSampler;  // uniform light selection is built in
    if (!scene || !sampler) raise("buildPipeline: scene and sampler are required");
    if (!mAccelerator->device()) raise("buildPipeline: the accelerator has no device");
    // the FixedSampler's filter becomes the film filter of every launch (its RDData payload)
    uint32_t filter = PG_FILTER_BOX;
    float alpha = 0.0f;
    if (const auto* fs = dynamic_cast<const FixedSampler*>(&renderDriver)) {
        const std::string& t = fs->filter.filterType;
        if (t == "triangle") filter = PG_FILTER_TRIANGLE;
        else if (t == "gaussian") filter = PG_FILTER_GAUSSIAN, alpha = fs->filter.alpha;
        else if (t != "box") raise("buildPipeline: unknown FilterType " + t);
        if (filter == PG_FILTER_GAUSSIAN && !(alpha > 0.0f)) raise("buildPipeline: gaussian filter needs Alpha > 0");
    } else {
        raise("buildPipeline: only the FixedSampler render driver is supported");
    }
    FlatScene flat;
    flatten(scene, identity34(), flat);
    return std::make_unique<HIPPipeline>(mAccelerator->device(), std::move(flat), integrator, std::move(sampler), filter,
                                         alpha);
}

SceneExport exportScene(const std::shared_ptr<Node>& scene) {
    if (!scene) raise("exportScene: null scene");
    FlatScene flat;
    flatten(scene, identity34(), flat);
    SceneExport e;
    e.verts = std::move(flat.verts);
    e.index = std::move(flat.index);
    e.triObject = std::move(flat.triObject);
    e.objectMaterial = std::move(flat.objectMaterial);
    e.materials = std::move(flat.materials);
    e.quads = std::move(flat.quads);
    e.points = std::move(flat.points);
    for (int k = 0; k < 3; ++k) e.env[k] = flat.env[k];
    for (const auto& [n, m] : flat.sensors) e.sensors.push_back({n, m});
    if (!flat.textures.empty()) {
        e.texcoords = std::move(flat.texcoords);
        e.textures = std::move(flat.textures);
        e.materialTextures = std::move(flat.materialTextures);
    }
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

}  // namespace Piper

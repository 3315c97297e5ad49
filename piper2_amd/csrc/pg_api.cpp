// C ABI of libpiper_hip.so: devices, scenes, films, renders and parity entry points.
// Declarations and the reference interfaces they replace: include/piper_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/piper_hip.h"
#include "pg_bvh.hpp"
#include "pg_guide_kernels.hpp"
#include "pg_kernels.hpp"

namespace {

thread_local std::string g_lastError;

struct PgError : std::runtime_error {
    int code;
    PgError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define PG_HIP(call)                                                                                  \
    do {                                                                                              \
        const hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                         \
            throw PgError(e_ == hipErrorOutOfMemory ? PG_ERR_OOM : PG_ERR_DEVICE,                     \
                          std::string(#call) + ": " + hipGetErrorString(e_));                         \
    } while (0)

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return PG_OK;
    } catch (const PgError& e) {
        g_lastError = e.what();
        return e.code;
    } catch (const std::invalid_argument& e) {
        g_lastError = e.what();
        return PG_ERR_INVALID;
    } catch (const std::exception& e) {
        g_lastError = e.what();
        return PG_ERR_DEVICE;
    }
}

void require(bool ok, const char* msg) {
    if (!ok) throw PgError(PG_ERR_INVALID, msg);
}

// Device buffer owned by the library.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void reserve(size_t n) {
        if (n <= bytes) return;
        release();
        PG_HIP(hipMalloc(&p, n));
        bytes = n;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

void upload(DevBuf& b, const void* host, size_t n, hipStream_t st) {
    b.reserve(std::max<size_t>(n, 16));
    if (n) PG_HIP(hipMemcpyAsync(b.p, host, n, hipMemcpyHostToDevice, st));
}

// Render progress the host reads back: the pool counters, then the recorded-path counts of up to
// kStatusItems work items from the oldest unresolved one.
constexpr size_t kStatusItems = 64;
constexpr size_t kStatusBytes = pg::kCountsBytes + kStatusItems * 4;
// ... followed by the pool's generation limit on its way to the device (8 B, padded).
constexpr size_t kPinnedBytes = kStatusBytes + 64;

// Instantiated graph, destroyed with its owner.
struct GraphExec {
    hipGraphExec_t exec = nullptr;
    GraphExec() = default;
    GraphExec(const GraphExec&) = delete;
    GraphExec& operator=(const GraphExec&) = delete;
    ~GraphExec() {
        if (exec) (void)hipGraphExecDestroy(exec);
    }
};

// Traversal scratch of pg_trace_noalloc: the stack entries that overflow a lane's LDS slice spill
// here instead of into a per-call allocation. Each GPU the library runs on holds its own copy
// from module load until the process ends. Host code never uses the array's name as an address:
// it takes the device's copy once per device (scratchAddress, at pg_device_create) and passes
// that pointer on. Every pg_trace_noalloc call on one GPU shares that GPU's copy, so calls that
// overlap in time on one GPU race.
constexpr size_t kScratchBytes = size_t(16) << 20;
static __device__ __align__(256) uint2 scratch[kScratchBytes / sizeof(uint2)];

constexpr int kMaxOrdinals = 64;
std::mutex g_scratchMu;
void* g_scratchPtr[kMaxOrdinals] = {};

// The device's copy of `scratch`, looked up once per device; `ordinal` must be current.
void* scratchAddress(int ordinal) {
    std::lock_guard<std::mutex> g(g_scratchMu);
    if (ordinal < 0 || ordinal >= kMaxOrdinals) throw PgError(PG_ERR_INVALID, "device ordinal beyond 63");
    if (!g_scratchPtr[ordinal]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || ptr == nullptr)
            throw PgError(PG_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratchPtr[ordinal] = ptr;
    }
    return g_scratchPtr[ordinal];
}

}  // namespace

struct pg_device {
    // Lifetime: scenes and films keep their device alive; pg_device_destroy only marks it closed
    // while any of them remains, and the last child to go frees it.
    int children = 0;
    bool closed = false;
    int ordinal = 0;
    int numCUs = 256;
    hipStream_t stream = nullptr;
    std::mutex mu;
    bool injectFault = false;
    std::string platform;
    void* scratch = nullptr;  // this GPU's copy of the fixed traversal scratch (pg_trace_noalloc)
    DevBuf sobol, sobolT, sobolB, sobolInv, sobolDelta;
    // render workspace
    uint32_t capacity = 0;
    DevBuf recL, recY, quad;  // film record ring and per-pixel footprint sums
    DevBuf itemDone;          // paths recorded per work item
    DevBuf mix, hist;          // guiding: mixtures per queue entry, vertex history per path slot
    DevBuf batchGuides, keyWb, keyTheta, groupBuf;  // guided batches: per-item guides, grouped inference
    DevBuf paths, qext0, qext1, qshadow, qmis, qdone, counts, totals, spill, travCounts;
    DevBuf batchScenes, batchRc, batchOffsets;  // work-item tables of a batched render
    void* status = nullptr;   // pinned host copy of the pool counters and item progress
    size_t spillBytes = 0;
    std::vector<hipEvent_t> events;
    size_t eventsUsed = 0;

    hipEvent_t event() {
        if (eventsUsed == events.size()) {
            hipEvent_t e;
            PG_HIP(hipEventCreate(&e));
            events.push_back(e);
        }
        return events[eventsUsed++];
    }
    void checkFault() {
        if (injectFault) {
            injectFault = false;
            throw PgError(PG_ERR_DEVICE, "injected device fault");
        }
    }
};

struct pg_scene {
    pg_device* dev = nullptr;
    std::vector<float> triVerts;   // 9 floats per triangle: world space, or a moving instance's local space
    std::vector<uint32_t> triMat;  // material per original triangle
    uint32_t numStatic = 0;        // triangles [0, numStatic) are static, in world space
    std::vector<uint32_t> instFirst;    // moving instance k: triangles [instFirst[k], instFirst[k + 1])
    std::vector<pg::DevInstance> inst;  // host copy of the instance table (root, keys, swept box)
    std::vector<float> instBox;         // 6 floats per instance: its mesh's local bounds
    DevBuf instBuf;
    float staticLo[3] = {0, 0, 0}, staticHi[3] = {0, 0, 0};  // static triangles and light quads
    pg::WideBvh bvh;
    pg::BvhParams bvhParams;  // of the current build
    uint32_t builder = PG_BVH_HOST;
    pg::DevScene ds{};
    DevBuf tris, triMatLeaf, nodes, mats, quads, lights, points, triUV, textures, texels, matTex;
    pg::LaunchTune tune;            // launch settings of this scene's renders (pg_autotune)
    uint32_t stackNeed = 1;         // deepest traversal stack a lane can need
    uint32_t spillPerLane = 1;      // of them beyond the LDS part (tune.ldsStack)
    float lo[3] = {0, 0, 0}, invExt[3] = {1, 1, 1};  // scene box (triangles and light quads) -> [0,1]^3
};

struct pg_film {
    pg_device* dev = nullptr;
    uint32_t width = 0, height = 0;
    DevBuf rgbw32, rgbw64;
};

struct pg_guide {
    pg_device* dev = nullptr;
    pg_guide_params p{};
    DevBuf theta, m, v, wb;                  // FP32 masters, Adam moments, BF16 copy
    DevBuf scratch, partial, tiles, csum, loss, grads;
    DevBuf trainRec, trainCount;              // training records of the current render
    const void* tilesFor = nullptr;           // scratch pointer the uploaded tile table refers to
    uint32_t tilesNpad = 0;
    uint32_t step = 0;                        // Adam steps taken
    double recordsPerPath = 3.0;              // training records a training path gave last render
};

namespace {

std::mutex g_lifetime;

void freeDevice(pg_device* dev) {
    (void)hipSetDevice(dev->ordinal);
    (void)hipStreamSynchronize(dev->stream);
    for (auto e : dev->events) (void)hipEventDestroy(e);
    if (dev->status) (void)hipHostFree(dev->status);
    (void)hipStreamDestroy(dev->stream);
    delete dev;
}
void addChild(pg_device* dev) {
    std::lock_guard<std::mutex> g(g_lifetime);
    ++dev->children;
}
void releaseChild(pg_device* dev) {
    std::lock_guard<std::mutex> g(g_lifetime);
    if (--dev->children == 0 && dev->closed) freeDevice(dev);
}

// Guide initialisation: uniform(+-sqrt(6 / (in + out))) weights from a splitmix64 stream, zero
// biases, zero padding (the always-zero input features 135, 138, 139, 142, 143 and output rows
// 34, 35, 38, 39, 42, 43, 46, 47).
std::vector<float> guideInitParams(uint32_t seed) {
    std::vector<float> t(pg::kGParamCount, 0.0f);
    uint64_t state = 0x9E3779B97F4A7C15ull * (seed + 1);
    auto next = [&] {
        uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return static_cast<double>(z >> 11) * 0x1p-53;
    };
    const int outs[5] = {64, 64, 64, 64, 48}, ins[5] = {144, 64, 64, 64, 64};
    size_t k = 0;
    for (int l = 0; l < 5; ++l) {
        const double lim = std::sqrt(6.0 / (ins[l] + outs[l]));
        for (int o = 0; o < outs[l]; ++o)
            for (int i = 0; i < ins[l]; ++i) {
                const double u = (2.0 * next() - 1.0) * lim;
                const bool padIn = i == 135 || i == 138 || i == 139 || i == 142 || i == 143;
                const bool padOut = o >= 32 && (o & 3) >= 2;
                const bool pad = (l == 0 && padIn) || (l == 4 && padOut);
                t[k++] = pad ? 0.0f : static_cast<float>(u);
            }
        k += outs[l];  // biases stay zero
    }
    return t;
}

// One gradient evaluation over n device records; with update, one Adam step.
void guideStep(pg_guide* g, const float* rec, uint32_t n, bool update) {
    hipStream_t st = g->dev->stream;
    const pg::GuideBatchLayout b = pg::guideBatchLayout(n);
    g->scratch.reserve(pg::guideTrainScratchBytes(b));
    g->partial.reserve(static_cast<size_t>(b.chunks) * pg::kGParamCount * 4);
    if (g->tilesFor != g->scratch.p || g->tilesNpad != b.npad) {
        std::vector<pg::WGradTile> tiles(pg::kGTiles);
        pg::buildGuideTiles(g->scratch.p, b, tiles.data());
        g->tiles.reserve(tiles.size() * sizeof(pg::WGradTile));
        PG_HIP(hipStreamSynchronize(st));
        PG_HIP(hipMemcpy(g->tiles.p, tiles.data(), tiles.size() * sizeof(pg::WGradTile), hipMemcpyHostToDevice));
        g->tilesFor = g->scratch.p, g->tilesNpad = b.npad;
    }
    pg::guideConfigureLds();
    PG_HIP(hipMemsetAsync(g->loss.p, 0, 4, st));
    pg::launchGuideGradients(st, rec, n, b, g->wb.as<uint16_t>(), g->theta.as<float>(), g->scratch.p, g->csum.as<float>(),
                             g->partial.as<float>(), g->loss.as<float>(), g->tiles.as<pg::WGradTile>(),
                             static_cast<uint32_t>(g->dev->numCUs), g->p.mis_fraction < 0.0f);
    pg::AdamParams ap{g->p.lr, g->p.beta1, g->p.beta2, g->p.eps, 1.0f, 1.0f, update ? 1u : 0u};
    if (update) {
        ++g->step;
        ap.c1 = 1.0f - std::pow(g->p.beta1, static_cast<float>(g->step));
        ap.c2 = 1.0f - std::pow(g->p.beta2, static_cast<float>(g->step));
    }
    pg::launchGuideAdam(st, g->partial.as<float>(), b.chunks, g->theta.as<float>(), g->m.as<float>(), g->v.as<float>(),
                        g->wb.as<uint16_t>(), g->grads.as<float>(), ap);
    PG_HIP(hipGetLastError());
}

void guideSetLds() { pg::guideConfigureLds(); }

// ---- moving instances on the host ----------------------------------------------------------
// The device's instanceAt in double: keys seg and seg + 1 at fraction f (linear scale, skew,
// shift and translation; slerp of the unit quaternions, normalised lerp above cosine 0.9995),
// M = parent * T R S.
void motionMatrix(const pg::DevInstance& I, uint32_t seg, double f, double m[12]) {
    const float* a = I.srt[seg];
    const float* b = I.srt[I.keys > 1 ? seg + 1 : seg];
    double k[16];
    for (int i = 0; i < 16; ++i) k[i] = a[i] * (1.0 - f) + b[i] * f;
    double q0[4], q1[4], q[4], c = 0.0;
    for (int i = 0; i < 4; ++i) q0[i] = a[9 + i], q1[i] = b[9 + i], c += q0[i] * q1[i];
    if (c < 0.0) {
        c = -c;
        for (double& v : q1) v = -v;
    }
    if (c > 0.9995) {
        double n = 0.0;
        for (int i = 0; i < 4; ++i) q[i] = q0[i] * (1.0 - f) + q1[i] * f, n += q[i] * q[i];
        for (double& v : q) v /= std::sqrt(n);
    } else {
        double pp[4], n = 0.0;
        for (int i = 0; i < 4; ++i) pp[i] = q1[i] - c * q0[i], n += pp[i] * pp[i];
        const double phi = f * std::acos(c), cs = std::cos(phi), sn = std::sin(phi) / std::sqrt(n);
        for (int i = 0; i < 4; ++i) q[i] = cs * q0[i] + sn * pp[i];
    }
    const double w = q[0], x = q[1], y = q[2], z = q[3];
    const double r[9] = {1 - 2 * (y * y + z * z), 2 * (x * y - w * z),     2 * (x * z + w * y),
                         2 * (x * y + w * z),     1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                         2 * (x * z - w * y),     2 * (y * z + w * x),     1 - 2 * (x * x + y * y)};
    double l[12];
    for (int i = 0; i < 3; ++i) {
        l[4 * i] = r[3 * i] * k[0];
        l[4 * i + 1] = r[3 * i] * k[3] + r[3 * i + 1] * k[1];
        l[4 * i + 2] = r[3 * i] * k[4] + r[3 * i + 1] * k[5] + r[3 * i + 2] * k[2];
        l[4 * i + 3] = r[3 * i] * k[6] + r[3 * i + 1] * k[7] + r[3 * i + 2] * k[8] + k[13 + i];
    }
    const float* P = I.parent;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 4; ++j)
            m[4 * i + j] = P[4 * i] * l[j] + P[4 * i + 1] * l[4 + j] + P[4 * i + 2] * l[8 + j] + (j == 3 ? P[4 * i + 3] : 0.0);
}

// World bounds of instance I over the times in [0, 1] (the range of a sample's time) at which it is
// visible: per key segment, the local box's corners at 65 equally spaced fractions, grown by half
// a step times a bound on the corners' speed, |d(world)/df| <= |P|_F (|dT| + 2 theta max|S p| +
// |dS p|) with theta the quaternions' angle (the rotation turns at 2 theta per unit fraction).
// Returns false if no such time exists.
bool sweptBox(const pg::DevInstance& I, const float* box, float lo[3], float hi[3]) {
    constexpr int kSteps = 64;
    double blo[3] = {INFINITY, INFINITY, INFINITY}, bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
    double pn = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) pn += static_cast<double>(I.parent[4 * i + j]) * I.parent[4 * i + j];
    pn = std::sqrt(pn);
    bool any = false;
    const uint32_t segs = I.keys > 1 ? I.keys - 1 : 1;
    for (uint32_t seg = 0; seg < segs; ++seg) {
        double fa = 0.0, fb = 0.0;
        if (I.keys > 1) {
            const double span = (static_cast<double>(I.t1) - I.t0) / (I.keys - 1);
            const double ta = I.t0 + seg * span, tb = ta + span;
            const double wa = std::max({ta, 0.0, static_cast<double>(I.t0)});
            const double wb = std::min({tb, 1.0, static_cast<double>(I.t1)});
            if (wa > wb) continue;
            fa = (wa - ta) / span, fb = (wb - ta) / span;
        }
        any = true;
        const float* a = I.srt[seg];
        const float* b = I.srt[I.keys > 1 ? seg + 1 : seg];
        double dq = 0.0;
        for (int i = 0; i < 4; ++i) dq += static_cast<double>(a[9 + i]) * b[9 + i];
        const double theta = std::acos(std::min(1.0, std::fabs(dq))) * 1.05;
        const double dt = std::sqrt((b[13] - a[13]) * (b[13] - a[13]) + (b[14] - a[14]) * (b[14] - a[14]) +
                                    (b[15] - a[15]) * (b[15] - a[15]));
        double speed = 0.0, reach = 0.0;
        for (int c = 0; c < 8; ++c) {
            const double p[3] = {box[(c & 1) ? 3 : 0], box[(c & 2) ? 4 : 1], box[(c & 4) ? 5 : 2]};
            auto sp = [&](const float* k, double out[3]) {
                out[0] = k[0] * p[0] + k[3] * p[1] + k[4] * p[2] + k[6];
                out[1] = k[1] * p[1] + k[5] * p[2] + k[7];
                out[2] = k[2] * p[2] + k[8];
            };
            double sa[3], sb[3];
            sp(a, sa), sp(b, sb);
            const double na = std::sqrt(sa[0] * sa[0] + sa[1] * sa[1] + sa[2] * sa[2]);
            const double nb = std::sqrt(sb[0] * sb[0] + sb[1] * sb[1] + sb[2] * sb[2]);
            const double nd = std::sqrt((sb[0] - sa[0]) * (sb[0] - sa[0]) + (sb[1] - sa[1]) * (sb[1] - sa[1]) +
                                        (sb[2] - sa[2]) * (sb[2] - sa[2]));
            speed = std::max(speed, pn * (dt + 2.0 * theta * std::max(na, nb) + nd));
        }
        const double step = (fb - fa) / kSteps;
        for (int i = 0; i <= kSteps; ++i) {
            double m[12];
            motionMatrix(I, seg, fa + step * i, m);
            for (int c = 0; c < 8; ++c) {
                const double p[3] = {box[(c & 1) ? 3 : 0], box[(c & 2) ? 4 : 1], box[(c & 4) ? 5 : 2]};
                for (int r = 0; r < 3; ++r) {
                    const double v = m[4 * r] * p[0] + m[4 * r + 1] * p[1] + m[4 * r + 2] * p[2] + m[4 * r + 3];
                    blo[r] = std::min(blo[r], v), bhi[r] = std::max(bhi[r], v);
                    reach = std::max(reach, std::fabs(v));
                }
            }
        }
        const double pad = speed * step * 0.5 + 1e-5 * reach;
        for (int r = 0; r < 3; ++r) blo[r] -= pad, bhi[r] += pad;
    }
    for (int r = 0; r < 3; ++r) lo[r] = static_cast<float>(blo[r]), hi[r] = static_cast<float>(bhi[r]);
    return any;
}

// Scene box for the guide's position inputs: static triangles and quads, and every visible
// instance's swept bounds.
void updateSceneBox(pg_scene* s) {
    float blo[3], bhi[3];
    for (int k = 0; k < 3; ++k) blo[k] = s->staticLo[k], bhi[k] = s->staticHi[k];
    for (const pg::DevInstance& I : s->inst)
        if (I.visible)
            for (int k = 0; k < 3; ++k) blo[k] = std::min(blo[k], I.lo[k]), bhi[k] = std::max(bhi[k], I.hi[k]);
    for (int k = 0; k < 3; ++k) {
        if (!(bhi[k] > blo[k])) blo[k] = 0.0f, bhi[k] = 1.0f;
        s->lo[k] = blo[k];
        s->invExt[k] = 1.0f / (bhi[k] - blo[k]);
    }
}

// An instance drawn with its local coordinates as world coordinates until its motion is set.
pg::DevInstance fixedInstance(const float* box) {
    pg::DevInstance I{};
    I.keys = 1, I.visible = 1;
    I.t0 = -FLT_MAX, I.t1 = FLT_MAX, I.rate = 0.0f;
    const float id[16] = {1, 1, 1, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0};
    std::memcpy(I.srt[0], id, sizeof(id));
    const float par[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    std::memcpy(I.parent, par, sizeof(par));
    for (int k = 0; k < 3; ++k) I.lo[k] = box[k], I.hi[k] = box[3 + k];
    return I;
}

// Moving instances: one BLAS per instance (host builder, the static tree's width and leaf size)
// over its local-space triangles, appended to the static tree's node and triangle arrays with
// node and triangle references rebased; the instance's leaf triangles keep their original
// triangle number (tie rule, texture coordinates) and carry the instance index + 1 in v1.w.
void appendInstances(pg_scene* s, const pg::BvhParams& bp, hipStream_t st) {
    const uint32_t W = bp.width;
    const size_t nodeFloats = 7 * static_cast<size_t>(W), nodeBytes = nodeFloats * 4;
    const uint32_t staticNodes = s->numStatic ? s->bvh.numNodes : 0;
    std::vector<float> nodes, tris;
    std::vector<uint32_t> mats;
    uint32_t nodeBase = staticNodes, triBase = s->numStatic;
    uint32_t maxDepth = s->numStatic ? s->bvh.maxDepth : 0, leaves = s->numStatic ? s->bvh.numLeaves : 0;
    for (size_t k = 0; k < s->inst.size(); ++k) {
        const uint32_t first = s->instFirst[k], cnt = s->instFirst[k + 1] - first;
        pg::WideBvh b;
        pg::buildWideBvh(&s->triVerts[9 * static_cast<size_t>(first)], cnt, bp, b);
        require(b.width == W, "instance BVH width differs from the scene's");
        for (uint32_t n = 0; n < b.numNodes; ++n) {
            uint32_t* child = reinterpret_cast<uint32_t*>(&b.nodes[n * nodeFloats + 6 * W]);
            for (uint32_t c = 0; c < W; ++c) {
                const uint32_t ref = child[c];
                if (ref == pg::kInvalidChild) continue;
                if (ref & pg::kLeafBit) {
                    const uint32_t f = (ref & pg::kLeafFirstMask) + triBase;
                    require(f <= pg::kLeafFirstMask, "too many triangles for the leaf references");
                    child[c] = (ref & ~pg::kLeafFirstMask) | f;
                } else {
                    child[c] = ref + nodeBase;
                }
            }
        }
        const uint32_t tag = static_cast<uint32_t>(k) + 1u;
        for (uint32_t i = 0; i < cnt; ++i) {
            float* o = &b.tris[12 * static_cast<size_t>(i)];
            uint32_t prim;
            std::memcpy(&prim, &o[3], 4);
            prim += first;
            std::memcpy(&o[3], &prim, 4);
            std::memcpy(&o[7], &tag, 4);
            mats.push_back(s->triMat[first + b.order[i]]);
        }
        nodes.insert(nodes.end(), b.nodes.begin(), b.nodes.end());
        tris.insert(tris.end(), b.tris.begin(), b.tris.end());
        s->inst[k].root = nodeBase;
        nodeBase += b.numNodes;
        triBase += cnt;
        maxDepth = std::max(maxDepth, b.maxDepth);
        leaves += b.numLeaves;
    }
    DevBuf nN, nT, nM;
    nN.reserve(static_cast<size_t>(nodeBase) * nodeBytes);
    nT.reserve(static_cast<size_t>(triBase) * 48);
    nM.reserve(static_cast<size_t>(triBase) * 4);
    if (s->numStatic) {
        PG_HIP(hipMemcpyAsync(nN.p, s->nodes.p, static_cast<size_t>(staticNodes) * nodeBytes, hipMemcpyDeviceToDevice, st));
        PG_HIP(hipMemcpyAsync(nT.p, s->tris.p, static_cast<size_t>(s->numStatic) * 48, hipMemcpyDeviceToDevice, st));
        PG_HIP(hipMemcpyAsync(nM.p, s->triMatLeaf.p, static_cast<size_t>(s->numStatic) * 4, hipMemcpyDeviceToDevice, st));
    }
    PG_HIP(hipMemcpyAsync(static_cast<char*>(nN.p) + static_cast<size_t>(staticNodes) * nodeBytes, nodes.data(),
                          nodes.size() * 4, hipMemcpyHostToDevice, st));
    PG_HIP(hipMemcpyAsync(static_cast<char*>(nT.p) + static_cast<size_t>(s->numStatic) * 48, tris.data(), tris.size() * 4,
                          hipMemcpyHostToDevice, st));
    PG_HIP(hipMemcpyAsync(static_cast<char*>(nM.p) + static_cast<size_t>(s->numStatic) * 4, mats.data(), mats.size() * 4,
                          hipMemcpyHostToDevice, st));
    PG_HIP(hipStreamSynchronize(st));
    std::swap(s->nodes.p, nN.p), std::swap(s->nodes.bytes, nN.bytes);
    std::swap(s->tris.p, nT.p), std::swap(s->tris.bytes, nT.bytes);
    std::swap(s->triMatLeaf.p, nM.p), std::swap(s->triMatLeaf.bytes, nM.bytes);
    s->bvh.width = W;
    s->bvh.numNodes = nodeBase;
    s->bvh.numLeaves = leaves;
    s->bvh.maxDepth = maxDepth;
    upload(s->instBuf, s->inst.data(), s->inst.size() * sizeof(pg::DevInstance), st);
    PG_HIP(hipStreamSynchronize(st));
}

void buildBvh(pg_scene* s, const pg_bvh_params* params) {
    pg::BvhParams bp;
    if (params) {
        bp.width = params->width;
        bp.leafSize = params->leaf_size;
        bp.bins = params->sah_bins;
    }
    const uint32_t nt = s->numStatic;  // the static BVH; moving instances get BLASes of their own
    hipStream_t st = s->dev->stream;
    const bool onDevice = params && params->builder == PG_BVH_DEVICE && nt > 0;
    require(!params || params->builder == PG_BVH_HOST || params->builder == PG_BVH_DEVICE, "unknown BVH builder");
    if (onDevice) {  // binned SAH on the GPU; host keeps only the shape
        require(bp.width == 4 || bp.width == 8, "BVH width must be 4 or 8");
        require(bp.leafSize >= 1 && bp.leafSize <= pg::kMaxLeafSize, "leaf size must be 1..8");
        DevBuf verts, mats;
        upload(verts, s->triVerts.data(), static_cast<size_t>(nt) * 36, st);
        upload(mats, s->triMat.data(), static_cast<size_t>(nt) * 4, st);
        pg::DeviceBvh db;
        db.alloc = [&](pg::DeviceBvh::Buffer b, size_t bytes) -> void* {
            DevBuf& d = b == pg::DeviceBvh::kNodes ? s->nodes : b == pg::DeviceBvh::kTris ? s->tris : s->triMatLeaf;
            d.reserve(bytes);
            return d.p;
        };
        try {
            pg::buildWideBvhDevice(st, verts.as<float>(), mats.as<uint32_t>(), nt, bp, db);
        } catch (const std::runtime_error& e) {
            throw PgError(PG_ERR_DEVICE, e.what());
        }
        s->bvh = pg::WideBvh{};
        s->bvh.width = db.width;
        s->bvh.numNodes = db.numNodes;
        s->bvh.numLeaves = db.numLeaves;
        s->bvh.maxDepth = db.maxDepth;
        s->bvh.pad = db.pad;
    } else {
        pg::buildWideBvh(s->triVerts.data(), nt, bp, s->bvh);
        std::vector<uint32_t> leafMat(std::max<uint32_t>(nt, 1));
        for (uint32_t i = 0; i < nt; ++i) leafMat[i] = s->triMat[s->bvh.order[i]];
        upload(s->tris, s->bvh.tris.data(), s->bvh.tris.size() * 4, st);
        upload(s->triMatLeaf, leafMat.data(), leafMat.size() * 4, st);
        upload(s->nodes, s->bvh.nodes.data(), s->bvh.nodes.size() * 4, st);
        PG_HIP(hipStreamSynchronize(st));
        s->bvh.nodes.clear(), s->bvh.nodes.shrink_to_fit();  // the device copy is the one in use
        s->bvh.tris.clear(), s->bvh.tris.shrink_to_fit();
    }
    if (nt == 0) s->bvh.numNodes = s->bvh.numLeaves = s->bvh.maxDepth = 0;
    if (!s->inst.empty()) appendInstances(s, bp, st);
    s->bvhParams = bp;
    s->builder = onDevice ? PG_BVH_DEVICE : PG_BVH_HOST;
    s->ds.tris = s->tris.as<float4>();
    s->ds.triMat = s->triMatLeaf.as<uint32_t>();
    s->ds.nodes = s->nodes.p;
    s->ds.numTris = static_cast<uint32_t>(s->triMat.size());
    s->ds.numNodes = s->bvh.numNodes;
    s->ds.width = s->bvh.width;
    s->ds.hasBvh = nt > 0 ? 1u : 0u;
    s->ds.inst = s->inst.empty() ? nullptr : s->instBuf.as<pg::DevInstance>();
    s->ds.numInst = static_cast<uint32_t>(s->inst.size());
    s->stackNeed = (s->bvh.width - 1) * s->bvh.maxDepth + 2;
    s->spillPerLane = std::max<uint32_t>(s->stackNeed > s->tune.ldsStack ? s->stackNeed - s->tune.ldsStack : 0u, 1u);
}

pg::LaunchTune tuneFrom(const pg_launch_config& c) {
    require(c.lds_stack >= 4 && c.lds_stack <= pg::kLdsStackMax, "lds_stack must be 4..16");
    require(c.trav_waves <= 64, "trav_waves must be 0..64");
    require(c.shade_block == 128 || c.shade_block == 256, "shade_block must be 128 or 256");
    require(c.infer_waves == 4 || c.infer_waves == 8 || c.infer_waves == 16, "infer_waves must be 4, 8 or 16");
    return pg::LaunchTune{c.lds_stack, c.trav_waves, c.shade_block, c.infer_waves};
}
pg_launch_config configOf(const pg::LaunchTune& t) {
    return pg_launch_config{t.ldsStack, t.travWaves, t.shadeBlock, t.inferWaves};
}
void setTune(pg_scene* s, const pg::LaunchTune& t) {
    s->tune = t;
    s->spillPerLane = std::max<uint32_t>(s->stackNeed > t.ldsStack ? s->stackNeed - t.ldsStack : 0u, 1u);
}

uint32_t log2Ceil(uint32_t v) {
    uint32_t m = 0;
    while ((1u << m) < v) ++m;
    return m;
}

pg::RenderConsts makeConsts(const pg_render_params& p) {
    require(p.width > 0 && p.height > 0, "film size must be positive");
    require(p.rect[2] > 0 && p.rect[3] > 0 && p.rect[0] + p.rect[2] <= p.width && p.rect[1] + p.rect[3] <= p.height,
            "render rect must lie inside the film");
    require(p.spp_end > p.spp_begin, "empty sample range");
    require(p.max_depth <= 127, "max_depth above 127 would exceed the 1024 Sobol dimensions");
    pg::RenderConsts rc{};
    rc.width = p.width, rc.height = p.height;
    uint32_t spp = p.spp ? p.spp : 1;
    spp = 1u << log2Ceil(spp);  // ScrambledSobolSampler.cpp:43-50
    require(p.spp_end <= spp, "sample range beyond samples per pixel");
    rc.spp = spp;
    rc.rectX = p.rect[0], rc.rectY = p.rect[1], rc.rectW = p.rect[2], rc.rectH = p.rect[3];
    rc.sppBegin = p.spp_begin, rc.sppCount = p.spp_end - p.spp_begin;
    rc.maxDepth = p.max_depth, rc.minRR = p.min_rr, rc.flags = p.flags;
    rc.log2Res = log2Ceil(std::max(p.width, p.height));  // ScrambledSobolSampler.cpp:66-72
    require(rc.log2Res <= pg::kMaxLog2Res, "film larger than 65536");
    const uint64_t paths = static_cast<uint64_t>(rc.rectW) * rc.rectH * rc.sppCount;
    require(paths < (1ull << 32), "more than 2^32 paths in one call: split the sample range");
    require(static_cast<uint64_t>(rc.spp) << (2 * rc.log2Res) < (1ull << 52),
            "sample index beyond the 52 Sobol matrix columns");
    std::memcpy(rc.ndc, p.ndc, sizeof(rc.ndc));
    require(p.filter <= pg::kFilterGaussian, "unknown film filter");
    require(p.filter != pg::kFilterGaussian || p.filter_alpha > 0.0f, "gaussian filter needs alpha > 0");
    rc.filter = p.filter;
    rc.negAlpha = -p.filter_alpha;
    rc.sub = p.filter == pg::kFilterGaussian ? std::exp(-p.filter_alpha) : 0.0f;
    return rc;
}

void ensureWorkspace(pg_device* dev, uint32_t cap, size_t spillBytes) {
    if (cap > dev->capacity) {
        const size_t P = cap;
        dev->paths.reserve(P * sizeof(pg::PathRec));
        dev->qext0.reserve(P * 4), dev->qext1.reserve(P * 4), dev->qdone.reserve(P * 4);
        dev->qshadow.reserve(2 * P * sizeof(pg::ShadowRec)), dev->qmis.reserve(P * sizeof(pg::MisRec));
        dev->capacity = cap;
    }
    dev->counts.reserve(pg::kCountsBytes);
    dev->totals.reserve(pg::kTotNum * 8);
    dev->travCounts.reserve(64);
    dev->spill.reserve(std::max<size_t>(spillBytes, 64));
}

pg::PathState pathState(pg_device* dev) {
    return pg::PathState{dev->paths.as<pg::PathRec>()};
}
pg::Queues queues(pg_device* dev) {
    return pg::Queues{{dev->qext0.as<uint32_t>(), dev->qext1.as<uint32_t>()},
                      dev->qshadow.as<pg::ShadowRec>(),
                      dev->qmis.as<pg::MisRec>(),
                      dev->qdone.as<uint32_t>(),
                      dev->counts.as<uint32_t>(),
                      reinterpret_cast<pg::PoolPaths*>(dev->counts.as<char>() + pg::kCntNum * 4)};
}

uint32_t traversalLanes(pg_device* dev, uint64_t work) {
    // the most lanes a traversal launch can take; each launch trims its grid to what the device
    // holds of that kernel at once
    const uint64_t full = static_cast<uint64_t>(dev->numCUs) * pg::kTravWavesPerCuMax * 64;
    const uint64_t need = (work + 63) / 64 * 64;
    return static_cast<uint32_t>(std::max<uint64_t>(64, std::min(full, need)));
}
uint32_t shadeThreads(pg_device* dev, uint64_t work) {
    const uint64_t full = static_cast<uint64_t>(dev->numCUs) * 8 * 256;
    const uint64_t need = (work + 255) / 256 * 256;
    return static_cast<uint32_t>(std::max<uint64_t>(256, std::min(full, need)));
}

enum Stage { kRaygen, kExtend, kShade, kShadow, kMis, kFilm, kGuide, kTrain, kStages };

// Per-stage device time: a pair of events around every launch of a stage (off: no events).
struct StageTimer {
    pg_device* dev;
    hipStream_t st;
    bool on;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> marks[kStages];

    template <typename F>
    void operator()(Stage sg, F&& launch) {
        if (!on) {
            launch();
            return;
        }
        hipEvent_t a = dev->event(), b = dev->event();
        PG_HIP(hipEventRecord(a, st));
        launch();
        PG_HIP(hipEventRecord(b, st));
        marks[sg].push_back({a, b});
    }
    // after the stream has drained
    void collect(pg_stats* stats) const {
        double* ms[kStages] = {&stats->ms_raygen, &stats->ms_extend, &stats->ms_shade, &stats->ms_shadow,
                               &stats->ms_mis,    &stats->ms_film,   &stats->ms_guide, &stats->ms_train};
        for (int sg = 0; sg < kStages; ++sg)
            for (auto& m : marks[sg]) {
                float e = 0.0f;
                PG_HIP(hipEventElapsedTime(&e, m.first, m.second));
                *ms[sg] += e;
            }
    }
};

// Paths whose film records one segment holds; larger sample ranges are rendered in segments
// (whole pixels x a sample sub-range), each resolved into the film when it is done.
constexpr uint64_t kMaxSegmentPaths = 1ull << 25;

uint32_t segmentSpp(const pg::RenderConsts& rc) {
    const uint64_t pixels = static_cast<uint64_t>(rc.rectW) * rc.rectH;
    return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(rc.sppCount, kMaxSegmentPaths / pixels)));
}

// Work item: one scene's whole rect x a sample sub-range (a film segment), with path numbers
// [first, first + count) in its render call.
struct WorkItem {
    uint32_t scene;
    pg::RenderConsts rc;  // rect, the item's sample sub-range, depth, flags
    uint64_t first, count;
};

// Splits a scene's render into segments of rc.segSpp samples per pixel, appended to `items`.
void appendSegments(uint32_t scene, const pg::RenderConsts& rc, std::vector<WorkItem>& items, uint64_t& total) {
    const uint64_t pixels = static_cast<uint64_t>(rc.rectW) * rc.rectH;
    for (uint32_t s0 = rc.sppBegin; s0 < rc.sppBegin + rc.sppCount; s0 += rc.segSpp) {
        WorkItem it{scene, rc, total, 0};
        it.rc.sppBegin = s0;
        it.rc.sppCount = std::min(rc.segSpp, rc.sppBegin + rc.sppCount - s0);
        it.rc.segSpp = it.rc.sppCount;
        it.count = pixels * it.rc.sppCount;
        total += it.count;
        items.push_back(it);
    }
}

struct RunOptions {
    bool batch = false;           // per-item scene and consts from device tables
    pg::RenderConsts rc0{};       // single scene: the call's consts (segments by rc0.segSpp)
    pg_film* const* films = nullptr;  // per scene (entries may be null)
    uint32_t capacity = 1u << 21;
    pg_guide* guide = nullptr;    // single scene
    pg_guide* const* guides = nullptr;  // per scene (guided batch)
    float* pathLHost = nullptr;   // single scene only
    double* latencyMs = nullptr;  // per scene
    pg_stats* stats = nullptr;
    bool countTraversal = false;
    bool timed = false;           // per-stage events (not with graph)
    bool graph = false;           // replay captured groups of iterations
};

// Iterations between two reads of the render's progress (each read waits for the stream).
constexpr uint32_t kCheckEvery = 4;
// Iterations without any progress after which a render is abandoned as stuck.
constexpr uint32_t kStallLimit = 1u << 16;

// Renders work items through one pool of path slots. Every iteration restarts the slots of
// finished paths with the next path numbers (in order), then runs one bounce for every live
// path: extend, (guide inference), shade, shadow, (MIS), record. Film records go to a ring
// indexed by path number; an item is resolved into its film as soon as all its paths are
// recorded, and new paths may only take ring slots that no unresolved item still needs.
void runPaths(pg_device* dev, pg_scene* const* scenes, uint32_t nScenes, const std::vector<WorkItem>& items,
              uint64_t total, const RunOptions& opt) {
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = dev->stream;
    require(!items.empty() && total > 0, "nothing to render");
    require(items.size() <= pg::kMaxBatchItems, "too many film segments in one call: split it");
    uint64_t maxItem = 0;
    uint32_t maxDepth = 0, spillPerLane = 1;
    bool envBlack = true;
    size_t quadBytes = 16;
    for (const WorkItem& it : items) {
        maxItem = std::max(maxItem, it.count);
        maxDepth = std::max(maxDepth, it.rc.maxDepth);
        quadBytes = std::max<size_t>(quadBytes, static_cast<size_t>(it.rc.rectW) * it.rc.rectH *
                                                    (it.rc.filter == pg::kFilterBox ? 16 : 36) *
                                                    ((it.rc.flags & pg::kFilmFp64) ? 8 : 4));
    }
    const pg::LaunchTune lt = scenes[0]->tune;  // a batch launches with its first scene's settings
    for (uint32_t i = 0; i < nScenes; ++i) {
        const uint32_t need = scenes[i]->stackNeed;
        spillPerLane = std::max(spillPerLane, need > lt.ldsStack ? need - lt.ldsStack : 1u);
        const float* env = scenes[i]->ds.env;
        envBlack = envBlack && !(env[0] > 0.0f || env[1] > 0.0f || env[2] > 0.0f);
    }
    const uint32_t cap = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(opt.capacity, total)));
    const uint32_t lanes = traversalLanes(dev, cap);
    const uint32_t threads = shadeThreads(dev, cap);
    ensureWorkspace(dev, cap, static_cast<size_t>(lanes) * spillPerLane * 8);
    // Ring: the largest item plus room for the spread of live path numbers (paths finish out of
    // order). A ring too small only throttles the pool; one item must always fit.
    const uint64_t ringSize =
        opt.pathLHost ? total : std::min(total, 2 * maxItem + 8 * static_cast<uint64_t>(cap));
    dev->recL.reserve(ringSize * 16), dev->recY.reserve(ringSize * 4);
    dev->quad.reserve(quadBytes);
    dev->itemDone.reserve(std::max<size_t>(items.size(), kStatusItems) * 4);
    PG_HIP(hipMemsetAsync(dev->itemDone.p, 0, items.size() * 4, st));
    const pg::RecordRing ring{dev->recL.as<float4>(), dev->recY.as<float>(), ringSize, dev->itemDone.as<uint32_t>()};

    pg::BatchTable bt{};
    std::vector<pg::DevScene> tabScene;
    std::vector<pg::RenderConsts> tabRc;
    std::vector<uint64_t> tabOffsets;
    if (opt.batch) {
        tabScene.resize(items.size()), tabRc.resize(items.size()), tabOffsets.resize(items.size() + 1);
        for (size_t k = 0; k < items.size(); ++k) {
            tabScene[k] = scenes[items[k].scene]->ds;
            tabRc[k] = items[k].rc;
            tabOffsets[k] = items[k].first;
        }
        tabOffsets[items.size()] = total;
        upload(dev->batchScenes, tabScene.data(), tabScene.size() * sizeof(pg::DevScene), st);
        upload(dev->batchRc, tabRc.data(), tabRc.size() * sizeof(pg::RenderConsts), st);
        upload(dev->batchOffsets, tabOffsets.data(), tabOffsets.size() * 8, st);
        bt = pg::BatchTable{dev->batchScenes.as<pg::DevScene>(), dev->batchRc.as<pg::RenderConsts>(),
                            dev->batchOffsets.as<uint64_t>(), static_cast<uint32_t>(items.size())};
        for (uint32_t i = 0; i < nScenes; ++i) bt.motion |= scenes[i]->ds.numInst != 0 ? 1u : 0u;
    }
    const pg::BatchTable* btp = opt.batch ? &bt : nullptr;
    const pg::DevScene& ds0 = scenes[0]->ds;  // the batch's BVH width; per-path scenes come from bt
    bool anyTextured = false;
    for (uint32_t i = 0; i < nScenes; ++i) anyTextured = anyTextured || scenes[i]->ds.textured != 0;
    const pg::RenderConsts& rc0 = opt.batch ? items[0].rc : opt.rc0;

    const pg::PathState ps = pathState(dev);
    const pg::Queues q = queues(dev);
    // Guiding: one guide (single scene) or one per scene (guided batch; scenes may share a guide).
    // Mixtures per extension-queue entry, per-slot vertex history of training paths, training
    // records per guide. Only vertices that found light give records, so the share of a guide's
    // paths that train follows its record yield in the previous render: every stride-th path.
    pg::GuideShade gs{};
    pg::GuideRecordOut tr{};
    std::vector<pg_guide*> keys;                       // distinct guides of the call
    std::vector<uint32_t> itemKey(items.size(), 0), keyStride;
    auto guideOf = [&](uint32_t scene) { return opt.guides ? opt.guides[scene] : opt.guide; };
    pg::GroupBuffers gb{};
    const uint16_t* const* keyWb = nullptr;
    const float* const* keyTheta = nullptr;
    if (opt.guide || opt.guides) {
        std::vector<uint64_t> keyPaths;
        for (size_t k = 0; k < items.size(); ++k) {
            pg_guide* g = guideOf(items[k].scene);
            require(g && g->dev == dev, "every scene of a guided render needs a guide on the render's device");
            const auto at = std::find(keys.begin(), keys.end(), g);
            itemKey[k] = static_cast<uint32_t>(at - keys.begin());
            if (at == keys.end()) keys.push_back(g), keyPaths.push_back(0);
            keyPaths[itemKey[k]] += items[k].count;
        }
        require(keys.size() <= pg::kMaxBatchGuides, "too many distinct guides in one batch");
        gs.enabled = 1;
        dev->mix.reserve(static_cast<size_t>(cap) * pg::kGMixFloatsHost * 4);
        gs.mix = dev->mix.as<float>();
        keyStride.assign(keys.size(), 1);
        for (size_t k = 0; k < keys.size(); ++k) {
            pg_guide* g = keys[k];
            if (g->p.train_records == 0) continue;
            gs.train = 1;
            const double want = static_cast<double>(keyPaths[k]) * g->recordsPerPath / g->p.train_records;
            keyStride[k] = static_cast<uint32_t>(std::min(std::max(want, 1.0), 1e9));
            g->trainRec.reserve(static_cast<size_t>(g->p.train_records) * 64);
            g->trainCount.reserve(16);
            PG_HIP(hipMemsetAsync(g->trainCount.p, 0, 4, st));
        }
        if (gs.train) {
            gs.histDepth = std::max<uint32_t>(1, std::min<uint32_t>(maxDepth, 255));
            dev->hist.reserve(static_cast<size_t>(cap) * gs.histDepth * 6 * 16);
            gs.hist = dev->hist.as<float4>();
        }
        auto records = [](pg_guide* g) {
            return g->p.train_records ? pg::GuideRecordOut{g->trainRec.as<float4>(), g->trainCount.as<uint32_t>(), g->p.train_records}
                                      : pg::GuideRecordOut{nullptr, nullptr, 0};
        };
        if (!opt.batch) {
            const pg_scene* s = scenes[0];
            gs.alpha = keys[0]->p.mis_fraction;
            gs.lo = make_float3(s->lo[0], s->lo[1], s->lo[2]);
            gs.invExt = make_float3(s->invExt[0], s->invExt[1], s->invExt[2]);
            gs.trainStride = keyStride[0];
            gs.trainPhase = keyStride[0] / 2;
            tr = records(keys[0]);
        } else {
            std::vector<pg::ItemGuide> tab(items.size());
            for (size_t k = 0; k < items.size(); ++k) {
                pg_guide* g = keys[itemKey[k]];
                const pg_scene* s = scenes[items[k].scene];
                const uint32_t stride = keyStride[itemKey[k]];
                tab[k] = pg::ItemGuide{g->wb.as<uint16_t>(), g->theta.as<float>(), records(g),
                                       make_float3(s->lo[0], s->lo[1], s->lo[2]),
                                       make_float3(s->invExt[0], s->invExt[1], s->invExt[2]),
                                       g->p.mis_fraction, itemKey[k], stride, stride / 2};
            }
            upload(dev->batchGuides, tab.data(), tab.size() * sizeof(pg::ItemGuide), st);
            bt.guide = dev->batchGuides.as<pg::ItemGuide>();
            std::vector<const uint16_t*> wbs;
            std::vector<const float*> thetas;
            for (pg_guide* g : keys) wbs.push_back(g->wb.as<uint16_t>()), thetas.push_back(g->theta.as<float>());
            upload(dev->keyWb, wbs.data(), wbs.size() * sizeof(void*), st);
            upload(dev->keyTheta, thetas.data(), thetas.size() * sizeof(void*), st);
            keyWb = dev->keyWb.as<const uint16_t*>();
            keyTheta = dev->keyTheta.as<const float*>();
            dev->groupBuf.reserve(pg::groupBufferBytes(cap));
            gb = pg::groupBuffers(dev->groupBuf.p, cap);
        }
        pg::guideConfigureLds();
    }
    for (uint32_t i = 0; i < nScenes; ++i) {  // FP64 films get their buffer on first use
        pg_film* f = opt.films ? opt.films[i] : nullptr;
        bool fp64 = false;
        for (const WorkItem& it : items) fp64 = fp64 || (it.scene == i && (it.rc.flags & pg::kFilmFp64));
        if (f && fp64 && !f->rgbw64.p) {
            f->rgbw64.reserve(static_cast<size_t>(f->width) * f->height * 32);
            PG_HIP(hipMemsetAsync(f->rgbw64.p, 0, f->rgbw64.bytes, st));
        }
    }
    PG_HIP(hipMemsetAsync(dev->totals.p, 0, pg::kTotNum * 8, st));
    uint32_t* trav = opt.countTraversal ? dev->travCounts.as<uint32_t>() : nullptr;
    if (trav) PG_HIP(hipMemsetAsync(dev->travCounts.p, 0, 64, st));
    dev->eventsUsed = 0;
    StageTimer timeStage{dev, st, opt.timed, {}};
    hipEvent_t start = nullptr;
    std::vector<hipEvent_t> sceneDone(nScenes, nullptr);
    std::vector<size_t> lastItem(nScenes, 0);
    for (size_t k = 0; k < items.size(); ++k) lastItem[items[k].scene] = k;
    if (opt.latencyMs) {
        start = dev->event();
        PG_HIP(hipEventRecord(start, st));
    }

    size_t nextResolve = 0;
    auto resolveItem = [&](size_t k) {
        const WorkItem& it = items[k];
        pg_film* f = opt.films ? opt.films[it.scene] : nullptr;
        if (f) {
            const bool fp64 = it.rc.flags & pg::kFilmFp64;
            timeStage(kFilm, [&] {
                pg::launchResolve(st, it.rc, ring, it.first, dev->quad.p, fp64 ? nullptr : f->rgbw32.as<float>(),
                                  fp64 ? f->rgbw64.as<double>() : nullptr);
            });
        }
        if (opt.latencyMs && lastItem[it.scene] == k) {
            sceneDone[it.scene] = dev->event();
            PG_HIP(hipEventRecord(sceneDone[it.scene], st));
        }
    };
    // new paths may take ring slots up to the oldest unresolved item's first path number + ring
    auto limitNow = [&] {
        return nextResolve < items.size() ? std::min(total, items[nextResolve].first + ringSize) : total;
    };

    timeStage(kRaygen, [&] { pg::launchPoolInit(st, q, cap); });
    uint2* spill = dev->spill.as<uint2>();
    // generation limit: written to device memory from pinned host memory after every progress read
    auto* limitHost = reinterpret_cast<uint64_t*>(static_cast<char*>(dev->status) + kStatusBytes);
    auto pushLimit = [&] {
        *limitHost = limitNow();
        PG_HIP(hipMemcpyAsync(&q.pool->limit, limitHost, 8, hipMemcpyHostToDevice, st));
    };
    pushLimit();
    // one bounce for every live path, after restarting the free slots
    auto iterate = [&](uint32_t cur) {
        const uint32_t nxt = cur ^ 1u;
        timeStage(kRaygen, [&] {
            pg::launchIterBegin(st, q, cur);
            pg::launchRegen(st, ds0, ps, q, rc0, cur, threads, btp);
        });
        timeStage(kExtend, [&] { pg::launchExtend(st, ds0, ps, q, cur, lanes, spill, spillPerLane, trav, btp, lt); });
        if (!keys.empty())  // grouped: 2 inference blocks of 8 waves per CU fit (LDS)
            timeStage(kGuide, [&] {
                const uint32_t blocks = static_cast<uint32_t>(dev->numCUs) * 2;
                if (opt.batch)
                    pg::launchGuideInferGrouped(st, bt, ps, q.ext[cur], q.counts + cur, cap,
                                                static_cast<uint32_t>(keys.size()), gb, keyWb, keyTheta,
                                                dev->mix.as<float>(), blocks);
                else
                    pg::launchGuideInferPaths(st, ds0, ps, q.ext[cur], q.counts + cur, cap, gs.lo, gs.invExt,
                                              keys[0]->wb.as<uint16_t>(), keys[0]->theta.as<float>(),
                                              dev->mix.as<float>(), static_cast<uint32_t>(dev->numCUs), lt.inferWaves);
            });
        timeStage(kShade, [&] {
            pg::launchShade(st, ds0, ps, q, rc0, cur, nxt, threads, dev->totals.as<uint64_t>(), gs, btp, anyTextured, lt);
        });
        timeStage(kShadow, [&] { pg::launchShadow(st, ds0, ps, q, lanes, spill, spillPerLane, btp, lt); });
        if (!envBlack) timeStage(kMis, [&] { pg::launchMis(st, ds0, ps, q, lanes, spill, spillPerLane, btp, lt); });
        timeStage(kFilm, [&] { pg::launchRecord(st, ps, q, rc0, threads, ring, gs, tr, btp); });
    };
    // Iterations run in groups of kCheckEvery (even, so every group starts on extension queue 0).
    // With opt.graph the group is captured once and replayed as one graph launch.
    GraphExec group;
    if (opt.graph) {
        require(!opt.timed, "stage timing is not available with graph launches");
        PG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipGraph_t g = nullptr;
        try {
            for (uint32_t k = 0; k < kCheckEvery; ++k) iterate(k & 1u);
        } catch (...) {
            (void)hipStreamEndCapture(st, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        PG_HIP(hipStreamEndCapture(st, &g));
        const hipError_t e = hipGraphInstantiate(&group.exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        PG_HIP(e);
    }
    uint64_t iterations = 0, lastNext = ~0ull;
    uint32_t lastFree = ~0u, stalled = 0;
    auto* statusCounts = static_cast<const uint32_t*>(dev->status);
    auto* statusPool = reinterpret_cast<const pg::PoolPaths*>(static_cast<const char*>(dev->status) + pg::kCntNum * 4);
    auto* statusItems = reinterpret_cast<const uint32_t*>(static_cast<const char*>(dev->status) + pg::kCountsBytes);
    for (;;) {
        if (group.exec) {
            PG_HIP(hipGraphLaunch(group.exec, st));
        } else {
            for (uint32_t k = 0; k < kCheckEvery; ++k) iterate(k & 1u);
        }
        PG_HIP(hipGetLastError());
        iterations += kCheckEvery;
        // progress: pool counters and the recorded-path counts of the oldest unresolved items
        const size_t window = std::min(kStatusItems, items.size() - std::min(nextResolve, items.size()));
        PG_HIP(hipMemcpyAsync(dev->status, dev->counts.p, pg::kCountsBytes, hipMemcpyDeviceToHost, st));
        if (window)
            PG_HIP(hipMemcpyAsync(static_cast<char*>(dev->status) + pg::kCountsBytes,
                                  dev->itemDone.as<uint32_t>() + nextResolve, window * 4, hipMemcpyDeviceToHost, st));
        PG_HIP(hipStreamSynchronize(st));
        const size_t base = nextResolve;
        while (nextResolve < items.size() && nextResolve - base < window &&
               statusItems[nextResolve - base] == items[nextResolve].count)
            resolveItem(nextResolve++);
        const uint64_t next = statusPool->next;
        const uint32_t freeSlots = statusCounts[pg::kCntDone];
        if (next == total && freeSlots == cap) break;  // every path started, finished and recorded
        stalled = (next == lastNext && freeSlots == lastFree) ? stalled + kCheckEvery : 0;
        if (stalled > kStallLimit) throw PgError(PG_ERR_DEVICE, "render made no progress");
        lastNext = next, lastFree = freeSlots;
        pushLimit();
    }
    while (nextResolve < items.size()) resolveItem(nextResolve++);  // all recorded by now
    PG_HIP(hipGetLastError());

    if (opt.pathLHost) {  // the ring holds every record (ringSize == total): scatter to caller order
        std::vector<float> recHost(total * 4);
        PG_HIP(hipMemcpyAsync(recHost.data(), ring.L, total * 16, hipMemcpyDeviceToHost, st));
        PG_HIP(hipStreamSynchronize(st));
        const pg::RenderConsts& rc = opt.rc0;
        for (const WorkItem& it : items) {
            const uint64_t pixels = static_cast<uint64_t>(rc.rectW) * rc.rectH;
            for (uint64_t p = 0; p < pixels; ++p)
                for (uint32_t k = 0; k < it.rc.sppCount; ++k) {
                    const float* r = &recHost[4 * (it.first + p * it.rc.sppCount + k)];
                    float* o = opt.pathLHost + 3 * (p * rc.sppCount + (it.rc.sppBegin - rc.sppBegin) + k);
                    o[0] = r[0], o[1] = r[1], o[2] = r[2];
                }
        }
    }
    uint64_t trained = 0;
    for (size_t k = 0; k < keys.size(); ++k) {  // online training on the records this render collected
        pg_guide* g = keys[k];
        if (g->p.train_records == 0) continue;
        uint32_t got = 0;
        PG_HIP(hipMemcpyAsync(&got, g->trainCount.p, 4, hipMemcpyDeviceToHost, st));
        PG_HIP(hipStreamSynchronize(st));
        const uint32_t n = std::min(got, g->p.train_records);
        const uint64_t stride = keyStride[k], phase = stride / 2;
        double trainPaths = 0.0;  // paths g with g % stride == phase, over the guide's items
        for (size_t i = 0; i < items.size(); ++i)
            if (itemKey[i] == k) trainPaths += static_cast<double>((items[i].count + stride - 1 - phase) / stride);
        if (trainPaths > 0.0) g->recordsPerPath = std::min(std::max(static_cast<double>(got) / trainPaths, 1e-6), 16.0);
        for (uint32_t b = 0; b + 64 <= n; b += g->p.train_batch) {
            const uint32_t cnt = std::min(g->p.train_batch, n - b);
            timeStage(kTrain, [&] { guideStep(g, g->trainRec.as<float>() + static_cast<size_t>(b) * 16, cnt, true); });
        }
        trained += n;
    }
    PG_HIP(hipStreamSynchronize(st));
    if (opt.latencyMs)
        for (uint32_t i = 0; i < nScenes; ++i) {
            float e = 0.0f;
            PG_HIP(hipEventElapsedTime(&e, start, sceneDone[i]));
            opt.latencyMs[i] = e;
        }
    if (pg_stats* stats = opt.stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->train_records = trained;
        uint64_t tot[pg::kTotNum];
        PG_HIP(hipMemcpy(tot, dev->totals.p, sizeof(tot), hipMemcpyDeviceToHost));
        stats->paths = total;
        stats->extension_rays = total + tot[pg::kTotExt];
        stats->shadow_rays = tot[pg::kTotShadow];
        stats->mis_rays = tot[pg::kTotMis];
        stats->extend_launches = iterations;
        if (trav) {
            unsigned long long tc[2];
            PG_HIP(hipMemcpy(tc, dev->travCounts.p, sizeof(tc), hipMemcpyDeviceToHost));
            stats->nodes_visited = tc[0];
            stats->tris_tested = tc[1];
        }
        timeStage.collect(stats);
        stats->ms_total =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
}

void render(pg_scene* s, const pg_render_params* params, pg_film* film, float* pathLHost, pg_stats* stats,
            pg_guide* guide = nullptr) {
    require(s && params, "null scene or params");
    pg_device* dev = s->dev;
    require(!guide || guide->dev == dev, "guide and scene live on different devices");
    std::lock_guard<std::mutex> lock(dev->mu);
    dev->checkFault();
    pg::RenderConsts rc = makeConsts(*params);
    require(!film || (film->width == rc.width && film->height == rc.height), "film size differs from params");
    rc.segSpp = segmentSpp(rc);
    std::vector<WorkItem> items;
    uint64_t total = 0;
    appendSegments(0, rc, items, total);
    RunOptions opt;
    opt.rc0 = rc;
    pg_film* films[1] = {film};
    opt.films = films;
    opt.capacity = params->wavefront_paths ? params->wavefront_paths : (1u << 21);
    opt.guide = guide;
    opt.pathLHost = pathLHost;
    opt.stats = stats;
    opt.countTraversal = rc.flags & pg::kCountTraversal;
    opt.graph = rc.flags & pg::kUseGraph;
    opt.timed = stats != nullptr && !(rc.flags & pg::kNoStageTiming) && !opt.graph;
    runPaths(dev, &s, 1, items, total, opt);
}

// Batched render: the paths of all scenes run through one path pool (their segments numbered
// one after another), so the pool mixes scenes, depths and resolutions and small scenes no longer
// leave the device idle. Per path the arithmetic is that of render(), and every scene is resolved
// segment by segment in the same order, so each film is bit-identical to its own render.
void renderBatch(pg_device* dev, pg_scene* const* scenes, const pg_render_params* params, uint32_t n,
                 pg_film* const* films, double* latencyMs, pg_stats* stats, pg_guide* const* guides = nullptr) {
    require(dev && scenes && params && n > 0, "empty batch");
    std::lock_guard<std::mutex> lock(dev->mu);
    dev->checkFault();
    std::vector<WorkItem> items;
    uint64_t total = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const pg_scene* s = scenes[i];
        require(s && s->dev == dev, "every batched scene must live on the batch's device");
        require(s->ds.width == scenes[0]->ds.width, "batched scenes must share one BVH width");
        pg::RenderConsts rc = makeConsts(params[i]);
        require(!(rc.flags & pg::kCountTraversal), "traversal counting is not available in batched renders");
        const pg_film* f = films ? films[i] : nullptr;
        require(!f || (f->dev == dev && f->width == rc.width && f->height == rc.height),
                "film size differs from params");
        rc.segSpp = segmentSpp(rc);
        appendSegments(i, rc, items, total);
    }
    RunOptions opt;
    opt.batch = true;
    opt.films = films;
    opt.capacity = params[0].wavefront_paths ? params[0].wavefront_paths : (1u << 21);
    opt.latencyMs = latencyMs;
    opt.stats = stats;
    opt.guides = guides;
    opt.graph = params[0].flags & pg::kUseGraph;
    opt.timed = stats != nullptr && !(params[0].flags & pg::kNoStageTiming) && !opt.graph;
    runPaths(dev, scenes, n, items, total, opt);
}

// Autotuner: times a slice of the render (the whole rect, the first samples of its range, about
// samplePaths paths) for every BVH width and path-pool capacity it tries, leaves the scene built
// at the fastest width and reports the fastest capacity. Each trial runs twice and the second run
// is timed, so workspace growth is not counted.
void autotune(pg_scene* s, const pg_render_params* params, uint64_t samplePaths, pg_tune_result* out) {
    require(s && params && out, "null argument");
    const pg::RenderConsts rc = makeConsts(*params);
    const uint64_t pixels = static_cast<uint64_t>(rc.rectW) * rc.rectH;
    const uint32_t k = static_cast<uint32_t>(
        std::max<uint64_t>(1, std::min<uint64_t>(rc.sppCount, std::max<uint64_t>(samplePaths, 1) / pixels)));
    pg_render_params p = *params;
    p.spp_end = p.spp_begin + k;
    p.flags &= ~(pg::kCountTraversal | pg::kFilmFp64);
    const uint64_t paths = pixels * k;
    const uint32_t widths[2] = {4, 8};
    const uint32_t caps[4] = {1u << 21, 1u << 22, 1u << 23, 1u << 24};
    double bestMs = 1e300;
    uint32_t bestWidth = s->ds.width, bestCap = caps[0], trials = 0;
    for (uint32_t w : widths) {
        if (s->ds.width != w && !s->triMat.empty()) {
            std::lock_guard<std::mutex> lock(s->dev->mu);
            pg_bvh_params bp{w, s->bvhParams.leafSize, s->bvhParams.bins, s->builder};
            buildBvh(s, &bp);
        }
        for (uint32_t c : caps) {
            if (c > caps[0] && c / 2 > paths) continue;  // a pool larger than the work changes nothing
            p.wavefront_paths = c;
            double ms = 0.0;
            for (int rep = 0; rep < 2; ++rep) {
                const auto t0 = std::chrono::steady_clock::now();
                render(s, &p, nullptr, nullptr, nullptr);
                ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            }
            ++trials;
            if (ms < bestMs) bestMs = ms, bestWidth = s->ds.width, bestCap = c;
        }
    }
    if (s->ds.width != bestWidth && !s->triMat.empty()) {
        std::lock_guard<std::mutex> lock(s->dev->mu);
        pg_bvh_params bp{bestWidth, s->bvhParams.leafSize, s->bvhParams.bins, s->builder};
        buildBvh(s, &bp);
    }
    p.wavefront_paths = bestCap;
    // Launch settings, one after another (coordinate descent) at the chosen width and capacity:
    // traversal LDS stack depth x traversal waves per CU, then the shade block, then the waves per
    // guide-inference block (timed on guided renders with a fresh, untrained guide: the network's
    // cost does not depend on its weights). The best of two runs per trial.
    auto timeTrial = [&](const pg::LaunchTune& t, pg_guide* g) {
        setTune(s, t);
        double best = 1e300;
        for (int rep = 0; rep < 2; ++rep) {
            const auto t0 = std::chrono::steady_clock::now();
            render(s, &p, nullptr, nullptr, nullptr, g);
            best = std::min(best, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        }
        ++trials;
        return best;
    };
    pg::LaunchTune tune = s->tune;
    double tuneMs = timeTrial(tune, nullptr);
    for (uint32_t depth : {6u, 8u, 10u, 12u, 16u})
        for (uint32_t waves : {0u, 16u, 20u, 24u}) {
            pg::LaunchTune t = tune;
            t.ldsStack = depth, t.travWaves = waves;
            if (t.ldsStack == tune.ldsStack && t.travWaves == tune.travWaves) continue;
            const double ms = timeTrial(t, nullptr);
            if (ms < tuneMs) tuneMs = ms, tune = t;
        }
    for (uint32_t blk : {128u, 256u}) {
        pg::LaunchTune t = tune;
        t.shadeBlock = blk;
        if (blk == tune.shadeBlock) continue;
        const double ms = timeTrial(t, nullptr);
        if (ms < tuneMs) tuneMs = ms, tune = t;
    }
    bestMs = std::min(bestMs, tuneMs);
    pg_guide_params gp{7u, 1e-2f, 0.9f, 0.99f, 1e-8f, 0.5f, 1u << 14, 0u};
    pg_guide* g = nullptr;
    if (pg_guide_create(s->dev, &gp, &g) == PG_OK) {
        double inferMs = 1e300;
        uint32_t inferBest = tune.inferWaves;
        for (uint32_t w : {4u, 8u, 16u}) {
            pg::LaunchTune t = tune;
            t.inferWaves = w;
            const double ms = timeTrial(t, g);
            if (ms < inferMs) inferMs = ms, inferBest = w;
        }
        tune.inferWaves = inferBest;
        pg_guide_destroy(g);
    }
    setTune(s, tune);
    out->bvh_width = bestWidth;
    out->wavefront_paths = bestCap;
    out->sample_paths = paths;
    out->trials = trials;
    out->mpaths_per_s = static_cast<double>(paths) / bestMs / 1e3;
    out->launch = configOf(tune);
}

}  // namespace

extern "C" {

const char* pg_last_error(void) { return g_lastError.c_str(); }

int pg_device_create(int ordinal, pg_device** out) {
    return guarded([&] {
        require(out != nullptr, "null output");
        int n = 0;
        PG_HIP(hipGetDeviceCount(&n));
        require(ordinal >= 0 && ordinal < n, "no such device");
        auto dev = std::make_unique<pg_device>();
        dev->ordinal = ordinal;
        PG_HIP(hipSetDevice(ordinal));
        hipDeviceProp_t prop;
        PG_HIP(hipGetDeviceProperties(&prop, ordinal));
        dev->numCUs = prop.multiProcessorCount;
        dev->platform = std::string("AMD HIP ") + prop.gcnArchName;
        PG_HIP(hipStreamCreateWithFlags(&dev->stream, hipStreamNonBlocking));
        dev->scratch = scratchAddress(ordinal);
        PG_HIP(hipHostMalloc(&dev->status, kPinnedBytes, hipHostMallocDefault));
        std::vector<uint32_t> m;
        std::vector<uint64_t> inv, delta;
        pg::buildSobolTables(m, inv, delta);
        upload(dev->sobol, m.data(), m.size() * 4, dev->stream);
        std::vector<uint32_t> mt(m.size());
        for (uint32_t d = 0; d < pg::kSobolDims; ++d)
            for (uint32_t j = 0; j < pg::kSobolCols; ++j) mt[j * pg::kSobolDims + d] = m[d * pg::kSobolCols + j];
        upload(dev->sobolT, mt.data(), mt.size() * 4, dev->stream);
        std::vector<uint32_t> mb;
        pg::buildSobolByteTable(m, mb);
        upload(dev->sobolB, mb.data(), mb.size() * 4, dev->stream);
        upload(dev->sobolInv, inv.data(), inv.size() * 8, dev->stream);
        upload(dev->sobolDelta, delta.data(), delta.size() * 8, dev->stream);
        PG_HIP(hipStreamSynchronize(dev->stream));
        *out = dev.release();
    });
}

void pg_device_destroy(pg_device* dev) {
    if (!dev) return;
    std::lock_guard<std::mutex> g(g_lifetime);
    dev->closed = true;
    if (dev->children == 0) freeDevice(dev);
}

const char* pg_device_platform(pg_device* dev) { return dev ? dev->platform.c_str() : ""; }

int pg_debug_inject_fault(pg_device* dev) {
    return guarded([&] {
        require(dev != nullptr, "null device");
        std::lock_guard<std::mutex> lock(dev->mu);
        dev->injectFault = true;
    });
}

namespace {
float remapRoughness(float roughness) {  // MicrofacetDistribution::remapRoughness, Classic/Kernel.cpp:378-381
    const float x = std::log(std::fmax(1e-3f, roughness));
    return std::fmax(1e-3f, 1.62142f + (0.819955f + (0.1734f + (0.0171201f + 0.000640711f * x) * x) * x) * x);
}

int sceneCreate(pg_device* dev, const pg_scene_desc* desc, const uint32_t* instanceFirst, uint32_t numInstances,
                pg_scene** out) {
    return guarded([&] {
        require(dev && desc && out, "null argument");
        std::lock_guard<std::mutex> lock(dev->mu);
        dev->checkFault();
        const uint32_t nt = desc->num_triangles;
        require(nt == 0 || (desc->vertices && desc->indices && desc->tri_object && desc->object_material),
                "missing mesh arrays");
        require(desc->num_materials > 0 && desc->materials, "at least one material is required");
        require(desc->num_quads <= 4096 && (desc->num_quads == 0 || desc->quads), "bad light quads");
        require(desc->num_point_lights <= 4096 && (desc->num_point_lights == 0 || desc->point_lights), "bad point lights");
        auto s = std::make_unique<pg_scene>();
        s->dev = dev;
        s->triVerts.resize(static_cast<size_t>(nt) * 9);
        s->triMat.resize(nt);
        for (uint32_t i = 0; i < nt; ++i) {
            for (int v = 0; v < 3; ++v) {
                const uint32_t j = desc->indices[3 * i + v];
                require(j < desc->num_vertices, "triangle index out of range");
                for (int k = 0; k < 3; ++k) s->triVerts[9 * i + 3 * v + k] = desc->vertices[3 * j + k];
            }
            const uint32_t obj = desc->tri_object[i];
            require(obj < desc->num_objects, "object id out of range");
            const uint32_t mat = desc->object_material[obj];
            require(mat < desc->num_materials, "material id out of range");
            s->triMat[i] = mat;
        }
        s->numStatic = nt;
        if (numInstances) {
            require(instanceFirst && numInstances <= pg::kMaxInstances, "bad moving instances");
            s->instFirst.assign(instanceFirst, instanceFirst + numInstances);
            s->instFirst.push_back(nt);
            for (uint32_t k = 0; k < numInstances; ++k)
                require(s->instFirst[k] < s->instFirst[k + 1], "every moving instance needs its own non-empty triangle range");
            s->numStatic = s->instFirst[0];
            s->instBox.resize(6 * static_cast<size_t>(numInstances));
            for (uint32_t k = 0; k < numInstances; ++k) {
                float* b = &s->instBox[6 * static_cast<size_t>(k)];
                for (int a = 0; a < 3; ++a) b[a] = INFINITY, b[3 + a] = -INFINITY;
                for (size_t v = 3 * static_cast<size_t>(s->instFirst[k]); v < 3 * static_cast<size_t>(s->instFirst[k + 1]); ++v)
                    for (int a = 0; a < 3; ++a)
                        b[a] = std::min(b[a], s->triVerts[3 * v + a]), b[3 + a] = std::max(b[3 + a], s->triVerts[3 * v + a]);
                s->inst.push_back(fixedInstance(b));
            }
        }
        std::vector<pg::DevMaterial> mats(desc->num_materials);
        for (uint32_t i = 0; i < desc->num_materials; ++i) {
            const pg_material& m = desc->materials[i];
            require(m.type <= PG_MAT_SUBSTRATE, "unknown material type");
            require(m.type != PG_MAT_GLASS || m.alpha >= 0.0f, "glass roughness must not be negative");
            require(m.type != PG_MAT_CONDUCTOR || m.alpha > 0.0f, "conductor alpha must be positive");
            require(m.type != PG_MAT_LAMBERT || (m.alpha >= 0.0f && m.alpha <= 1.0f), "matte roughness must be in [0, 1]");
            uint32_t type = m.type;
            float alpha = m.alpha;
            float spec[3] = {m.reserved[0], m.reserved[1], m.reserved[2]};
            const bool twoRough = m.type == PG_MAT_GLASS || m.type == PG_MAT_SUBSTRATE;  // RoughnessX, RoughnessY
            const float roughY = desc->roughness_y ? desc->roughness_y[i] : m.alpha;
            const uint32_t texY = desc->roughness_y_textures ? desc->roughness_y_textures[i] : PG_TEX_NONE;
            require(twoRough || (roughY == m.alpha && texY == PG_TEX_NONE),
                    "only glass and substrate have a second roughness");
            require(m.type != PG_MAT_GLASS || roughY >= 0.0f, "glass roughness must not be negative");
            require(texY == PG_TEX_NONE || texY < desc->num_textures, "material texture index out of range");
            float alphaY = alpha;
            // glassInit (Classic/Kernel.cpp:630-655): smooth only when both roughnesses are zero
            if (m.type == PG_MAT_GLASS && (m.alpha > 0.0f || roughY > 0.0f)) type = pg::kDevRoughGlass;
            if (m.type == PG_MAT_PLASTIC || m.type == PG_MAT_SUBSTRATE || type == pg::kDevRoughGlass) {
                alpha = remapRoughness(m.alpha);  // MicrofacetDistribution::remapRoughness
                alphaY = twoRough ? remapRoughness(roughY) : alpha;
            } else if (m.type == PG_MAT_LAMBERT && m.alpha > 0.0f) {
                // matteInit (Classic/Kernel.cpp:260-266) picks OrenNayar; its constructor (:99-106)
                // takes sigma = roughness x pi radians
                const float sigma = m.alpha * 3.14159265358979323846f;
                const float s2 = sigma * sigma;
                type = pg::kDevOrenNayar;
                alpha = 1.0f - s2 / (2.0f * (s2 + 0.33f));
                spec[0] = 0.45f * s2 / (s2 + 0.09f), spec[1] = 0.0f, spec[2] = 0.0f;
            }
            mats[i] = pg::DevMaterial{type,   {m.albedo[0], m.albedo[1], m.albedo[2]}, alpha, {spec[0], spec[1], spec[2]},
                                      alphaY, m.alpha, roughY, texY};
        }
        // textures: per-triangle corner texcoords, the texture table with one texel pool, and the
        // binding of each material's colour inputs
        std::vector<pg::DevTexture> texTab(std::max<uint32_t>(desc->num_textures, 1));
        std::vector<uint8_t> texelPool;
        require(desc->num_textures == 0 || desc->textures, "textures missing");
        for (uint32_t i = 0; i < desc->num_textures; ++i) {
            const pg_texture& t = desc->textures[i];
            require(t.channels == 1 || t.channels == 2 || t.channels == 4, "texture channels must be 1, 2 or 4");
            pg::DevTexture& d = texTab[i];
            d = pg::DevTexture{};
            d.type = t.type, d.channels = t.channels;
            std::memcpy(d.black, t.black, 16), std::memcpy(d.white, t.white, 16);
            if (t.type == PG_TEX_CHECKERBOARD) {
                require(std::isfinite(t.scale), "checkerboard scale must be finite");
                d.scale = t.scale;
            } else {
                require(t.type == PG_TEX_IMAGE, "unknown texture type");
                require(t.width > 0 && t.height > 0 && t.width <= 65536 && t.height <= 65536 && t.texels,
                        "image texture needs width, height and texels");
                require(t.wrap == PG_WRAP_REPEAT || t.wrap == PG_WRAP_MIRROR, "unknown texture wrap mode");
                d.width = t.width, d.height = t.height, d.wrap = t.wrap;
                d.offset = static_cast<uint32_t>(texelPool.size());
                const size_t bytes = static_cast<size_t>(t.width) * t.height * t.channels;
                require(texelPool.size() + bytes < (1ull << 32), "texel pool exceeds 4 GiB");
                texelPool.insert(texelPool.end(), t.texels, t.texels + bytes);
            }
        }
        std::vector<uint4> matTex(desc->num_materials, uint4{pg::kTexNone, pg::kTexNone, pg::kTexNone, 0u});
        bool textured = false;
        if (desc->material_textures) {
            for (uint32_t i = 0; i < desc->num_materials; ++i) {
                const pg_material_textures& b = desc->material_textures[i];
                for (uint32_t t : {b.albedo, b.reserved, b.alpha})
                    require(t == PG_TEX_NONE || t < desc->num_textures, "material texture index out of range");
                for (uint32_t t : {b.albedo, b.reserved})
                    require(t == PG_TEX_NONE || desc->textures[t].channels >= 3, "colour textures need 4 channels");
                require(b.alpha == PG_TEX_NONE || desc->materials[i].type != PG_MAT_MIRROR, "a mirror has no roughness");
                const uint32_t mt = desc->materials[i].type;
                require(b.reserved == PG_TEX_NONE || mt == PG_MAT_PLASTIC || mt == PG_MAT_SUBSTRATE || mt == PG_MAT_GLASS,
                        "only plastic, substrate and glass have a second colour input");
                matTex[i] = uint4{b.albedo, b.reserved, b.alpha, desc->materials[i].type};
            }
        }
        // without its own RoughnessY a glass or substrate is isotropic: RoughnessY is RoughnessX, texture too
        const bool ownY = desc->roughness_y || desc->roughness_y_textures;
        for (uint32_t i = 0; i < desc->num_materials; ++i) {
            const uint32_t mt = desc->materials[i].type;
            if (!ownY && (mt == PG_MAT_GLASS || mt == PG_MAT_SUBSTRATE)) mats[i].texY = matTex[i].z;
        }
        for (uint32_t i = 0; i < desc->num_materials; ++i) {
            const uint4 b = matTex[i];
            if (b.x != PG_TEX_NONE || b.y != PG_TEX_NONE || b.z != PG_TEX_NONE || mats[i].texY != PG_TEX_NONE) {
                matTex[i].w = desc->materials[i].type;
                mats[i].type |= pg::kMatTextured;
                textured = true;
            }
        }
        std::vector<float> triUV;
        if (textured && desc->texcoords) {
            triUV.resize(static_cast<size_t>(nt) * 6);
            for (uint32_t i = 0; i < nt; ++i)
                for (int k = 0; k < 3; ++k) {
                    const uint32_t v = desc->indices[3 * i + k];
                    triUV[6 * i + 2 * k] = desc->texcoords[2 * v];
                    triUV[6 * i + 2 * k + 1] = desc->texcoords[2 * v + 1];
                }
        }
        // area lights: light l is the next area_light_quads[l] quads, or every quad a light of its own
        std::vector<pg::DevAreaLight> lights;
        if (desc->area_light_quads) {
            require(desc->num_area_lights <= 4096, "bad area lights");
            uint32_t first = 0;
            for (uint32_t l = 0; l < desc->num_area_lights; ++l) {
                const uint32_t c = desc->area_light_quads[l];
                require(c > 0 && c <= desc->num_quads - first, "area light primitive counts must cover the quads");
                lights.push_back({first, c});
                first += c;
            }
            require(first == desc->num_quads, "area light primitive counts must cover the quads");
        } else {
            for (uint32_t i = 0; i < desc->num_quads; ++i) lights.push_back({i, 1});
        }
        std::vector<pg::DevQuad> quads(desc->num_quads);
        for (uint32_t i = 0; i < desc->num_quads; ++i) {
            const pg_quad& q = desc->quads[i];
            pg::DevQuad& d = quads[i];
            std::memcpy(d.origin, q.origin, 12), std::memcpy(d.u, q.u, 12), std::memcpy(d.v, q.v, 12);
            std::memcpy(d.radiance, q.radiance, 12);
            // BasicShapes.cpp:70-90: normal = U x V / |U x V|, area, largest-determinant axis
            const float det[3] = {q.u[1] * q.v[2] - q.u[2] * q.v[1], q.u[2] * q.v[0] - q.u[0] * q.v[2],
                                  q.u[0] * q.v[1] - q.u[1] * q.v[0]};
            d.area = std::sqrt(det[0] * det[0] + det[1] * det[1] + det[2] * det[2]);
            require(d.area > 0.0f, "degenerate light quad");
            for (int k = 0; k < 3; ++k) d.normal[k] = det[k] / d.area;
            const float a0 = std::fabs(det[0]), a1 = std::fabs(det[1]), a2 = std::fabs(det[2]);
            d.maxDetComp = (a0 >= a1 && a0 >= a2) ? 0u : ((a1 >= a0 && a1 >= a2) ? 1u : 2u);
        }
        // Plane (BasicShapes.cpp:59-94): the light's area is the float sum of its primitives' areas in
        // order, a primitive's CDF entry the running sum over that total, its probability its area over it;
        // DiffuseAreaLight has one radiance, so the primitives of one light must share it
        for (uint32_t l = 0; l < lights.size(); ++l) {
            const pg::DevAreaLight L = lights[l];
            float total = 0.0f;
            for (uint32_t k = 0; k < L.count; ++k) {
                pg::DevQuad& d = quads[L.first + k];
                require(std::memcmp(d.radiance, quads[L.first].radiance, 12) == 0,
                        "the primitives of one area light must share its radiance");
                total = total + d.area;
                d.cdf = total;
            }
            for (uint32_t k = 0; k < L.count; ++k) {
                pg::DevQuad& d = quads[L.first + k];
                d.light = l;
                d.cdf = d.cdf / total;
                d.pdf = d.area / total;
                d.area = total;
            }
        }
        hipStream_t st = dev->stream;
        upload(s->mats, mats.data(), mats.size() * sizeof(pg::DevMaterial), st);
        upload(s->quads, quads.data(), quads.size() * sizeof(pg::DevQuad), st);
        upload(s->lights, lights.data(), lights.size() * sizeof(pg::DevAreaLight), st);
        upload(s->points, desc->point_lights, desc->num_point_lights * sizeof(pg_point_light), st);
        pg::DevScene& ds = s->ds;
        ds.mats = s->mats.as<pg::DevMaterial>();
        ds.quads = s->quads.as<pg::DevQuad>();
        ds.numQuads = desc->num_quads;
        ds.lights = s->lights.as<pg::DevAreaLight>();
        ds.numLights = static_cast<uint32_t>(lights.size());
        ds.points = s->points.as<float>();
        ds.numPoints = desc->num_point_lights;
        upload(s->textures, texTab.data(), texTab.size() * sizeof(pg::DevTexture), st);
        upload(s->texels, texelPool.data(), texelPool.size(), st);
        upload(s->matTex, matTex.data(), matTex.size() * sizeof(uint4), st);
        upload(s->triUV, triUV.data(), triUV.size() * sizeof(float), st);
        ds.textures = s->textures.as<pg::DevTexture>();
        ds.texels = s->texels.as<uint8_t>();
        ds.matTex = s->matTex.as<uint4>();
        ds.triUV = triUV.empty() ? nullptr : s->triUV.as<float>();
        ds.textured = textured ? 1u : 0u;
        ds.sobol = dev->sobol.as<uint32_t>();
        ds.sobolT = dev->sobolT.as<uint32_t>();
        ds.sobolB = dev->sobolB.as<uint32_t>();
        ds.sobolInv = dev->sobolInv.as<uint64_t>();
        ds.sobolDelta = dev->sobolDelta.as<uint64_t>();
        const pg_camera& c = desc->camera;
        std::memcpy(ds.cam.base, c.position, 12), std::memcpy(ds.cam.lookAt, c.look_at, 12);
        std::memcpy(ds.cam.up, c.up, 12);
        ds.cam.sizeX = c.sensor_size[0], ds.cam.sizeY = c.sensor_size[1];
        ds.cam.focal = c.focal_length, ds.cam.aperture = c.aperture_radius;
        float blo[3] = {INFINITY, INFINITY, INFINITY}, bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
        auto grow = [&](const float* p) {
            for (int k = 0; k < 3; ++k) blo[k] = std::min(blo[k], p[k]), bhi[k] = std::max(bhi[k], p[k]);
        };
        for (size_t i = 0; i < 9 * static_cast<size_t>(s->numStatic); i += 3) grow(&s->triVerts[i]);
        for (const auto& d : quads) {
            float c4[3];
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b) {
                    for (int k = 0; k < 3; ++k) c4[k] = d.origin[k] + a * d.u[k] + b * d.v[k];
                    grow(c4);
                }
        }
        for (int k = 0; k < 3; ++k) s->staticLo[k] = blo[k], s->staticHi[k] = bhi[k];
        updateSceneBox(s.get());
        std::memcpy(ds.env, desc->env_radiance, 12);
        const pg_bvh_params defaults{8, 4, 32, PG_BVH_DEVICE};  // BVH8, leaves of <= 4, 32 bins, built on the GPU
        buildBvh(s.get(), &defaults);
        addChild(dev);
        *out = s.release();
    });
}
}  // namespace

int pg_scene_create(pg_device* dev, const pg_scene_desc* desc, pg_scene** out) {
    return sceneCreate(dev, desc, nullptr, 0, out);
}

int pg_scene_create_instanced(pg_device* dev, const pg_scene_desc* desc, const uint32_t* instance_first,
                              uint32_t num_instances, pg_scene** out) {
    return sceneCreate(dev, desc, instance_first, num_instances, out);
}

int pg_scene_set_motion(pg_scene* s, uint32_t instance, const pg_motion* m) {
    return guarded([&] {
        require(s && m, "null argument");
        require(instance < s->inst.size(), "instance index out of range");
        require(m->num_keys >= 1 && m->num_keys <= PG_MAX_MOTION_KEYS, "a motion has 1 to 8 keys");
        std::lock_guard<std::mutex> lock(s->dev->mu);
        s->dev->checkFault();
        pg::DevInstance I = s->inst[instance];
        I.keys = m->num_keys;
        I.visible = m->visible ? 1u : 0u;
        if (I.keys > 1) {
            require(std::isfinite(m->t0) && std::isfinite(m->t1) && m->t1 > m->t0, "key times need t0 < t1");
            I.t0 = m->t0, I.t1 = m->t1;
            I.rate = static_cast<float>(I.keys - 1) / (m->t1 - m->t0);
        } else {
            I.t0 = -FLT_MAX, I.t1 = FLT_MAX, I.rate = 0.0f;
        }
        for (uint32_t k = 0; k < I.keys; ++k) {
            const float* src = m->keys[k];
            for (int i = 0; i < 16; ++i) require(std::isfinite(src[i]), "motion keys must be finite");
            const float n = std::sqrt(src[9] * src[9] + src[10] * src[10] + src[11] * src[11] + src[12] * src[12]);
            require(n > 0.0f, "a motion key has a zero quaternion");
            std::memcpy(I.srt[k], src, 64);
            for (int i = 9; i < 13; ++i) I.srt[k][i] = src[i] / n;
        }
        for (int i = 0; i < 12; ++i) require(std::isfinite(m->parent[i]), "parent transform must be finite");
        std::memcpy(I.parent, m->parent, sizeof(I.parent));
        if (I.visible && !sweptBox(I, &s->instBox[6 * static_cast<size_t>(instance)], I.lo, I.hi)) I.visible = 0;
        s->inst[instance] = I;
        updateSceneBox(s);
        hipStream_t st = s->dev->stream;
        upload(s->instBuf, s->inst.data(), s->inst.size() * sizeof(pg::DevInstance), st);
        PG_HIP(hipStreamSynchronize(st));
    });
}

int pg_scene_set_launch(pg_scene* s, const pg_launch_config* c) {
    return guarded([&] {
        require(s && c, "null argument");
        const pg::LaunchTune t = tuneFrom(*c);
        std::lock_guard<std::mutex> lock(s->dev->mu);
        setTune(s, t);
    });
}

int pg_scene_get_launch(pg_scene* s, pg_launch_config* c) {
    return guarded([&] {
        require(s && c, "null argument");
        *c = configOf(s->tune);
    });
}

int pg_scene_build_bvh(pg_scene* s, const pg_bvh_params* params) {
    return guarded([&] {
        require(s && params, "null argument");
        std::lock_guard<std::mutex> lock(s->dev->mu);
        s->dev->checkFault();
        buildBvh(s, params);
    });
}

void pg_scene_destroy(pg_scene* s) {
    if (!s) return;
    pg_device* dev = s->dev;
    {
        std::lock_guard<std::mutex> lock(dev->mu);
        (void)hipStreamSynchronize(dev->stream);
        delete s;
    }
    releaseChild(dev);
}

int pg_scene_bvh_info(pg_scene* s, uint64_t info[6]) {
    return guarded([&] {
        require(s && info, "null argument");
        info[0] = s->bvh.numNodes;
        info[1] = s->bvh.numLeaves;
        info[2] = s->bvh.maxDepth;
        info[3] = s->bvh.width;
        info[4] = 7ull * 4 * s->bvh.width;
        info[5] = 48;
    });
}

int pg_scene_bvh_download(pg_scene* s, void* nodes, uint64_t nodeBytes, void* tris, uint64_t triBytes) {
    return guarded([&] {
        require(s != nullptr, "null scene");
        std::lock_guard<std::mutex> lock(s->dev->mu);
        const size_t nb = static_cast<size_t>(s->bvh.numNodes) * 7 * 4 * s->bvh.width;
        const size_t tb = s->triMat.size() * 48;
        require((!nodes || nodeBytes >= nb) && (!tris || triBytes >= tb), "buffers too small");
        if (nodes && nb) PG_HIP(hipMemcpy(nodes, s->nodes.p, nb, hipMemcpyDeviceToHost));
        if (tris && tb) PG_HIP(hipMemcpy(tris, s->tris.p, tb, hipMemcpyDeviceToHost));
    });
}

int pg_film_create(pg_device* dev, uint32_t width, uint32_t height, pg_film** out) {
    return guarded([&] {
        require(dev && out && width > 0 && height > 0, "bad film arguments");
        std::lock_guard<std::mutex> lock(dev->mu);
        dev->checkFault();
        auto f = std::make_unique<pg_film>();
        f->dev = dev;
        f->width = width, f->height = height;
        f->rgbw32.reserve(static_cast<size_t>(width) * height * 16);
        PG_HIP(hipMemsetAsync(f->rgbw32.p, 0, f->rgbw32.bytes, dev->stream));
        PG_HIP(hipStreamSynchronize(dev->stream));
        addChild(dev);
        *out = f.release();
    });
}

int pg_film_clear(pg_film* f) {
    return guarded([&] {
        require(f != nullptr, "null film");
        std::lock_guard<std::mutex> lock(f->dev->mu);
        f->dev->checkFault();
        PG_HIP(hipMemsetAsync(f->rgbw32.p, 0, f->rgbw32.bytes, f->dev->stream));
        if (f->rgbw64.p) PG_HIP(hipMemsetAsync(f->rgbw64.p, 0, f->rgbw64.bytes, f->dev->stream));
        PG_HIP(hipStreamSynchronize(f->dev->stream));
    });
}

void pg_film_destroy(pg_film* f) {
    if (!f) return;
    pg_device* dev = f->dev;
    {
        std::lock_guard<std::mutex> lock(dev->mu);
        (void)hipStreamSynchronize(dev->stream);
        delete f;
    }
    releaseChild(dev);
}

int pg_film_download(pg_film* f, float* rgbw, double* rgbw64) {
    return guarded([&] {
        require(f != nullptr, "null film");
        std::lock_guard<std::mutex> lock(f->dev->mu);
        f->dev->checkFault();
        const size_t px = static_cast<size_t>(f->width) * f->height;
        PG_HIP(hipStreamSynchronize(f->dev->stream));
        if (rgbw) PG_HIP(hipMemcpy(rgbw, f->rgbw32.p, px * 16, hipMemcpyDeviceToHost));
        if (rgbw64) {
            if (f->rgbw64.p) PG_HIP(hipMemcpy(rgbw64, f->rgbw64.p, px * 32, hipMemcpyDeviceToHost));
            else std::memset(rgbw64, 0, px * 32);
        }
    });
}

int pg_render(pg_scene* s, const pg_render_params* params, pg_film* film, pg_stats* stats) {
    return guarded([&] { render(s, params, film, nullptr, stats); });
}

int pg_render_guided(pg_scene* s, pg_guide* guide, const pg_render_params* params, pg_film* film, pg_stats* stats) {
    return guarded([&] {
        require(guide != nullptr, "null guide");
        render(s, params, film, nullptr, stats, guide);
    });
}

int pg_render_paths(pg_scene* s, const pg_render_params* params, pg_film* film, float* pathL, pg_stats* stats) {
    return guarded([&] {
        require(pathL != nullptr, "null path_L");
        render(s, params, film, pathL, stats);
    });
}

int pg_render_batch(pg_device* dev, pg_scene* const* scenes, const pg_render_params* params, uint32_t n,
                    pg_film* const* films, double* latency_ms, pg_stats* stats) {
    return guarded([&] { renderBatch(dev, scenes, params, n, films, latency_ms, stats); });
}

int pg_render_batch_guided(pg_device* dev, pg_scene* const* scenes, pg_guide* const* guides,
                           const pg_render_params* params, uint32_t n, pg_film* const* films, double* latency_ms,
                           pg_stats* stats) {
    return guarded([&] {
        require(guides != nullptr, "null guides");
        renderBatch(dev, scenes, params, n, films, latency_ms, stats, guides);
    });
}

int pg_autotune(pg_scene* scene, const pg_render_params* params, uint64_t sample_paths, pg_tune_result* out) {
    return guarded([&] { autotune(scene, params, sample_paths, out); });
}

int pg_trace(pg_scene* s, const float* rays, uint32_t n, float* t, uint32_t* prim, float* uv) {
    return pg_trace_at(s, rays, nullptr, n, t, prim, uv);
}

int pg_trace_at(pg_scene* s, const float* rays, const float* times, uint32_t n, float* t, uint32_t* prim, float* uv) {
    return guarded([&] {
        require(s && (n == 0 || (rays && t && prim && uv)), "null argument");
        if (n == 0) return;
        pg_device* dev = s->dev;
        std::lock_guard<std::mutex> lock(dev->mu);
        dev->checkFault();
        hipStream_t st = dev->stream;
        DevBuf dr, dtimes, dt, dp, duv, spill;
        upload(dr, rays, static_cast<size_t>(n) * 32, st);
        if (times) upload(dtimes, times, static_cast<size_t>(n) * 4, st);
        dt.reserve(static_cast<size_t>(n) * 4), dp.reserve(static_cast<size_t>(n) * 4), duv.reserve(static_cast<size_t>(n) * 8);
        const uint32_t lanes = traversalLanes(dev, n);
        spill.reserve(static_cast<size_t>(lanes) * s->spillPerLane * 8);
        pg::launchTrace(st, s->ds, dr.as<float>(), n, dt.as<float>(), dp.as<uint32_t>(), duv.as<float>(), nullptr, lanes,
                        spill.as<uint2>(), s->spillPerLane, s->tune, times ? dtimes.as<float>() : nullptr);
        PG_HIP(hipGetLastError());
        PG_HIP(hipStreamSynchronize(st));
        PG_HIP(hipMemcpy(t, dt.p, static_cast<size_t>(n) * 4, hipMemcpyDeviceToHost));
        PG_HIP(hipMemcpy(prim, dp.p, static_cast<size_t>(n) * 4, hipMemcpyDeviceToHost));
        PG_HIP(hipMemcpy(uv, duv.p, static_cast<size_t>(n) * 8, hipMemcpyDeviceToHost));
    });
}

int pg_trace_noalloc(pg_scene* s, const float* rays, uint32_t n, float* t, uint32_t* prim, float* uv, void* stream) {
    return guarded([&] {
        require(s && (n == 0 || (rays && t && prim && uv)), "null argument");
        if (n == 0) return;
        pg_device* dev = s->dev;
        std::lock_guard<std::mutex> lock(dev->mu);
        dev->checkFault();
        PG_HIP(hipSetDevice(dev->ordinal));
        const size_t perLane = static_cast<size_t>(s->spillPerLane) * sizeof(uint2);
        const size_t fitLanes = kScratchBytes / perLane / 64 * 64;
        if (fitLanes < 64)
            throw PgError(PG_ERR_INVALID, "BVH too deep for the fixed traversal scratch: one wave needs " +
                                              std::to_string(perLane * 64) + " bytes, the scratch holds " +
                                              std::to_string(kScratchBytes));
        const uint32_t lanes = static_cast<uint32_t>(std::min<size_t>(traversalLanes(dev, n), fitLanes));
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : dev->stream;
        pg::launchTrace(st, s->ds, rays, n, t, prim, uv, nullptr, lanes, static_cast<uint2*>(dev->scratch),
                        s->spillPerLane, s->tune);
        PG_HIP(hipGetLastError());
        if (!stream) PG_HIP(hipStreamSynchronize(st));
    });
}

int pg_occluded(pg_scene* s, const float* rays, uint32_t n, uint8_t* occ) {
    return guarded([&] {
        require(s && (n == 0 || (rays && occ)), "null argument");
        if (n == 0) return;
        pg_device* dev = s->dev;
        std::lock_guard<std::mutex> lock(dev->mu);
        dev->checkFault();
        hipStream_t st = dev->stream;
        DevBuf dr, docc, spill;
        upload(dr, rays, static_cast<size_t>(n) * 32, st);
        docc.reserve(n);
        const uint32_t lanes = traversalLanes(dev, n);
        spill.reserve(static_cast<size_t>(lanes) * s->spillPerLane * 8);
        pg::launchTrace(st, s->ds, dr.as<float>(), n, nullptr, nullptr, nullptr, docc.as<uint8_t>(), lanes,
                        spill.as<uint2>(), s->spillPerLane, s->tune);
        PG_HIP(hipGetLastError());
        PG_HIP(hipStreamSynchronize(st));
        PG_HIP(hipMemcpy(occ, docc.p, n, hipMemcpyDeviceToHost));
    });
}

int pg_camera_rays(pg_scene* s, const pg_render_params* params, const uint32_t* pix, uint32_t n, float* rays,
                   float* film) {
    return guarded([&] {
        require(s && params && (n == 0 || (pix && rays && film)), "null argument");
        if (n == 0) return;
        const pg::RenderConsts rc = makeConsts(*params);
        for (uint32_t i = 0; i < n; ++i)
            require(pix[3 * i] < rc.width && pix[3 * i + 1] < rc.height && pix[3 * i + 2] < rc.spp, "pixel out of range");
        pg_device* dev = s->dev;
        std::lock_guard<std::mutex> lock(dev->mu);
        dev->checkFault();
        hipStream_t st = dev->stream;
        DevBuf dpix, drays, dfilm;
        upload(dpix, pix, static_cast<size_t>(n) * 12, st);
        drays.reserve(static_cast<size_t>(n) * 32), dfilm.reserve(static_cast<size_t>(n) * 8);
        pg::launchCameraRays(st, s->ds, rc, dpix.as<uint32_t>(), n, drays.as<float>(), dfilm.as<float>());
        PG_HIP(hipGetLastError());
        PG_HIP(hipStreamSynchronize(st));
        PG_HIP(hipMemcpy(rays, drays.p, static_cast<size_t>(n) * 32, hipMemcpyDeviceToHost));
        PG_HIP(hipMemcpy(film, dfilm.p, static_cast<size_t>(n) * 8, hipMemcpyDeviceToHost));
    });
}

int pg_sobol(pg_device* dev, uint32_t m, uint32_t dims, const uint32_t* pix, uint32_t n, float* out,
             uint64_t* indexOut) {
    return guarded([&] {
        require(dev && (n == 0 || (pix && out)), "null argument");
        require(m <= pg::kMaxLog2Res && dims <= pg::kSobolDims, "resolution or dimension out of range");
        if (n == 0) return;
        for (uint32_t i = 0; i < n; ++i) {
            require(pix[3 * i] < (1u << m) && pix[3 * i + 1] < (1u << m), "pixel outside the 2^m grid");
            require((static_cast<uint64_t>(pix[3 * i + 2]) << (2 * m)) < (1ull << pg::kSobolCols),
                    "sample index beyond the Sobol matrix columns");
        }
        std::lock_guard<std::mutex> lock(dev->mu);
        dev->checkFault();
        hipStream_t st = dev->stream;
        DevBuf dpix, dout, didx;
        upload(dpix, pix, static_cast<size_t>(n) * 12, st);
        dout.reserve(std::max<size_t>(static_cast<size_t>(n) * dims * 4, 4));
        didx.reserve(static_cast<size_t>(n) * 8);
        pg::DevScene ds{};
        ds.sobol = dev->sobol.as<uint32_t>();
        ds.sobolInv = dev->sobolInv.as<uint64_t>();
        ds.sobolDelta = dev->sobolDelta.as<uint64_t>();
        pg::launchSobol(st, ds, m, dims, dpix.as<uint32_t>(), n, dout.as<float>(), didx.as<uint64_t>());
        PG_HIP(hipGetLastError());
        PG_HIP(hipStreamSynchronize(st));
        if (dims) PG_HIP(hipMemcpy(out, dout.p, static_cast<size_t>(n) * dims * 4, hipMemcpyDeviceToHost));
        if (indexOut) PG_HIP(hipMemcpy(indexOut, didx.p, static_cast<size_t>(n) * 8, hipMemcpyDeviceToHost));
    });
}

int pg_debug_copy(pg_device* dev, uint64_t bytes, uint32_t reps, double* gbps) {
    return guarded([&] {
        require(dev && bytes >= 16 && bytes % 16 == 0 && reps > 0, "bad copy arguments");
        std::lock_guard<std::mutex> lock(dev->mu);
        dev->checkFault();
        hipStream_t st = dev->stream;
        DevBuf a, b;
        a.reserve(bytes), b.reserve(bytes);
        PG_HIP(hipMemsetAsync(a.p, 1, bytes, st));
        hipEvent_t e0 = dev->event(), e1 = dev->event();
        pg::launchCopy(st, a.p, b.p, bytes);  // warm-up
        PG_HIP(hipEventRecord(e0, st));
        for (uint32_t r = 0; r < reps; ++r) pg::launchCopy(st, a.p, b.p, bytes);
        PG_HIP(hipEventRecord(e1, st));
        PG_HIP(hipGetLastError());
        PG_HIP(hipStreamSynchronize(st));
        float ms = 0.0f;
        PG_HIP(hipEventElapsedTime(&ms, e0, e1));
        dev->eventsUsed = 0;
        if (gbps) *gbps = 2.0 * static_cast<double>(bytes) * reps / (ms * 1e-3) / 1e9;
    });
}

// ---------------------------------------------------------------------------------------------
// guiding network
// ---------------------------------------------------------------------------------------------
int pg_guide_create(pg_device* dev, const pg_guide_params* params, pg_guide** out) {
    return guarded([&] {
        require(dev && params && out, "null argument");
        require(params->mis_fraction < 1.0f, "mis_fraction must be below 1 (negative: learned)");
        require(params->lr > 0.0f && params->train_batch >= 64, "bad learning rate or batch");
        std::lock_guard<std::mutex> lock(dev->mu);
        dev->checkFault();
        auto g = std::make_unique<pg_guide>();
        g->dev = dev;
        g->p = *params;
        const std::vector<float> theta = guideInitParams(params->seed);
        upload(g->theta, theta.data(), theta.size() * 4, dev->stream);
        g->m.reserve(pg::kGParamCount * 4), g->v.reserve(pg::kGParamCount * 4), g->wb.reserve(pg::kGParamCount * 2);
        PG_HIP(hipMemsetAsync(g->m.p, 0, g->m.bytes, dev->stream));
        PG_HIP(hipMemsetAsync(g->v.p, 0, g->v.bytes, dev->stream));
        pg::launchGuideToBf16(dev->stream, g->theta.as<float>(), g->wb.as<uint16_t>());
        g->csum.reserve(16), g->loss.reserve(16), g->grads.reserve(pg::kGParamCount * 4);
        PG_HIP(hipGetLastError());
        PG_HIP(hipStreamSynchronize(dev->stream));
        addChild(dev);
        *out = g.release();
    });
}

void pg_guide_destroy(pg_guide* g) {
    if (!g) return;
    pg_device* dev = g->dev;
    {
        std::lock_guard<std::mutex> lock(dev->mu);
        (void)hipStreamSynchronize(dev->stream);
        delete g;
    }
    releaseChild(dev);
}

uint32_t pg_guide_param_count(void) { return pg::kGParamCount; }

int pg_guide_get_params(pg_guide* g, float* theta, uint32_t n) {
    return guarded([&] {
        require(g && theta && n == pg::kGParamCount, "bad arguments");
        std::lock_guard<std::mutex> lock(g->dev->mu);
        g->dev->checkFault();
        PG_HIP(hipStreamSynchronize(g->dev->stream));
        PG_HIP(hipMemcpy(theta, g->theta.p, n * 4, hipMemcpyDeviceToHost));
    });
}

int pg_guide_set_params(pg_guide* g, const float* theta, uint32_t n) {
    return guarded([&] {
        require(g && theta && n == pg::kGParamCount, "bad arguments");
        std::lock_guard<std::mutex> lock(g->dev->mu);
        g->dev->checkFault();
        hipStream_t st = g->dev->stream;
        upload(g->theta, theta, n * 4, st);
        PG_HIP(hipMemsetAsync(g->m.p, 0, g->m.bytes, st));
        PG_HIP(hipMemsetAsync(g->v.p, 0, g->v.bytes, st));
        pg::launchGuideToBf16(st, g->theta.as<float>(), g->wb.as<uint16_t>());
        g->step = 0;
        PG_HIP(hipGetLastError());
        PG_HIP(hipStreamSynchronize(st));
    });
}

int pg_guide_infer(pg_guide* g, const float* in, uint32_t n, float* mix) {
    return guarded([&] {
        require(g && (n == 0 || (in && mix)), "null argument");
        if (n == 0) return;
        std::lock_guard<std::mutex> lock(g->dev->mu);
        g->dev->checkFault();
        hipStream_t st = g->dev->stream;
        DevBuf din, dmix;
        upload(din, in, static_cast<size_t>(n) * 48, st);
        dmix.reserve(static_cast<size_t>(n) * pg::kGMixFloatsHost * 4);
        guideSetLds();
        pg::launchGuideInferArray(st, din.as<float>(), n, g->wb.as<uint16_t>(), g->theta.as<float>(), dmix.as<float>(),
                                  static_cast<uint32_t>(g->dev->numCUs) * 2);
        PG_HIP(hipGetLastError());
        PG_HIP(hipStreamSynchronize(st));
        PG_HIP(hipMemcpy(mix, dmix.p, static_cast<size_t>(n) * pg::kGMixFloatsHost * 4, hipMemcpyDeviceToHost));
    });
}

int pg_guide_gradients(pg_guide* g, const float* rec, uint32_t n, float* loss, float* grads) {
    return guarded([&] {
        require(g && rec && n > 0 && grads, "bad arguments");
        std::lock_guard<std::mutex> lock(g->dev->mu);
        g->dev->checkFault();
        hipStream_t st = g->dev->stream;
        DevBuf drec;
        upload(drec, rec, static_cast<size_t>(n) * 64, st);
        guideStep(g, drec.as<float>(), n, false);
        PG_HIP(hipStreamSynchronize(st));
        PG_HIP(hipMemcpy(grads, g->grads.p, pg::kGParamCount * 4, hipMemcpyDeviceToHost));
        if (loss) PG_HIP(hipMemcpy(loss, g->loss.p, 4, hipMemcpyDeviceToHost));
    });
}

int pg_guide_train(pg_guide* g, const float* rec, uint32_t n, float* loss) {
    return guarded([&] {
        require(g && rec && n > 0, "bad arguments");
        std::lock_guard<std::mutex> lock(g->dev->mu);
        g->dev->checkFault();
        hipStream_t st = g->dev->stream;
        DevBuf drec;
        upload(drec, rec, static_cast<size_t>(n) * 64, st);
        float last = 0.0f;
        for (uint32_t b = 0; b < n; b += g->p.train_batch) {
            const uint32_t cnt = std::min(g->p.train_batch, n - b);
            guideStep(g, drec.as<float>() + static_cast<size_t>(b) * 16, cnt, true);
        }
        PG_HIP(hipStreamSynchronize(st));
        PG_HIP(hipMemcpy(&last, g->loss.p, 4, hipMemcpyDeviceToHost));
        if (loss) *loss = last;
    });
}

int pg_sobol_matrices(uint32_t* out) {
    return guarded([&] {
        require(out != nullptr, "null output");
        std::vector<uint32_t> m;
        std::vector<uint64_t> inv, delta;
        pg::buildSobolTables(m, inv, delta);
        std::memcpy(out, m.data(), m.size() * 4);
    });
}

}  // extern "C"

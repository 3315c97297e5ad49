/* piper_hip.h - C ABI of the MI355X (gfx950) path-tracing library libpiper_hip.so.
 *
 * This is the drop-in boundary for Piper2's GPU render path. Piper2 renders through two plugin
 * interfaces, `Tracer` (Src/Interface/BuiltinComponent/Tracer.hpp:140-159) and `Accelerator`
 * (Src/Interface/Infrastructure/Accelerator.hpp:130-161); the C++ plugin pair in
 * piper2_amd/plugin/ implements those interfaces on top of the entry points below, so every
 * entry point names the reference interface it replaces.
 *
 * Conventions: every call returns PG_OK (0) or a negative PG_ERR_* code and never throws.
 * pg_last_error() returns a thread-local message for the last failure. The library owns device
 * memory behind opaque handles; the caller owns every host array it passes. One HIP stream per
 * device handle; calls on one handle are serialised by an internal mutex (the reference admits
 * one live TraceLauncher per pipeline, Tracer/Embree/Embree.cpp:993-996).
 */
#ifndef PIPER_HIP_H
#define PIPER_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PG_OK 0
#define PG_ERR_INVALID (-1)     /* bad argument or size */
#define PG_ERR_DEVICE (-2)      /* HIP runtime error, or a forced device fault (pg_debug_inject_fault) */
#define PG_ERR_OOM (-3)         /* device allocation failed */
#define PG_ERR_UNSUPPORTED (-4) /* feature or symbol not implemented by this library */

/* Reference-compatibility bits (pg_render_params.flags). Default 0 = corrected ("spec") math. */
#define PG_COMPAT_D1 0x1u  /* light CDF remap with the upper bound (Kernel/Sampling.hpp:61) */
#define PG_COMPAT_D3 0x2u  /* unnormalised shading tangents (Tracer/Embree/Kernel.cpp:180-186) */
#define PG_COMPAT_D9 0x4u  /* u2 drawn before u1 (Tracer/Embree/Kernel.cpp:117,145, MS ABI order) */
#define PG_COMPAT_D13 0x8u /* light pass-through decrements depth (PathTracing/Kernel.cpp:143) */
#define PG_FILM_FP64 0x100u /* accumulate into the film's FP64 RGBW buffer instead of FP32 */
#define PG_COUNT_TRAVERSAL 0x200u /* count BVH nodes visited / triangles tested by extension rays (slower) */
#define PG_NO_STAGE_TIMING 0x400u /* stats without per-stage device timing (no events around launches) */
#define PG_USE_GRAPH 0x800u /* capture the bounce loop's kernels once per call, replay them as hipGraph
                               launches (implies PG_NO_STAGE_TIMING) */

#define PG_FILTER_BOX 0u      /* weight 1 */
#define PG_FILTER_TRIANGLE 1u /* max(0, 1 - |dx|) max(0, 1 - |dy|) */
#define PG_FILTER_GAUSSIAN 2u /* max(exp(-a dx^2) - exp(-a), 0) max(exp(-a dy^2) - exp(-a), 0) */

#define PG_MAT_LAMBERT 0u   /* matte (Surface/Classic/Kernel.cpp:94-269): roughness = alpha in [0, 1];
                               0 gives Lambert, > 0 Oren-Nayar with sigma = roughness x pi */
#define PG_MAT_CONDUCTOR 1u /* GGX conductor, Schlick F0 (new: the reference has no conductor) */
#define PG_MAT_PLASTIC 2u   /* plastic (Surface/Classic/Kernel.cpp:435-473): Lambert diffuse = albedo plus
                               GGX with dielectric Fresnel (eta 1.46), specular = reserved[0..2],
                               roughness = alpha (remapRoughness applied, :378-381) */
#define PG_MAT_MIRROR 3u    /* mirror (Classic/Kernel.cpp:663-692): perfect reflection scaled by albedo */
#define PG_MAT_GLASS 4u     /* glass (glassInit :597-656, eta 1.5): reflection = albedo, transmission =
                               reserved[0..2], roughness = alpha: 0 gives smooth glass (FresnelSpecular
                               :499-541), > 0 microfacet reflection + transmission (remapRoughness) */
#define PG_MAT_SUBSTRATE 5u /* substrate (FresnelBlend, Classic/Kernel.cpp:694-779): diffuse = albedo,
                               specular = reserved[0..2], roughness = alpha (remapRoughness applied) */

#define PG_HIT_MISS 0xFFFFFFFFu
#define PG_HIT_QUAD 0x80000000u /* prim = PG_HIT_QUAD | quad index for a light-quad hit */

typedef struct pg_device pg_device;
typedef struct pg_scene pg_scene;
typedef struct pg_film pg_film;

typedef struct pg_material {
    uint32_t type;      /* PG_MAT_* */
    float albedo[3];    /* diffuse rho (ConstantData, ProceduralTexture/Kernel.cpp:21-26), conductor F0, or
                           mirror / glass reflection */
    float alpha;        /* matte: roughness (0 = Lambert); conductor: GGX alpha = roughness^2; plastic and
                           substrate: roughness; mirror, glass: 0 */
    float reserved[3];  /* plastic, substrate: specular rgb; glass: transmission rgb */
} pg_material;

typedef struct pg_quad { /* Plane primitive of an area light (BasicShapes/Shared.hpp:22-28, AreaLight/Shared.hpp:22-26) */
    float origin[3], u[3], v[3];
    float radiance[3];
} pg_quad;

#define PG_TEX_NONE 0xFFFFFFFFu  /* a material input without a texture: its constant is used */
#define PG_TEX_CHECKERBOARD 1u   /* CheckBoard (ProceduralTexture/Kernel.cpp:28-38): White where the
                                    parities of (u, v) scaled by Scale agree, Black elsewhere */
#define PG_TEX_IMAGE 2u          /* SampledTexture (SampledTexture/Kernel.cpp:37-82): 8-bit texels
                                    scaled by 1/255, bilinear, Repeat or Mirror wrap */
#define PG_WRAP_REPEAT 0u
#define PG_WRAP_MIRROR 1u
typedef struct pg_texture {
    uint32_t type;         /* PG_TEX_CHECKERBOARD or PG_TEX_IMAGE */
    uint32_t channels;     /* 1, 2 or 4 */
    float black[4], white[4];
    float scale;           /* checkerboard */
    uint32_t width, height, wrap; /* image: texels hold height rows of width * channels bytes */
    const uint8_t* texels;
} pg_texture;
/* Texture of each input of a material (PG_TEX_NONE or an index into pg_scene_desc.textures),
 * sampled at the hit's texture coordinate and replacing the constant there. Colour inputs need a
 * texture of at least 3 channels; the roughness input reads channel 0 and is a roughness as the
 * surface's config gives it (matte: 0 Lambert, > 0 Oren-Nayar; conductor: alpha = roughness^2;
 * plastic, substrate: remapRoughness; glass: 0 smooth, > 0 rough). */
typedef struct pg_material_textures {
    uint32_t albedo;   /* replaces pg_material.albedo */
    uint32_t reserved; /* replaces pg_material.reserved (plastic / substrate specular, glass transmission) */
    uint32_t alpha;    /* the roughness input (not for mirrors) */
} pg_material_textures;

typedef struct pg_point_light { /* PointLightData (Light/BasicLight/Shared.hpp:26-29): a delta light */
    float position[3];  /* the node's world position (pointInit, Light/BasicLight/Kernel.cpp:27-33) */
    float intensity[3]; /* radiant intensity; radiance at distance d is intensity / d^2 (:35-45) */
} pg_point_light;

typedef struct pg_camera { /* PCData (Sensor/PerspectiveCamera/Shared.hpp:22-29) plus the node position */
    float position[3], look_at[3], up[3];
    float sensor_size[2];  /* metres */
    float focal_length;    /* metres */
    float aperture_radius; /* metres; focal_length / (2 FStop), 0 = pinhole */
} pg_camera;

typedef struct pg_scene_desc {
    /* Triangle mesh buffer as TriangleIndexedGeometryDesc (Interface/BuiltinComponent/Tracer.hpp:51-63):
     * uint3 index[tri] + float3 vertex[vert]; instances are passed flattened to world space. */
    const float* vertices;
    uint32_t num_vertices;
    const uint32_t* indices;
    uint32_t num_triangles;
    const uint32_t* tri_object;      /* object id per triangle */
    const uint32_t* object_material; /* material id per object */
    uint32_t num_objects;
    const pg_material* materials;
    uint32_t num_materials;
    const pg_quad* quads; /* primitives of the area lights; light 0 is the environment (Embree.cpp:1175-1188) */
    uint32_t num_quads;
    pg_camera camera;
    float env_radiance[3]; /* constant environment (BasicLight/Kernel.cpp:59-89) */
    /* point lights; lights are selected uniformly over {environment, area lights, point lights} in
     * that order. Zero / NULL for none. */
    const pg_point_light* point_lights;
    uint32_t num_point_lights;
    /* textures (optional): per-vertex texture coordinates (2 floats per vertex; NULL gives (0, 0)
     * everywhere, calcTriangleMeshSurface, Tracer/Embree/Kernel.cpp:170-192), the textures, and one
     * pg_material_textures per material (NULL: no textured material) */
    const float* texcoords;
    const pg_texture* textures;
    uint32_t num_textures;
    const pg_material_textures* material_textures;
    /* area lights of several primitives (DiffuseAreaLight over a Plane with several Primitives,
     * BasicShapes.cpp:59-94): light l is the next area_light_quads[l] quads, which share one radiance;
     * the counts cover every quad. A light picks its primitive by area (select, Kernel/Sampling.hpp:52-63)
     * and its density is over the whole light's area. NULL: every quad is a light of its own. */
    const uint32_t* area_light_quads;
    uint32_t num_area_lights;
    /* anisotropic roughness (Glass and Substrate take RoughnessX and RoughnessY, Classic.cpp:98-117,
     * 177-195, Classic/Kernel.cpp:630-655, 769-773; pg_material.alpha and the alpha texture are
     * RoughnessX): per material its RoughnessY constant (NULL: pg_material.alpha) and a texture for it
     * (PG_TEX_NONE: the constant; NULL: none). Both NULL: every RoughnessY is its RoughnessX, texture
     * included. Glass is smooth only where both roughnesses are zero. Other surfaces have one
     * roughness: their RoughnessY must equal pg_material.alpha, with no texture. */
    const float* roughness_y;
    const uint32_t* roughness_y_textures;
} pg_scene_desc;

#define PG_BVH_HOST 0u   /* binned SAH on the host's cores (multithreaded) */
#define PG_BVH_DEVICE 1u /* binned SAH on the GPU (bins capped at 32) */
typedef struct pg_bvh_params {
    uint32_t width;     /* 4 or 8 children per node */
    uint32_t leaf_size; /* max triangles per leaf, 1..8 */
    uint32_t sah_bins;  /* 8..64 */
    uint32_t builder;   /* PG_BVH_HOST or PG_BVH_DEVICE */
} pg_bvh_params;

typedef struct pg_render_params {
    uint32_t width, height;      /* film resolution */
    uint32_t spp;                /* samples per pixel; rounded up to a power of two (ScrambledSobolSampler.cpp:43-50) */
    uint32_t rect[4];            /* render rect x, y, w, h (RenderRECT) */
    uint32_t spp_begin, spp_end; /* sample range [begin, end) rendered by this call */
    uint32_t max_depth, min_rr;  /* MaxTraceDepth, MinRussianRouletteDepth (PathTracing.cpp:38-39) */
    uint32_t flags;              /* PG_COMPAT_*, PG_FILM_FP64 */
    float ndc[4];                /* SensorNDCAffineTransform ox, oy, sx, sy (Sensor.hpp:38-74) */
    uint32_t wavefront_paths;    /* paths in flight per wave of work; 0 = library default */
    uint32_t filter;             /* PG_FILTER_*: the FixedSampler film filter (FixedSampler/Kernel.cpp:55-70) */
    float filter_alpha;          /* PG_FILTER_GAUSSIAN: Alpha (GaussianFilterData, FixedSampler.cpp:40-44) */
    uint32_t reserved;
} pg_render_params;

typedef struct pg_stats {
    uint64_t paths;
    uint64_t extension_rays, shadow_rays, mis_rays;
    uint64_t extend_launches;
    double ms_total;   /* host wall time of the call, stream synchronised */
    double ms_raygen, ms_extend, ms_shade, ms_shadow, ms_mis, ms_film; /* HIP-event time per stage */
    uint64_t nodes_visited, tris_tested; /* extension rays, with PG_COUNT_TRAVERSAL only */
    double ms_guide, ms_train;           /* guided renders: network inference, online training */
    uint64_t train_records;              /* guided renders: records trained on after the render */
} pg_stats;

/* ---- device (Accelerator: enumerateContexts / getNativePlatform, Accelerator.hpp:130-161) ---- */
int pg_device_create(int ordinal, pg_device** out);
void pg_device_destroy(pg_device* dev);
const char* pg_device_platform(pg_device* dev); /* "AMD HIP gfx950" (getNativePlatform) */
const char* pg_last_error(void);

/* ---- scene (Tracer::buildAcceleration / buildNode, Tracer.hpp:140-159; Embree.cpp:497-655) ---- */
/* Copies the scene to the device and builds its BVH there (width 8, leaves of <= 4, 32 bins,
 * PG_BVH_DEVICE); pg_scene_build_bvh rebuilds it with other parameters or on the host. */
int pg_scene_create(pg_device* dev, const pg_scene_desc* desc, pg_scene** out);
int pg_scene_build_bvh(pg_scene* scene, const pg_bvh_params* params);

/* ---- moving instances (Group children with keyframed transforms: EmbreeBranchNode and
 *      updateTransformSRT, Embree.cpp:706-835; ray time = sample dimension 2, Embree/Kernel.cpp:77) ---- */
#define PG_MAX_MOTION_KEYS 8
/* A scene whose triangles [instance_first[k], instance_first[k + 1]) (the last range ends at
 * desc->num_triangles) are moving instance k's mesh in the instance's own space; triangles before
 * instance_first[0] are static and in world space. instance_first is strictly increasing, at most
 * 256 instances. Each instance gets a BVH of its own; the ray meets it in its space at the ray's
 * time. Until pg_scene_set_motion is called an instance is drawn with its own coordinates. */
int pg_scene_create_instanced(pg_device* dev, const pg_scene_desc* desc, const uint32_t* instance_first,
                              uint32_t num_instances, pg_scene** out);
/* Transform of one instance for the next renders, as the reference's updateTransformSRT hands it
 * to Embree (rtcSetGeometryTimeRange / TimeStepCount / TransformQuaternion, Embree.cpp:756-772):
 * num_keys transforms at equal time steps from t0 to t1 (times normalised to the shutter, so a
 * sample's time lies in [0, 1]); the instance is visible for t0 <= time <= t1 (num_keys = 1: a
 * fixed transform, always visible). Between two keys scale, skew, shift and translation
 * interpolate linearly and the rotation spherically. keys: TransformSRT order (Kernel/Transform.hpp
 * :390-393): scale xyz, skew xy xz yz, shift xyz, quaternion w x y z, translation xyz; the matrix
 * is T R S. parent: row-major 3x4 static transform of the instance's parent nodes, applied after. */
typedef struct pg_motion {
    uint32_t num_keys;
    uint32_t visible; /* 0: absent from the frame (rtcDisableGeometry, Embree.cpp:718-724) */
    float t0, t1;
    float keys[PG_MAX_MOTION_KEYS][16];
    float parent[12];
} pg_motion;
int pg_scene_set_motion(pg_scene* scene, uint32_t instance, const pg_motion* motion);
void pg_scene_destroy(pg_scene* scene);
/* BVH shape after a build: {nodes, leaves, max depth, width, node bytes, triangle bytes}. */
int pg_scene_bvh_info(pg_scene* scene, uint64_t info[6]);
/* Copy the built BVH to the host: nodes (node bytes each), triangles (48 B each: v0,prim | v1 | v2). */
int pg_scene_bvh_download(pg_scene* scene, void* nodes, uint64_t node_bytes, void* tris, uint64_t tri_bytes);

/* ---- film (FixedSampler RGBW film, RenderDriver/FixedSampler/Kernel.cpp:22-52) ---- */
int pg_film_create(pg_device* dev, uint32_t width, uint32_t height, pg_film** out);
int pg_film_clear(pg_film* film);
void pg_film_destroy(pg_film* film);
/* rgbw: float[4*w*h] from the FP32 buffer; rgbw64: double[4*w*h] from the FP64 buffer (either may be NULL). */
int pg_film_download(pg_film* film, float* rgbw, double* rgbw64);

/* ---- render (TraceLauncher::launch + RenderDriver::renderFrame, Tracer.hpp:100-119,
 *      FixedSampler.cpp:68-133; per-sample kernel piperMain, Tracer/Embree/Kernel.cpp:47-100) ---- */
int pg_render(pg_scene* scene, const pg_render_params* params, pg_film* film, pg_stats* stats);
/* Same render, but also writes each path's radiance (rgb float) to path_L (host), path order
 * ((py-rect.y)*rect.w + (px-rect.x))*(spp_end-spp_begin) + (s-spp_begin). Parity aid. */
int pg_render_paths(pg_scene* scene, const pg_render_params* params, pg_film* film, float* path_L,
                    pg_stats* stats);
/* Batched render (no reference counterpart: the reference renders one pipeline at a time under a
 * lock, Tracer/Embree/Embree.cpp:993-996; it replaces n successive RenderDriver::renderFrame
 * calls, FixedSampler.cpp:68-133). The paths of n scenes (own resolutions, rects, sample
 * ranges, depths and flags) run through one wavefront pipeline; each film ends up exactly as
 * pg_render(scenes[i], &params[i], films[i]) leaves it. All scenes live on `dev` and share one
 * BVH width; films[i] may be NULL. latency_ms (optional, n entries): device time from the start
 * of the batch until scene i's film is complete. */
int pg_render_batch(pg_device* dev, pg_scene* const* scenes, const pg_render_params* params, uint32_t n,
                    pg_film* const* films, double* latency_ms, pg_stats* stats);

/* Launch settings of a scene's renders (no choice changes any result): the traversal stack
 * entries each lane keeps in LDS (4..16, default 10; deeper entries spill to device memory), the
 * traversal waves per CU at most (0, the default: as many as the device keeps resident), the shade
 * block (128 or 256 threads, default 256) and the waves per guide-inference block (4, 8 or 16,
 * default 8; guided batches always use 8). pg_autotune sets them; pg_scene_set_launch sets them
 * directly. */
typedef struct pg_launch_config {
    uint32_t lds_stack;
    uint32_t trav_waves;
    uint32_t shade_block;
    uint32_t infer_waves;
} pg_launch_config;
int pg_scene_set_launch(pg_scene* scene, const pg_launch_config* config);
int pg_scene_get_launch(pg_scene* scene, pg_launch_config* config);

/* Autotuner (no reference counterpart: the reference has no tunable launch shape). Times a slice
 * of `params` (the whole rect, the first samples of its range, about sample_paths paths) at BVH
 * widths 4 and 8 and path-pool capacities 2^21..2^24, then, one after another, the launch settings:
 * traversal LDS stack depth {6, 8, 10, 12, 16} x traversal waves per CU {resident, 16, 20, 24}, the
 * shade block {128, 256} and the waves per inference block {4, 8, 16} (on guided renders with a
 * fresh guide). Leaves the scene built at the fastest width with the fastest launch settings and
 * returns the fastest capacity; pass it as params->wavefront_paths. The films are untouched. */
typedef struct pg_tune_result {
    uint32_t bvh_width;        /* the scene's BVH width from now on */
    uint32_t wavefront_paths;  /* fastest path-pool capacity */
    uint64_t sample_paths;     /* paths per timed trial */
    uint32_t trials;
    uint32_t reserved;
    double mpaths_per_s;       /* throughput of the best trial */
    pg_launch_config launch;   /* the launch settings the scene now has */
} pg_tune_result;
int pg_autotune(pg_scene* scene, const pg_render_params* params, uint64_t sample_paths, pg_tune_result* out);

/* ---- parity entry points (Embree rtcIntersect1 / rtcOccluded1, Embree.cpp:300,384; sampler
 *      sobolStart / sobolGenerate, ScrambledSobolSampler/Kernel.cpp:52-67) ---- */
/* rays: float[8*n] = {o xyz, d xyz, tmin, tmax}; outputs t[n], prim[n], uv[2n] (host arrays). */
int pg_trace(pg_scene* scene, const float* rays, uint32_t n, float* t, uint32_t* prim, float* uv);
/* pg_trace at ray times times[n] (RTCRay::time, Embree.cpp:290), for scenes with moving instances;
 * pg_trace traces at time 0. */
int pg_trace_at(pg_scene* scene, const float* rays, const float* times, uint32_t n, float* t, uint32_t* prim,
                float* uv);
int pg_occluded(pg_scene* scene, const float* rays, uint32_t n, uint8_t* occluded);
/* pg_trace without any allocation: no counterpart in the reference. For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). rays, t, prim and uv are DEVICE pointers on the scene's GPU, laid out as in
 * pg_trace; the results equal pg_trace's. The traversal stack entries that overflow on-chip
 * memory spill to a fixed 16 MiB array in device memory (one per GPU, held until the process
 * ends) instead of a per-call buffer; a BVH so deep that one wave's spill exceeds it is rejected
 * with PG_ERR_INVALID. stream: a hipStream_t of the scene's GPU, the call then only enqueues the
 * work; NULL runs it on the device handle's stream and waits for it. All calls on one GPU share
 * that array: calls that overlap in time on one GPU race. */
int pg_trace_noalloc(pg_scene* scene, const float* rays, uint32_t n, float* t, uint32_t* prim, float* uv,
                     void* stream);
/* Camera rays for (px, py, sample) triples as piperMain makes them; rays as in pg_trace, film[2n]. */
int pg_camera_rays(pg_scene* scene, const pg_render_params* params, const uint32_t* pix, uint32_t n,
                   float* rays, float* film);
/* Sobol values: out[n*dims] for samples (px, py, s) at resolution 2^m: dims 0..dims-1 of index
 * sobolIntervalToIndex(m, s, px, py); index_out[n] (optional) receives the indices. */
int pg_sobol(pg_device* dev, uint32_t m, uint32_t dims, const uint32_t* pix, uint32_t n, float* out,
             uint64_t* index_out);
/* Generator matrices, uint32[1024*52] (the reference's SobolMatrix.dat layout, Shared.hpp:22-23). */
int pg_sobol_matrices(uint32_t* out);

/* ---- online neural path guiding (no reference counterpart: the reference has no guiding) ----
 * A 144-64-64-64-64-40 MLP (BF16 weights on the matrix cores, FP32 accumulation) maps a hit's
 * position (scene box -> [0,1]^3), shading normal and outgoing direction to an 8-lobe von
 * Mises-Fisher mixture, sampled with probability mis_fraction next to the BSDF (one-sample MIS). */
typedef struct pg_guide pg_guide;
typedef struct pg_guide_params {
    uint32_t seed;           /* weight initialisation */
    float lr, beta1, beta2, eps; /* Adam */
    float mis_fraction;      /* probability of drawing the continuation from the guide; negative: the
                                network learns it per vertex (an extra output in [0.05, 0.95], trained
                                with the combined density mis_fraction p_guide + (1 - mis_fraction) p_bsdf) */
    uint32_t train_batch;    /* records per Adam step (>= 64) */
    uint32_t train_records;  /* records a guided render collects for training (0: none) */
} pg_guide_params;
int pg_guide_create(pg_device* dev, const pg_guide_params* params, pg_guide** out);
void pg_guide_destroy(pg_guide* guide);
uint32_t pg_guide_param_count(void); /* 24880 FP32 parameters: per layer W[out][in] then b[out] */
int pg_guide_get_params(pg_guide* guide, float* theta, uint32_t n);
int pg_guide_set_params(pg_guide* guide, const float* theta, uint32_t n); /* also resets Adam */
/* in[12 n] = per sample (pos01 xyz, -, normal xyz, -, wo xyz, -); mix[52 n] = 8 x (mu xyz, kappa),
 * then 8 lobe weights pi, then 8 x pi * kappa / (2 pi (1 - exp(-2 kappa))), then the learned
 * selection probability and three zeros. */
int pg_guide_infer(pg_guide* guide, const float* in, uint32_t n, float* mix);
/* Training records rec[16 n] = (pos01 xyz, weight c, normal, BSDF pdf, wo, -, sampled direction, -).
 * Loss = -sum c log p(direction) / sum c, p the guide's density, or with a learned mis_fraction the
 * combined density alpha p_guide + (1 - alpha) p_bsdf with alpha the network's selection output.
 * gradients: FP32 dLoss/dtheta, no update. */
int pg_guide_gradients(pg_guide* guide, const float* rec, uint32_t n, float* loss, float* grads);
/* Adam steps over rec in batches of train_batch; *loss = loss of the last batch. */
int pg_guide_train(pg_guide* guide, const float* rec, uint32_t n, float* loss);
/* pg_render with guided continuation sampling (one-sample MIS between the guide mixture, with
 * probability mis_fraction, and the BSDF). With train_records > 0 the render also collects that
 * many training records from its own paths and takes Adam steps on them after the render, so a
 * progressive driver trains online pass by pass while every pass stays unbiased. */
int pg_render_guided(pg_scene* scene, pg_guide* guide, const pg_render_params* params, pg_film* film,
                     pg_stats* stats);
/* pg_render_batch with guided sampling: scene i uses guides[i] (scenes may share a guide; at most
 * 256 distinct guides). One persistent path pool runs every scene; each bounce groups the
 * extension queue by guide so every inference block runs one guide's weights. Each guide with
 * train_records > 0 trains after the batch on the records of its own scenes. With frozen guides
 * (train_records = 0) each film is exactly what pg_render_guided(scenes[i], guides[i], ...)
 * leaves. No reference counterpart (the reference has no guiding and no batching). */
int pg_render_batch_guided(pg_device* dev, pg_scene* const* scenes, pg_guide* const* guides,
                           const pg_render_params* params, uint32_t n, pg_film* const* films, double* latency_ms,
                           pg_stats* stats);

/* ---- diagnostics ---- */
/* Make the next device call on this handle fail with PG_ERR_DEVICE (host-forced fault path test). */
int pg_debug_inject_fault(pg_device* dev);
/* Run a plain device copy kernel of `bytes` (read + write) `reps` times; *gbps receives the
 * (read + write) bandwidth. Calibrates HBM byte counters; no reference counterpart. */
int pg_debug_copy(pg_device* dev, uint64_t bytes, uint32_t reps, double* gbps);

#ifdef __cplusplus
}
#endif

#endif /* PIPER_HIP_H */

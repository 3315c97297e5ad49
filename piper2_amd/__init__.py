"""piper2_amd: MI355X (gfx950) path-tracing hot path of Piper2, behind a C ABI.

The product is libpiper_hip.so (include/piper_hip.h). This module is a thin ctypes layer over it
for tests and benchmarks; it fails loudly when the library has not been built, and it never falls
back to any CPU implementation.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpiper_hip.so")

PG_OK = 0
PG_ERR_INVALID = -1
PG_ERR_DEVICE = -2
PG_ERR_OOM = -3
PG_ERR_UNSUPPORTED = -4

PG_COMPAT_D1 = 0x1
PG_COMPAT_D3 = 0x2
PG_COMPAT_D9 = 0x4
PG_COMPAT_D13 = 0x8
PG_FILM_FP64 = 0x100
PG_COUNT_TRAVERSAL = 0x200
PG_NO_STAGE_TIMING = 0x400
PG_USE_GRAPH = 0x800
PG_BVH_HOST = 0
PG_BVH_DEVICE = 1
PG_FILTER_BOX = 0
PG_FILTER_TRIANGLE = 1
PG_FILTER_GAUSSIAN = 2

PG_MAT_LAMBERT = 0
PG_MAT_CONDUCTOR = 1
PG_MAT_PLASTIC = 2
PG_MAT_MIRROR = 3
PG_MAT_GLASS = 4
PG_MAT_SUBSTRATE = 5
PG_TEX_NONE = 0xFFFFFFFF
PG_TEX_CHECKERBOARD = 1
PG_TEX_IMAGE = 2
PG_WRAP_REPEAT = 0
PG_WRAP_MIRROR = 1
PG_HIT_MISS = 0xFFFFFFFF
PG_HIT_QUAD = 0x80000000

F3 = ctypes.c_float * 3


class Material(ctypes.Structure):
    _fields_ = [("type", ctypes.c_uint32), ("albedo", F3), ("alpha", ctypes.c_float), ("reserved", F3)]


class Quad(ctypes.Structure):
    _fields_ = [("origin", F3), ("u", F3), ("v", F3), ("radiance", F3)]


class Camera(ctypes.Structure):
    _fields_ = [("position", F3), ("look_at", F3), ("up", F3), ("sensor_size", ctypes.c_float * 2),
                ("focal_length", ctypes.c_float), ("aperture_radius", ctypes.c_float)]


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("vertices", ctypes.c_void_p), ("num_vertices", ctypes.c_uint32),
        ("indices", ctypes.c_void_p), ("num_triangles", ctypes.c_uint32),
        ("tri_object", ctypes.c_void_p), ("object_material", ctypes.c_void_p), ("num_objects", ctypes.c_uint32),
        ("materials", ctypes.c_void_p), ("num_materials", ctypes.c_uint32),
        ("quads", ctypes.c_void_p), ("num_quads", ctypes.c_uint32),
        ("camera", Camera), ("env_radiance", F3),
        ("point_lights", ctypes.c_void_p), ("num_point_lights", ctypes.c_uint32),
        ("texcoords", ctypes.c_void_p), ("textures", ctypes.c_void_p), ("num_textures", ctypes.c_uint32),
        ("material_textures", ctypes.c_void_p),
        ("area_light_quads", ctypes.c_void_p), ("num_area_lights", ctypes.c_uint32),
        ("roughness_y", ctypes.c_void_p), ("roughness_y_textures", ctypes.c_void_p),
    ]


class Texture(ctypes.Structure):  # pg_texture
    _fields_ = [("type", ctypes.c_uint32), ("channels", ctypes.c_uint32), ("black", ctypes.c_float * 4),
                ("white", ctypes.c_float * 4), ("scale", ctypes.c_float), ("width", ctypes.c_uint32),
                ("height", ctypes.c_uint32), ("wrap", ctypes.c_uint32), ("texels", ctypes.c_void_p)]


class BvhParams(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("leaf_size", ctypes.c_uint32), ("sah_bins", ctypes.c_uint32),
                ("builder", ctypes.c_uint32)]


class RenderParams(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("spp", ctypes.c_uint32),
        ("rect", ctypes.c_uint32 * 4), ("spp_begin", ctypes.c_uint32), ("spp_end", ctypes.c_uint32),
        ("max_depth", ctypes.c_uint32), ("min_rr", ctypes.c_uint32), ("flags", ctypes.c_uint32),
        ("ndc", ctypes.c_float * 4), ("wavefront_paths", ctypes.c_uint32), ("filter", ctypes.c_uint32),
        ("filter_alpha", ctypes.c_float), ("reserved", ctypes.c_uint32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("paths", ctypes.c_uint64), ("extension_rays", ctypes.c_uint64), ("shadow_rays", ctypes.c_uint64),
        ("mis_rays", ctypes.c_uint64), ("extend_launches", ctypes.c_uint64),
        ("ms_total", ctypes.c_double), ("ms_raygen", ctypes.c_double), ("ms_extend", ctypes.c_double),
        ("ms_shade", ctypes.c_double), ("ms_shadow", ctypes.c_double), ("ms_mis", ctypes.c_double),
        ("ms_film", ctypes.c_double), ("nodes_visited", ctypes.c_uint64), ("tris_tested", ctypes.c_uint64),
        ("ms_guide", ctypes.c_double), ("ms_train", ctypes.c_double), ("train_records", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class LaunchConfig(ctypes.Structure):
    """pg_launch_config: launch settings of a scene's renders."""
    _fields_ = [("lds_stack", ctypes.c_uint32), ("trav_waves", ctypes.c_uint32), ("shade_block", ctypes.c_uint32),
                ("infer_waves", ctypes.c_uint32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class TuneResult(ctypes.Structure):
    _fields_ = [("bvh_width", ctypes.c_uint32), ("wavefront_paths", ctypes.c_uint32),
                ("sample_paths", ctypes.c_uint64), ("trials", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("mpaths_per_s", ctypes.c_double), ("launch", LaunchConfig)]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_ if name not in ("reserved", "launch")}
        d["launch"] = self.launch.as_dict()
        return d


MAX_MOTION_KEYS = 8


class Motion(ctypes.Structure):
    """pg_motion: transform keys of one moving instance for the next renders."""
    _fields_ = [("num_keys", ctypes.c_uint32), ("visible", ctypes.c_uint32), ("t0", ctypes.c_float),
                ("t1", ctypes.c_float), ("keys", (ctypes.c_float * 16) * MAX_MOTION_KEYS),
                ("parent", ctypes.c_float * 12)]


IDENTITY_SRT = (1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
IDENTITY_3X4 = (1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0)


def motion_struct(keys, t0=0.0, t1=1.0, visible=True, parent=None):
    """pg_motion from a list of 16-float TransformSRT keys (scale xyz, skew xy xz yz, shift xyz,
    quaternion w x y z, translation xyz)."""
    keys = [list(map(float, k)) for k in keys]
    if not 1 <= len(keys) <= MAX_MOTION_KEYS or any(len(k) != 16 for k in keys):
        raise ValueError("a motion has 1 to 8 keys of 16 floats")
    m = Motion()
    m.num_keys, m.visible, m.t0, m.t1 = len(keys), 1 if visible else 0, t0, t1
    for i, k in enumerate(keys):
        m.keys[i][:] = k
    m.parent[:] = list(map(float, parent if parent is not None else IDENTITY_3X4))
    return m


# Every symbol the C ABI declares, with (restype, argtypes).
_VP = ctypes.c_void_p
_U32 = ctypes.c_uint32
SIGNATURES = {
    "pg_device_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_VP)]),
    "pg_device_destroy": (None, [_VP]),
    "pg_device_platform": (ctypes.c_char_p, [_VP]),
    "pg_last_error": (ctypes.c_char_p, []),
    "pg_scene_create": (ctypes.c_int, [_VP, ctypes.POINTER(SceneDesc), ctypes.POINTER(_VP)]),
    "pg_scene_build_bvh": (ctypes.c_int, [_VP, ctypes.POINTER(BvhParams)]),
    "pg_scene_create_instanced": (ctypes.c_int, [_VP, ctypes.POINTER(SceneDesc), _VP, _U32, ctypes.POINTER(_VP)]),
    "pg_scene_set_motion": (ctypes.c_int, [_VP, _U32, ctypes.POINTER(Motion)]),
    "pg_scene_set_launch": (ctypes.c_int, [_VP, ctypes.POINTER(LaunchConfig)]),
    "pg_scene_get_launch": (ctypes.c_int, [_VP, ctypes.POINTER(LaunchConfig)]),
    "pg_scene_destroy": (None, [_VP]),
    "pg_scene_bvh_info": (ctypes.c_int, [_VP, _VP]),
    "pg_scene_bvh_download": (ctypes.c_int, [_VP, _VP, ctypes.c_uint64, _VP, ctypes.c_uint64]),
    "pg_film_create": (ctypes.c_int, [_VP, _U32, _U32, ctypes.POINTER(_VP)]),
    "pg_film_clear": (ctypes.c_int, [_VP]),
    "pg_film_destroy": (None, [_VP]),
    "pg_film_download": (ctypes.c_int, [_VP, _VP, _VP]),
    "pg_render": (ctypes.c_int, [_VP, ctypes.POINTER(RenderParams), _VP, ctypes.POINTER(Stats)]),
    "pg_render_paths": (ctypes.c_int, [_VP, ctypes.POINTER(RenderParams), _VP, _VP, ctypes.POINTER(Stats)]),
    "pg_trace": (ctypes.c_int, [_VP, _VP, _U32, _VP, _VP, _VP]),
    "pg_trace_at": (ctypes.c_int, [_VP, _VP, _VP, _U32, _VP, _VP, _VP]),
    "pg_occluded": (ctypes.c_int, [_VP, _VP, _U32, _VP]),
    "pg_trace_noalloc": (ctypes.c_int, [_VP, _VP, _U32, _VP, _VP, _VP, _VP]),
    "pg_camera_rays": (ctypes.c_int, [_VP, ctypes.POINTER(RenderParams), _VP, _U32, _VP, _VP]),
    "pg_sobol": (ctypes.c_int, [_VP, _U32, _U32, _VP, _U32, _VP, _VP]),
    "pg_sobol_matrices": (ctypes.c_int, [_VP]),
    "pg_debug_inject_fault": (ctypes.c_int, [_VP]),
    "pg_debug_copy": (ctypes.c_int, [_VP, ctypes.c_uint64, _U32, ctypes.POINTER(ctypes.c_double)]),
    "pg_guide_create": (ctypes.c_int, [_VP, _VP, ctypes.POINTER(_VP)]),
    "pg_guide_destroy": (None, [_VP]),
    "pg_guide_param_count": (_U32, []),
    "pg_guide_get_params": (ctypes.c_int, [_VP, _VP, _U32]),
    "pg_guide_set_params": (ctypes.c_int, [_VP, _VP, _U32]),
    "pg_guide_infer": (ctypes.c_int, [_VP, _VP, _U32, _VP]),
    "pg_guide_gradients": (ctypes.c_int, [_VP, _VP, _U32, _VP, _VP]),
    "pg_guide_train": (ctypes.c_int, [_VP, _VP, _U32, _VP]),
    "pg_render_guided": (ctypes.c_int, [_VP, _VP, ctypes.POINTER(RenderParams), _VP, ctypes.POINTER(Stats)]),
    "pg_render_batch": (ctypes.c_int, [_VP, _VP, ctypes.POINTER(RenderParams), _U32, _VP, _VP,
                                       ctypes.POINTER(Stats)]),
    "pg_render_batch_guided": (ctypes.c_int, [_VP, _VP, _VP, ctypes.POINTER(RenderParams), _U32, _VP, _VP,
                                              ctypes.POINTER(Stats)]),
    "pg_autotune": (ctypes.c_int, [_VP, ctypes.POINTER(RenderParams), ctypes.c_uint64, ctypes.POINTER(TuneResult)]),
}

_lib = None


class PgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"piper_hip error {code}: {msg}")
        self.code = code


def lib():
    """Load libpiper_hip.so (raises if it has not been built: there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code):
    if code != PG_OK:
        raise PgError(code, lib().pg_last_error().decode())
    return code


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class Device:
    def __init__(self, ordinal=0):
        self.h = ctypes.c_void_p()
        check(lib().pg_device_create(ordinal, ctypes.byref(self.h)))

    @property
    def platform(self):
        return lib().pg_device_platform(self.h).decode()

    def close(self):
        if self.h:
            lib().pg_device_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sobol(self, m, dims, pix):
        pix = np.ascontiguousarray(pix, dtype=np.uint32).reshape(-1, 3)
        n = pix.shape[0]
        out = np.zeros((n, dims), dtype=np.float32)
        idx = np.zeros(n, dtype=np.uint64)
        check(lib().pg_sobol(self.h, m, dims, _ptr(pix), n, _ptr(out), _ptr(idx)))
        return out, idx

    def inject_fault(self):
        check(lib().pg_debug_inject_fault(self.h))

    def copy_bandwidth(self, nbytes=1 << 30, reps=5):
        """GB/s (read + write) of a plain device copy kernel over nbytes."""
        g = ctypes.c_double()
        check(lib().pg_debug_copy(self.h, nbytes, reps, ctypes.byref(g)))
        return g.value


def sobol_matrices():
    out = np.zeros(1024 * 52, dtype=np.uint32)
    check(lib().pg_sobol_matrices(_ptr(out)))
    return out.reshape(1024, 52)


class Scene:
    """A scene on the device (see piper2_amd.scenes for the seeded generators)."""

    def __init__(self, device, sc, bvh_width=8, leaf_size=4, bins=32, builder=None):
        self.device = device
        self._keep = sc
        desc = scene_desc(sc)
        self.h = ctypes.c_void_p()
        first = sc.get("instance_first")
        if first is not None and len(first):  # moving instances (scenes.py: instance_first, motions)
            self._first = np.ascontiguousarray(first, dtype=np.uint32)
            check(lib().pg_scene_create_instanced(device.h, ctypes.byref(desc), _ptr(self._first), len(self._first),
                                                  ctypes.byref(self.h)))
        else:
            check(lib().pg_scene_create(device.h, ctypes.byref(desc), ctypes.byref(self.h)))
        if (bvh_width, leaf_size, bins) != (8, 4, 32) or builder is not None:
            self.build_bvh(bvh_width, leaf_size, bins, builder)
        for k, m in enumerate(sc.get("motions") or []):
            self.set_motion(k, **m)

    def set_launch(self, lds_stack=10, trav_waves=0, shade_block=256, infer_waves=8):
        """Launch settings of this scene's renders (pg_scene_set_launch)."""
        c = LaunchConfig(lds_stack, trav_waves, shade_block, infer_waves)
        check(lib().pg_scene_set_launch(self.h, ctypes.byref(c)))

    def get_launch(self):
        c = LaunchConfig()
        check(lib().pg_scene_get_launch(self.h, ctypes.byref(c)))
        return c.as_dict()

    def set_motion(self, instance, keys, t0=0.0, t1=1.0, visible=True, parent=None):
        """Keys of moving instance `instance` (pg_scene_set_motion)."""
        m = motion_struct(keys, t0, t1, visible, parent)
        check(lib().pg_scene_set_motion(self.h, instance, ctypes.byref(m)))

    def build_bvh(self, width=8, leaf_size=4, bins=32, builder=None):
        """Rebuild the BVH (builder: PG_BVH_DEVICE by default, or PG_BVH_HOST)."""
        p = BvhParams(width, leaf_size, bins, PG_BVH_DEVICE if builder is None else builder)
        check(lib().pg_scene_build_bvh(self.h, ctypes.byref(p)))

    def bvh_info(self):
        info = np.zeros(6, dtype=np.uint64)
        check(lib().pg_scene_bvh_info(self.h, _ptr(info)))
        keys = ["nodes", "leaves", "max_depth", "width", "node_bytes", "tri_bytes"]
        return {k: int(v) for k, v in zip(keys, info)}

    def bvh_download(self):
        info = self.bvh_info()
        nodes = np.zeros(info["nodes"] * info["node_bytes"] // 4, dtype=np.float32)
        tris = np.zeros(self._keep["indices"].shape[0] * 12, dtype=np.float32)
        check(lib().pg_scene_bvh_download(self.h, _ptr(nodes), nodes.nbytes, _ptr(tris), tris.nbytes))
        return nodes, tris

    def close(self):
        if self.h:
            lib().pg_scene_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def trace(self, rays):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = rays.shape[0]
        t = np.zeros(n, np.float32)
        prim = np.zeros(n, np.uint32)
        uv = np.zeros((n, 2), np.float32)
        check(lib().pg_trace(self.h, _ptr(rays), n, _ptr(t), _ptr(prim), _ptr(uv)))
        return t, prim, uv

    def trace_at(self, rays, times):
        """Closest hits at per-ray times (pg_trace_at)."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        times = np.ascontiguousarray(times, dtype=np.float32).reshape(-1)
        n = rays.shape[0]
        if times.shape[0] != n:
            raise ValueError("one time per ray")
        t = np.zeros(n, np.float32)
        prim = np.zeros(n, np.uint32)
        uv = np.zeros((n, 2), np.float32)
        check(lib().pg_trace_at(self.h, _ptr(rays), _ptr(times), n, _ptr(t), _ptr(prim), _ptr(uv)))
        return t, prim, uv

    def occluded(self, rays):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        out = np.zeros(rays.shape[0], np.uint8)
        check(lib().pg_occluded(self.h, _ptr(rays), rays.shape[0], _ptr(out)))
        return out

    def camera_rays(self, params, pix):
        pix = np.ascontiguousarray(pix, dtype=np.uint32).reshape(-1, 3)
        n = pix.shape[0]
        rays = np.zeros((n, 8), np.float32)
        film = np.zeros((n, 2), np.float32)
        check(lib().pg_camera_rays(self.h, ctypes.byref(params), _ptr(pix), n, _ptr(rays), _ptr(film)))
        return rays, film

    def render(self, params, film=None, path_l=False):
        st = Stats()
        if path_l:
            n = params.rect[2] * params.rect[3] * (params.spp_end - params.spp_begin)
            out = np.zeros((n, 3), np.float32)
            check(lib().pg_render_paths(self.h, ctypes.byref(params), film.h if film else None, _ptr(out),
                                        ctypes.byref(st)))
            return st.as_dict(), out
        check(lib().pg_render(self.h, ctypes.byref(params), film.h if film else None, ctypes.byref(st)))
        return st.as_dict()

    def autotune(self, params, sample_paths=1 << 25):
        """pg_autotune: rebuilds the BVH at the fastest width; returns the chosen configuration."""
        out = TuneResult()
        check(lib().pg_autotune(self.h, ctypes.byref(params), sample_paths, ctypes.byref(out)))
        return out.as_dict()

    def render_guided(self, guide, params, film=None):
        st = Stats()
        check(lib().pg_render_guided(self.h, guide.h, ctypes.byref(params), film.h if film else None,
                                     ctypes.byref(st)))
        return st.as_dict()


class GuideParams(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint32), ("lr", ctypes.c_float), ("beta1", ctypes.c_float),
                ("beta2", ctypes.c_float), ("eps", ctypes.c_float), ("mis_fraction", ctypes.c_float),
                ("train_batch", ctypes.c_uint32), ("train_records", ctypes.c_uint32)]


class Guide:
    """Online path-guiding network on the device (pg_guide_*)."""

    def __init__(self, device, seed=0, lr=1e-2, beta1=0.9, beta2=0.99, eps=1e-8, mis_fraction=0.5,
                 train_batch=1 << 14, train_records=1 << 18):
        self.device = device
        self.params = GuideParams(seed, lr, beta1, beta2, eps, mis_fraction, train_batch, train_records)
        self.h = ctypes.c_void_p()
        check(lib().pg_guide_create(device.h, ctypes.byref(self.params), ctypes.byref(self.h)))
        self.n = lib().pg_guide_param_count()

    def get_params(self):
        out = np.zeros(self.n, np.float32)
        check(lib().pg_guide_get_params(self.h, _ptr(out), self.n))
        return out

    def set_params(self, theta):
        theta = np.ascontiguousarray(theta, np.float32)
        check(lib().pg_guide_set_params(self.h, _ptr(theta), self.n))

    def infer(self, pos01, normal, wo):
        n = pos01.shape[0]
        inp = np.zeros((n, 12), np.float32)
        inp[:, 0:3], inp[:, 4:7], inp[:, 8:11] = pos01, normal, wo
        mix = np.zeros((n, 52), np.float32)
        check(lib().pg_guide_infer(self.h, _ptr(inp), n, _ptr(mix)))
        return mix

    @staticmethod
    def records(pos01, normal, wo, direction, c, bsdf_pdf=None):
        """Training records; bsdf_pdf: the BSDF's density of each direction (read with a learned
        mis_fraction)."""
        n = pos01.shape[0]
        rec = np.zeros((n, 16), np.float32)
        rec[:, 0:3], rec[:, 3], rec[:, 4:7], rec[:, 8:11], rec[:, 12:15] = pos01, c, normal, wo, direction
        if bsdf_pdf is not None:
            rec[:, 7] = bsdf_pdf
        return rec

    def gradients(self, rec):
        rec = np.ascontiguousarray(rec, np.float32)
        g = np.zeros(self.n, np.float32)
        loss = np.zeros(1, np.float32)
        check(lib().pg_guide_gradients(self.h, _ptr(rec), rec.shape[0], _ptr(loss), _ptr(g)))
        return float(loss[0]), g

    def train(self, rec):
        rec = np.ascontiguousarray(rec, np.float32)
        loss = np.zeros(1, np.float32)
        check(lib().pg_guide_train(self.h, _ptr(rec), rec.shape[0], _ptr(loss)))
        return float(loss[0])

    def close(self):
        if self.h:
            lib().pg_guide_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Film:
    def __init__(self, device, width, height):
        self.width, self.height = width, height
        self.h = ctypes.c_void_p()
        check(lib().pg_film_create(device.h, width, height, ctypes.byref(self.h)))

    def clear(self):
        check(lib().pg_film_clear(self.h))

    def download(self, fp64=False):
        if fp64:
            out = np.zeros((self.height, self.width, 4), np.float64)
            check(lib().pg_film_download(self.h, None, _ptr(out)))
        else:
            out = np.zeros((self.height, self.width, 4), np.float32)
            check(lib().pg_film_download(self.h, _ptr(out), None))
        return out

    def close(self):
        if self.h:
            lib().pg_film_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def render_batch(device, scenes, params, films=None, guides=None):
    """One batched render of several scenes (pg_render_batch, or pg_render_batch_guided when
    `guides` lists one Guide per scene). films: list of Film or None entries.
    Returns (stats dict, per-scene latency in ms as a numpy array)."""
    n = len(scenes)
    assert len(params) == n and (films is None or len(films) == n) and (guides is None or len(guides) == n)
    hs = (ctypes.c_void_p * n)(*[s.h.value for s in scenes])
    ps = (RenderParams * n)(*params)
    fs = (ctypes.c_void_p * n)(*[(f.h.value if f else None) for f in films]) if films else None
    lat = np.zeros(n, np.float64)
    st = Stats()
    if guides is None:
        check(lib().pg_render_batch(device.h, ctypes.cast(hs, _VP), ps, n, ctypes.cast(fs, _VP) if fs else None,
                                    _ptr(lat), ctypes.byref(st)))
    else:
        gs = (ctypes.c_void_p * n)(*[g.h.value for g in guides])
        check(lib().pg_render_batch_guided(device.h, ctypes.cast(hs, _VP), ctypes.cast(gs, _VP), ps, n,
                                           ctypes.cast(fs, _VP) if fs else None, _ptr(lat), ctypes.byref(st)))
    return st.as_dict(), lat


def scene_desc(sc):
    """Build a SceneDesc that points into the numpy arrays of a generated scene dict."""
    for k in ("vertices", "indices", "tri_object", "object_material", "materials", "quads"):
        assert sc[k].flags["C_CONTIGUOUS"], k
    mats = sc["materials"]  # (n, 8) float32: type, r, g, b, alpha (plastic: roughness), specular rgb
    mat_arr = (Material * len(mats))()
    for i, m in enumerate(mats):
        mat_arr[i] = Material(int(m[0]), F3(*m[1:4]), float(m[4]), F3(*m[5:8]))
    quads = sc["quads"]  # (n, 12)
    quad_arr = (Quad * max(len(quads), 1))()
    for i, q in enumerate(quads):
        quad_arr[i] = Quad(F3(*q[0:3]), F3(*q[3:6]), F3(*q[6:9]), F3(*q[9:12]))
    sc["_mat_arr"], sc["_quad_arr"] = mat_arr, quad_arr
    pts = sc.get("point_lights")  # optional (n, 6) float32: position, intensity (pg_point_light)
    if pts is not None and len(pts):
        pts = sc["_points"] = np.ascontiguousarray(pts, np.float32).reshape(-1, 6)
    else:
        pts = None
    # optional textures: "texcoords" (V, 2) float32, "textures" (list of dicts with the pg_texture
    # fields; "texels" a uint8 array), "material_textures" (M, 3) uint32 (albedo, reserved, roughness)
    texcoords = sc.get("texcoords")
    if texcoords is not None:
        texcoords = sc["_texcoords"] = np.ascontiguousarray(texcoords, np.float32).reshape(-1, 2)
    texs = sc.get("textures") or []
    tex_arr = (Texture * max(len(texs), 1))()
    keep = []
    for i, t in enumerate(texs):
        texels = None
        if t.get("texels") is not None:
            texels = np.ascontiguousarray(t["texels"], np.uint8)
            keep.append(texels)
        tex_arr[i] = Texture(int(t["type"]), int(t["channels"]), (ctypes.c_float * 4)(*t.get("black", [0] * 4)),
                             (ctypes.c_float * 4)(*t.get("white", [0] * 4)), float(t.get("scale", 0.0)),
                             int(t.get("width", 0)), int(t.get("height", 0)), int(t.get("wrap", 0)),
                             _ptr(texels) if texels is not None else None)
    mtex = sc.get("material_textures")
    if mtex is not None:
        mtex = sc["_mtex"] = np.ascontiguousarray(mtex, np.uint32).reshape(-1, 3)
    sc["_tex_arr"], sc["_texels"] = tex_arr, keep
    # optional "area_lights": primitives per area light, in quad order (pg_scene_desc.area_light_quads)
    al = sc.get("area_lights")
    if al is not None:
        al = sc["_area_lights"] = np.ascontiguousarray(al, np.uint32).reshape(-1)
    # optional anisotropic roughness: "roughness_y" (M,) float32 RoughnessY per material (glass and
    # substrate; materials[:, 4] is RoughnessX) and "roughness_y_textures" (M,) uint32
    ry = sc.get("roughness_y")
    if ry is not None:
        ry = sc["_roughness_y"] = np.ascontiguousarray(ry, np.float32).reshape(-1)
    ryt = sc.get("roughness_y_textures")
    if ryt is not None:
        ryt = sc["_roughness_y_textures"] = np.ascontiguousarray(ryt, np.uint32).reshape(-1)
    c = sc["camera"]  # 14 floats: pos, lookAt, up, sizeX, sizeY, focal, aperture, 0
    cam = Camera(F3(*c[0:3]), F3(*c[3:6]), F3(*c[6:9]), (ctypes.c_float * 2)(c[9], c[10]), float(c[11]), float(c[12]))
    return SceneDesc(
        _ptr(sc["vertices"]), sc["vertices"].shape[0], _ptr(sc["indices"]), sc["indices"].shape[0],
        _ptr(sc["tri_object"]), _ptr(sc["object_material"]), sc["object_material"].shape[0],
        ctypes.cast(mat_arr, ctypes.c_void_p), len(mats), ctypes.cast(quad_arr, ctypes.c_void_p), len(quads),
        cam, F3(*sc["env"]), _ptr(pts) if pts is not None else None, 0 if pts is None else pts.shape[0],
        _ptr(texcoords) if texcoords is not None else None, ctypes.cast(tex_arr, ctypes.c_void_p) if texs else None,
        len(texs), _ptr(mtex) if mtex is not None else None, _ptr(al) if al is not None else None,
        0 if al is None else al.shape[0], _ptr(ry) if ry is not None else None,
        _ptr(ryt) if ryt is not None else None)


def render_params(sc, spp=None, rect=None, spp_range=None, flags=0, wavefront_paths=0, film_filter=0,
                  filter_alpha=0.0):
    """RenderParams for a generated scene: film size, depth and RR from its config; film_filter is
    PG_FILTER_BOX, PG_FILTER_TRIANGLE or PG_FILTER_GAUSSIAN (with filter_alpha)."""
    w, h = sc["width"], sc["height"]
    spp = spp or sc["spp"]
    rect = rect or (0, 0, w, h)
    s0, s1 = spp_range or (0, spp)
    p = RenderParams()
    p.width, p.height, p.spp = w, h, spp
    p.rect[:] = list(rect)
    p.spp_begin, p.spp_end = s0, s1
    p.max_depth, p.min_rr, p.flags = sc["max_depth"], sc["min_rr"], flags
    p.ndc[:] = list(sc["ndc"])
    p.wavefront_paths = wavefront_paths
    p.filter, p.filter_alpha = film_filter, filter_alpha
    return p

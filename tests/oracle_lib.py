"""ctypes binding of the CPU oracle (oracle/liboracle.so). Test infrastructure only."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

_VP = ctypes.c_void_p
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError("oracle/liboracle.so missing: run __graft_entry__.build()")
        L = ctypes.CDLL(ORACLE_PATH)
        L.or_sobol_sample.restype = ctypes.c_float
        L.or_sobol_sample.argtypes = [ctypes.c_uint64, ctypes.c_uint32]
        L.or_interval_to_index.restype = ctypes.c_uint64
        L.or_interval_to_index.argtypes = [ctypes.c_uint32] * 4
        L.or_sobol_matrices.argtypes = [_VP]
        L.or_scene_create.restype = _VP
        L.or_scene_create.argtypes = [_VP, ctypes.c_uint32, _VP, ctypes.c_uint32, _VP, _VP, ctypes.c_uint32, _VP,
                                      ctypes.c_uint32, _VP, ctypes.c_uint32, _VP, _VP]
        L.or_scene_destroy.argtypes = [_VP]
        L.or_scene_set_wide_bvh.argtypes = [_VP, ctypes.c_uint32, _VP, ctypes.c_uint64, _VP, ctypes.c_uint64]
        L.or_scene_set_instances.argtypes = [_VP, _VP, ctypes.c_uint32]
        L.or_scene_set_motion.argtypes = [_VP, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_float,
                                          ctypes.c_float, _VP, _VP]
        L.or_trace_at.argtypes = [_VP, _VP, _VP, ctypes.c_uint32, _VP, _VP, _VP]
        L.or_scene_set_points.argtypes = [_VP, _VP, ctypes.c_uint32]
        L.or_scene_set_area_lights.argtypes = [_VP, _VP, ctypes.c_uint32]
        L.or_scene_set_area_lights.restype = ctypes.c_int
        L.or_scene_set_textures.argtypes = [_VP, _VP, _VP]
        L.or_scene_add_texture.restype = ctypes.c_uint32
        L.or_scene_add_texture.argtypes = [_VP, _VP, _VP, _VP]
        L.or_texture_sample.argtypes = [_VP, ctypes.c_uint32, ctypes.c_float, ctypes.c_float, _VP]
        L.or_trace.argtypes = [_VP, _VP, ctypes.c_uint32, _VP, _VP, _VP, ctypes.c_int, _VP]
        L.or_occluded.argtypes = [_VP, _VP, ctypes.c_uint32, _VP]
        L.or_camera_rays.argtypes = [_VP, _VP, _VP, _VP, ctypes.c_uint32, _VP, _VP]
        L.or_render.argtypes = [_VP, _VP, _VP, _VP, _VP, ctypes.c_int, _VP]
        L.or_bsdf_eval.argtypes = [_VP, ctypes.c_int, _VP, _VP, _VP, _VP]
        L.or_bsdf_sample.argtypes = [_VP, ctypes.c_int, _VP, ctypes.c_float, ctypes.c_float, _VP, _VP, _VP]
        L.or_bsdf_eval_y.argtypes = [_VP, ctypes.c_float, ctypes.c_int, _VP, _VP, _VP, _VP]
        L.or_bsdf_sample_y.argtypes = [_VP, ctypes.c_float, ctypes.c_int, _VP, ctypes.c_float, ctypes.c_float, _VP, _VP,
                                       _VP]
        L.or_scene_set_roughness_y.argtypes = [_VP, _VP, _VP]
        L.or_scene_set_roughness_y.restype = ctypes.c_int
        for name in ("or_ggx_d_aniso", "or_ggx_lambda"):
            getattr(L, name).argtypes = [ctypes.c_float] * 5
            getattr(L, name).restype = ctypes.c_float
        L.or_ggx_sample_wh.argtypes = [_VP, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, _VP]
        L.or_fresnel_dielectric.restype = ctypes.c_float
        L.or_fresnel_dielectric.argtypes = [ctypes.c_float] * 3
        L.or_remap_roughness.restype = ctypes.c_float
        L.or_remap_roughness.argtypes = [ctypes.c_float]
        L.or_ggx_d.restype = ctypes.c_float
        L.or_ggx_d.argtypes = [ctypes.c_float, ctypes.c_float]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_VP) if a is not None else None


def sobol_matrices():
    out = np.zeros(1024 * 52, np.uint32)
    lib().or_sobol_matrices(_p(out))
    return out.reshape(1024, 52)


def sobol_sample(index, dim):
    return lib().or_sobol_sample(int(index), int(dim))


def interval_to_index(m, s, px, py):
    return lib().or_interval_to_index(m, s, px, py)


class OracleScene:
    def __init__(self, sc):
        self.sc = sc
        self._arrs = [np.ascontiguousarray(sc[k]) for k in
                      ("vertices", "indices", "tri_object", "object_material", "materials", "quads", "camera", "env")]
        v, i, to, om, m, q, c, e = self._arrs
        self.h = lib().or_scene_create(_p(v), v.shape[0], _p(i), i.shape[0], _p(to), _p(om), om.shape[0], _p(m),
                                       m.shape[0], _p(q), q.shape[0], _p(c), _p(e))
        first = sc.get("instance_first")
        if first is not None and len(first):
            self._first = np.ascontiguousarray(first, np.uint32)
            lib().or_scene_set_instances(self.h, _p(self._first), len(self._first))
            for k, m in enumerate(sc.get("motions") or []):
                self.set_motion(k, **m)
        al = sc.get("area_lights")  # optional: primitives per area light (pg_scene_desc.area_light_quads)
        if al is not None:
            self._al = np.ascontiguousarray(al, np.uint32)
            assert lib().or_scene_set_area_lights(self.h, _p(self._al), len(self._al)) == 0
        ry, ryt = sc.get("roughness_y"), sc.get("roughness_y_textures")
        pts = sc.get("point_lights")
        if pts is not None and len(pts):
            self._pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 6)
            lib().or_scene_set_points(self.h, _p(self._pts), self._pts.shape[0])
        if sc.get("textures"):
            self._keep = []
            for t in sc["textures"]:
                it = np.array([t["type"], t["channels"], t.get("width", 0), t.get("height", 0), t.get("wrap", 0)],
                              np.uint32)
                ft = np.array([*t.get("black", [0] * 4), *t.get("white", [0] * 4), t.get("scale", 0.0)], np.float32)
                tx = np.ascontiguousarray(t["texels"], np.uint8) if t.get("texels") is not None else None
                self._keep += [it, ft, tx]
                lib().or_scene_add_texture(self.h, _p(it), _p(ft), _p(tx))
            triuv = None
            if sc.get("texcoords") is not None:
                uv = np.asarray(sc["texcoords"], np.float32).reshape(-1, 2)
                triuv = np.ascontiguousarray(uv[sc["indices"].astype(np.int64)].reshape(-1, 6))
            mt = np.ascontiguousarray(sc["material_textures"], np.uint32).reshape(-1, 3)
            self._keep += [triuv, mt]
            lib().or_scene_set_textures(self.h, _p(triuv), _p(mt))
        if ry is not None or ryt is not None:  # anisotropic roughness of glass and substrate
            self._ry = None if ry is None else np.ascontiguousarray(ry, np.float32)
            self._ryt = None if ryt is None else np.ascontiguousarray(ryt, np.uint32)
            assert lib().or_scene_set_roughness_y(self.h, _p(self._ry), _p(self._ryt)) == 0

    def set_motion(self, instance, keys, t0=0.0, t1=1.0, visible=True, parent=None):
        """Keys of moving instance `instance` (as Scene.set_motion)."""
        keys = np.ascontiguousarray(keys, np.float32).reshape(-1, 16)
        par = np.ascontiguousarray(parent if parent is not None else [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float32)
        lib().or_scene_set_motion(self.h, instance, keys.shape[0], int(bool(visible)), float(t0), float(t1), _p(keys),
                                  _p(par))

    def trace_at(self, rays, times):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        times = np.ascontiguousarray(times, np.float32).reshape(-1)
        n = rays.shape[0]
        t, prim, uv = np.zeros(n, np.float32), np.zeros(n, np.uint32), np.zeros((n, 2), np.float32)
        lib().or_trace_at(self.h, _p(rays), _p(times), n, _p(t), _p(prim), _p(uv))
        return t, prim, uv

    def use_wide_bvh(self, width, nodes, tris):
        """Traverse the GPU library's own wide BVH (Scene.bvh_download()) from now on, in the GPU's
        order, so trace counts and render counts are the GPU kernel's."""
        self._wide = (np.ascontiguousarray(nodes), np.ascontiguousarray(tris))
        n_nodes = self._wide[0].nbytes // (28 * width)
        lib().or_scene_set_wide_bvh(self.h, width, _p(self._wide[0]), n_nodes, _p(self._wide[1]),
                                    self._wide[1].nbytes // 48)

    def texture_sample(self, tex, u, v):
        out = np.zeros(4, np.float32)
        lib().or_texture_sample(self.h, int(tex), np.float32(u), np.float32(v), _p(out))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_scene_destroy(self.h)
            self.h = None

    def trace(self, rays, brute=False, counts=False):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        t, prim, uv = np.zeros(n, np.float32), np.zeros(n, np.uint32), np.zeros((n, 2), np.float32)
        cnt = np.zeros(2, np.uint64)
        lib().or_trace(self.h, _p(rays), n, _p(t), _p(prim), _p(uv), int(brute), _p(cnt) if counts else None)
        return (t, prim, uv, cnt) if counts else (t, prim, uv)

    def occluded(self, rays):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.zeros(rays.shape[0], np.uint8)
        lib().or_occluded(self.h, _p(rays), rays.shape[0], _p(out))
        return out

    def camera_rays(self, pix):
        pix = np.ascontiguousarray(pix, np.uint32).reshape(-1, 3)
        sc = self.sc
        ip = np.array([sc["width"], sc["height"], sc["spp"]], np.uint32)
        ndc = np.array(sc["ndc"], np.float32)
        rays = np.zeros((pix.shape[0], 8), np.float32)
        film = np.zeros((pix.shape[0], 2), np.float32)
        lib().or_camera_rays(self.h, _p(ip), _p(ndc), _p(pix), pix.shape[0], _p(rays), _p(film))
        return rays, film

    def render(self, rect=None, spp_range=None, flags=0, threads=None, path_l=False, film=True, spp=None,
               max_depth=None, min_rr=None, film_filter=0, filter_alpha=0.0):
        sc = self.sc
        w, h = sc["width"], sc["height"]
        spp = spp or sc["spp"]
        rect = rect or (0, 0, w, h)
        s0, s1 = spp_range or (0, spp)
        md = sc["max_depth"] if max_depth is None else max_depth
        rr = sc["min_rr"] if min_rr is None else min_rr
        ip = np.array([w, h, spp, md, rr, flags, *rect, s0, s1, film_filter], np.uint32)
        ndc = np.array([*sc["ndc"], filter_alpha], np.float32)
        f = np.zeros((h, w, 4), np.float64) if film else None
        pl = np.zeros((rect[2] * rect[3] * (s1 - s0), 3), np.float32) if path_l else None
        rc = np.zeros(5, np.uint64)
        lib().or_render(self.h, _p(ip), _p(ndc), _p(f), _p(pl), int(threads or os.cpu_count() or 1), _p(rc))
        return {"film": f, "path_l": pl, "rays": rc[:3], "nodes_visited": int(rc[3]), "tris_tested": int(rc[4])}

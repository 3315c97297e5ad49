"""Seeded scene generators for the benchmark and parity workloads.

Each generator returns a dict of numpy arrays in the layout of pg_scene_desc (include/piper_hip.h):
  vertices (V,3) f32, indices (T,3) u32, tri_object (T,) u32, object_material (O,) u32,
  materials (M,8) f32 = [type, r, g, b, alpha, specular rgb] (plastic: alpha holds the roughness), quads (Q,12) f32 = [origin, U, V, radiance],
  camera (14,) f32 = [position, look_at, up, sensor_w, sensor_h, focal, aperture, 0], env (3,) f32,
plus the render settings (width, height, spp, max_depth, min_rr, ndc).

Instances (the boxes and icospheres) are flattened to world space. The camera's up vector is made
orthogonal to the view direction, so the reference camera's unnormalised right/up vectors
(Sensor/PerspectiveCamera/Kernel.cpp:31-32) stay unit length.
"""
import math

import numpy as np

LAMBERT, CONDUCTOR, PLASTIC, MIRROR, GLASS, SUBSTRATE = 0, 1, 2, 3, 4, 5


def calc_render_rect(width, height, device_aspect, fit_fill=True):
    """calcRenderRECT (Interface/BuiltinComponent/Sensor.hpp:38-74): NDC affine transform and rect."""
    iar = np.float32(width) / np.float32(height)
    iiar = np.float32(1.0) / iar
    idar = np.float32(1.0) / np.float32(device_aspect)
    dar = np.float32(device_aspect)
    if fit_fill:
        rect = (0, 0, width, height)
        if iar > dar:
            ndc = (0.0, (idar - iiar) * 0.5 * dar, 1.0, iiar * dar)
        else:
            ndc = ((dar - iar) * 0.5 * idar, 0.0, iar * idar, 1.0)
        return tuple(float(np.float32(v)) for v in ndc), rect
    raise NotImplementedError("OverScan fit mode")


def make_camera(pos, look_at, up_hint, vfov_deg, aspect, focal=0.035, fstop=math.inf):
    pos, look_at = np.asarray(pos, np.float64), np.asarray(look_at, np.float64)
    fwd = look_at - pos
    focus = np.linalg.norm(fwd)
    fwd /= focus
    up = np.asarray(up_hint, np.float64)
    up = up - fwd * np.dot(up, fwd)
    up /= np.linalg.norm(up)
    film_dist = 1.0 / (1.0 / focal - 1.0 / focus)
    sh = 2.0 * film_dist * math.tan(math.radians(vfov_deg) / 2.0)
    sw = sh * aspect
    aperture = 0.0 if math.isinf(fstop) else focal / (2.0 * fstop)
    return np.array([*pos, *look_at, *up, sw, sh, focal, aperture, 0.0], np.float32)


def _quad_tris(p0, p1, p2, p3):
    return [p0, p1, p2, p3], [[0, 1, 2], [0, 2, 3]]


def _box(center_xz, y0, size, yaw):
    """Axis-aligned box of `size` (sx, sy, sz) rotated by yaw about +y, base at height y0: 12 tris."""
    sx, sy, sz = size
    c, s = math.cos(yaw), math.sin(yaw)
    local = np.array([[x, y, z] for y in (0.0, sy) for z in (-sz / 2, sz / 2) for x in (-sx / 2, sx / 2)])
    rot = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    v = local @ rot.T + np.array([center_xz[0], y0, center_xz[1]])
    # vertex index = y*4 + z*2 + x
    faces = [[0, 1, 3, 2], [4, 6, 7, 5], [0, 2, 6, 4], [1, 5, 7, 3], [0, 4, 5, 1], [2, 3, 7, 6]]
    tris = []
    for f in faces:
        tris += [[f[0], f[1], f[2]], [f[0], f[2], f[3]]]
    return v, np.array(tris)


def icosphere(subdiv):
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
         [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
         [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    v = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = v[a] + v[b]
                v.append(m / np.linalg.norm(m))
                cache[key] = len(v) - 1
            return cache[key]

        nf = []
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = nf
    return np.array(v), np.array(f)


def _value_noise(rng, x, z, cells_x, cells_z):
    lattice = rng.uniform(-1.0, 1.0, size=(cells_z + 2, cells_x + 2))
    gx, gz = x * cells_x, z * cells_z
    ix, iz = np.floor(gx).astype(int), np.floor(gz).astype(int)
    fx, fz = gx - ix, gz - iz
    sx, sz = fx * fx * (3 - 2 * fx), fz * fz * (3 - 2 * fz)
    a = lattice[iz, ix] * (1 - sx) + lattice[iz, ix + 1] * sx
    b = lattice[iz + 1, ix] * (1 - sx) + lattice[iz + 1, ix + 1] * sx
    return a * (1 - sz) + b * sz


def fbm_height(rng, u, v, octaves=4, amplitude=0.2, base=(8, 4)):
    """fBm value noise over the unit square (u, v in [0, 1]), peak amplitude `amplitude`."""
    h = np.zeros_like(u)
    norm = 0.0
    for o in range(octaves):
        w = 0.5 ** o
        h += w * _value_noise(rng, u, v, base[0] << o, base[1] << o)
        norm += w
    return amplitude * h / norm


class _Builder:
    def __init__(self):
        self.v, self.t, self.obj, self.objmat, self.mats = [], [], [], [], []
        self.nv = 0

    def material(self, kind, albedo, alpha=0.0, spec=(0.0, 0.0, 0.0)):
        self.mats.append([kind, *albedo, alpha, *spec])
        return len(self.mats) - 1

    def add(self, verts, tris, mat):
        oid = len(self.objmat)
        self.objmat.append(mat)
        self.v.append(np.asarray(verts, np.float64))
        self.t.append(np.asarray(tris, np.int64) + self.nv)
        self.obj.append(np.full(len(tris), oid, np.uint32))
        self.nv += len(verts)
        return oid

    def finish(self, quads, camera, **settings):
        sc = {
            "vertices": np.ascontiguousarray(np.concatenate(self.v).astype(np.float32)),
            "indices": np.ascontiguousarray(np.concatenate(self.t).astype(np.uint32)),
            "tri_object": np.ascontiguousarray(np.concatenate(self.obj)),
            "object_material": np.ascontiguousarray(np.array(self.objmat, np.uint32)),
            "materials": np.ascontiguousarray(np.array(self.mats, np.float32)),
            "quads": np.ascontiguousarray(np.array(quads, np.float32).reshape(-1, 12)),
            "camera": camera,
            "env": np.zeros(3, np.float32),
        }
        sc.update(settings)
        return sc


def cornell(seed=1, width=256, height=256, spp=16, max_depth=5, plastic=False, rough_walls=False,
            specular=False, point_lights=False, textured=False, glass_roughness=0.0, multi_light=False,
            anisotropic=False):
    """Config 1: closed box of 5 quads (open toward the camera), 2 seeded boxes, one ceiling light.

    plastic=True makes the two boxes plastic (seeded specular colour and roughness) and adds a
    tilted one-quad plastic panel that paths reach on both faces; rough_walls=True gives the walls
    seeded matte roughness (Oren-Nayar); specular=True makes the floor substrate (its albedo as the
    diffuse colour) and adds a smooth glass sphere (outward winding, so rays enter through front
    faces and leave through back faces) and a mirror panel before the back wall, tilted to show the
    ceiling (glass_roughness > 0 makes the sphere rough glass); point_lights=True adds two point lights (lights 2 and 3 after the environment and the
    ceiling quad); textured=True adds texture coordinates (a planar projection of every vertex),
    a checkerboard on the back wall, 8-bit images on the floor (repeat) and the left wall (mirror),
    a checkerboard roughness (0 and 0.6: Lambert and Oren-Nayar cells) on the ceiling and, with
    plastic=True, a checkerboard specular colour on the first box and an image roughness on the
    second; multi_light=True replaces the ceiling light by one Plane light of three primitives of
    unequal area (two halves of a 0.5 x 0.5 quad cut at 0.6 of its width and a 0.2 x 0.2 piece
    beside them) and adds a second light of two strips high on the back wall ("area_lights" [3, 2]);
    anisotropic=True (with specular=True) gives the glass sphere RoughnessY 0.25 and the substrate
    floor RoughnessY 0.8 ("roughness_y"; RoughnessX stays glass_roughness and 0.3, so the sphere is
    rough even at glass_roughness 0) and, with textured=True too, the sphere's RoughnessY from the
    roughness image ("roughness_y_textures"). The light and the rest of the geometry stay as in the
    plain scene."""
    rng = np.random.default_rng(seed)
    prng = np.random.default_rng(seed + 1000)
    b = _Builder()
    walls = {
        "floor": [[-1, -1, -1], [1, -1, -1], [1, -1, 1], [-1, -1, 1]],
        "ceiling": [[-1, 1, -1], [-1, 1, 1], [1, 1, 1], [1, 1, -1]],
        "back": [[-1, -1, -1], [-1, 1, -1], [1, 1, -1], [1, -1, -1]],
        "left": [[-1, -1, -1], [-1, -1, 1], [-1, 1, 1], [-1, 1, -1]],
        "right": [[1, -1, -1], [1, 1, -1], [1, 1, 1], [1, -1, 1]],
    }
    for name, corners in walls.items():
        alb = np.sort(rng.uniform(0.1, 0.8, 3))
        if name == "left":  # red-dominant
            alb = np.array([alb[2], alb[0], alb[1]])
        elif name == "right":  # green-dominant
            alb = np.array([alb[0], alb[2], alb[1]])
        else:
            rng.shuffle(alb)
        v, t = _quad_tris(*corners)
        if specular and name == "floor":
            b.add(v, t, b.material(SUBSTRATE, alb, 0.3, [0.25, 0.25, 0.25]))
        else:
            b.add(v, t, b.material(LAMBERT, alb, prng.uniform(0.2, 0.9) if rough_walls else 0.0))
    for _ in range(2):
        size = rng.uniform(0.3, 0.6, 3)
        yaw = rng.uniform(0.0, math.pi / 2)
        pos = rng.uniform(-0.6, 0.6, 2)
        v, t = _box(pos, -1.0, size, yaw)
        alb = rng.uniform(0.1, 0.8, 3)
        if plastic:
            mat = b.material(PLASTIC, alb, prng.uniform(0.05, 0.6), prng.uniform(0.2, 0.9, 3))
        else:
            mat = b.material(LAMBERT, alb)
        b.add(v, t, mat)
    if plastic:
        panel = [[-0.9, 0.1, -0.2], [-0.3, 0.1, -0.6], [-0.3, 0.7, -0.5], [-0.9, 0.7, -0.1]]
        v, t = _quad_tris(*panel)
        b.add(v, t, b.material(PLASTIC, [0.5, 0.3, 0.2], 0.2, [0.8, 0.8, 0.8]))
    if specular:
        sv, st = icosphere(2)
        b.add(sv * 0.28 + np.array([0.3, 0.05, 0.35]), st,
              b.material(GLASS, [1.0, 1.0, 1.0], glass_roughness, [0.95, 0.95, 0.95]))
        mirror = [[-0.7, -0.3, -0.55], [-0.1, -0.3, -0.55], [-0.1, 0.393, -0.95], [-0.7, 0.393, -0.95]]  # tilted up 30 deg
        v, t = _quad_tris(*mirror)
        b.add(v, t, b.material(MIRROR, [0.9, 0.9, 0.9]))
    quads = [[-0.25, 0.999, -0.25, 0.5, 0.0, 0.0, 0.0, 0.0, 0.5, 15.0, 15.0, 15.0]]
    if multi_light:
        quads = [[-0.25, 0.999, -0.25, 0.3, 0.0, 0.0, 0.0, 0.0, 0.5, 15.0, 15.0, 15.0],
                 [0.05, 0.999, -0.25, 0.2, 0.0, 0.0, 0.0, 0.0, 0.5, 15.0, 15.0, 15.0],
                 [-0.6, 0.999, 0.3, 0.2, 0.0, 0.0, 0.0, 0.0, 0.2, 15.0, 15.0, 15.0],
                 [-0.8, 0.6, -0.999, 0.3, 0.0, 0.0, 0.0, 0.2, 0.0, 4.0, 5.0, 6.0],
                 [0.5, 0.55, -0.999, 0.25, 0.0, 0.0, 0.0, 0.3, 0.0, 4.0, 5.0, 6.0]]
    aspect = width / height
    cam = make_camera([0, 0, 3.8], [0, 0, 0], [0, 1, 0], 40.0, aspect)
    ndc, _ = calc_render_rect(width, height, float(cam[9] / cam[10]))
    sc = b.finish(quads, cam, width=width, height=height, spp=spp, max_depth=max_depth, min_rr=max_depth,
                  ndc=ndc, name="cornell" + ("_plastic" if plastic else "") + ("_rough" if rough_walls else "")
                  + ("_specular" if specular else "") + ("_points" if point_lights else "")
                  + ("_multilight" if multi_light else ""))
    if multi_light:
        sc["area_lights"] = np.array([3, 2], np.uint32)
    if point_lights:
        sc["point_lights"] = np.array([[0.45, 0.6, 0.2, 1.5, 1.4, 1.2], [-0.55, -0.3, 0.45, 0.6, 0.7, 1.0]], np.float32)
    if textured:
        v = sc["vertices"].astype(np.float64)
        sc["texcoords"] = np.ascontiguousarray(np.stack([0.7 * v[:, 0] + 0.3 * v[:, 2] + 1.1,
                                                         0.8 * v[:, 1] - 0.2 * v[:, 2] + 0.9], 1).astype(np.float32))
        trng = np.random.default_rng(seed + 2000)
        none = 0xFFFFFFFF
        sc["textures"] = [
            {"type": 1, "channels": 4, "black": [0.1, 0.1, 0.15, 1.0], "white": [0.8, 0.75, 0.6, 1.0], "scale": 3.0},
            {"type": 2, "channels": 4, "width": 13, "height": 9, "wrap": 0,
             "texels": trng.integers(0, 256, (9, 13, 4), dtype=np.uint8)},
            {"type": 2, "channels": 4, "width": 7, "height": 5, "wrap": 1,
             "texels": trng.integers(0, 256, (5, 7, 4), dtype=np.uint8)},
            {"type": 1, "channels": 4, "black": [0.05, 0.05, 0.05, 1.0], "white": [0.9, 0.2, 0.2, 1.0], "scale": 5.0},
        ]
        sc["textures"].append({"type": 1, "channels": 1, "black": [0.0, 0, 0, 0], "white": [0.6, 0, 0, 0],
                               "scale": 4.0})  # roughness: Lambert / Oren-Nayar cells
        sc["textures"].append({"type": 2, "channels": 1, "width": 6, "height": 4, "wrap": 0,
                               "texels": trng.integers(20, 200, (4, 6, 1), dtype=np.uint8)})  # roughness image
        mt = np.full((len(sc["materials"]), 3), none, np.uint32)
        mt[sc["object_material"][2], 0] = 0  # back wall
        mt[sc["object_material"][0], 0] = 1  # floor
        mt[sc["object_material"][3], 0] = 2  # left wall
        mt[sc["object_material"][1], 2] = 4  # ceiling roughness: checkerboard of 0 and 0.6
        if plastic:
            mt[sc["object_material"][5], 1] = 3  # the first box's specular colour
            mt[sc["object_material"][6], 2] = 5  # the second box's roughness
        sc["material_textures"] = mt
        sc["name"] += "_textured"
    if anisotropic:
        assert specular
        ry = sc["materials"][:, 4].copy()
        glass = sc["object_material"][-2]  # the sphere (the mirror is the last object)
        floor = sc["object_material"][0]
        ry[glass], ry[floor] = 0.25, 0.8
        sc["roughness_y"] = ry
        if textured:
            ryt = np.full(len(ry), 0xFFFFFFFF, np.uint32)
            ryt[glass] = 5
            sc["roughness_y_textures"] = ryt
        sc["name"] += "_anisotropic"
    return sc


def terrain_spheres(seed=2, width=1920, height=1080, spp=64, max_depth=8, min_rr=3, grid=(1000, 500),
                    n_spheres=200, n_lights=16):
    """Config 2: fBm height-field terrain + 200 icospheres (subdivision 3), 16 emissive quads."""
    rng = np.random.default_rng(seed)
    b = _Builder()
    gx, gz = grid
    u, v = np.meshgrid(np.linspace(0.0, 1.0, gx), np.linspace(0.0, 1.0, gz))
    hgt = fbm_height(rng, u, v)
    verts = np.stack([-10.0 + 20.0 * u, hgt, -5.0 + 10.0 * v], axis=-1).reshape(-1, 3)
    i = np.arange(gz - 1)[:, None] * gx + np.arange(gx - 1)[None, :]
    i = i.reshape(-1)
    tris = np.concatenate([np.stack([i, i + gx, i + 1], -1), np.stack([i + 1, i + gx, i + gx + 1], -1)])
    b.add(verts, tris, b.material(LAMBERT, rng.uniform(0.1, 0.9, 3)))

    sv, st = icosphere(3)
    for _ in range(n_spheres):
        r = rng.uniform(0.1, 0.4)
        x, z = rng.uniform(-9.0, 9.0), rng.uniform(-4.5, 4.5)
        cx = int(round((x + 10.0) / 20.0 * (gx - 1)))
        cz = int(round((z + 5.0) / 10.0 * (gz - 1)))
        y = hgt[cz, cx] + 0.7 * r
        if rng.uniform() < 0.5:
            mat = b.material(LAMBERT, rng.uniform(0.1, 0.9, 3))
        else:
            rough = rng.uniform(0.05, 0.6)
            mat = b.material(CONDUCTOR, rng.uniform(0.5, 1.0, 3), rough * rough)  # alpha = roughness^2
        b.add(sv * r + np.array([x, y, z]), st, mat)

    quads = []
    for _ in range(n_lights):
        sx, sz = rng.uniform(0.5, 1.0, 2)
        x, z, y = rng.uniform(-8.0, 8.0), rng.uniform(-4.0, 4.0), rng.uniform(2.0, 3.0)
        rad = rng.uniform(5.0, 50.0, 3)
        quads.append([x - sx / 2, y, z - sz / 2, sx, 0.0, 0.0, 0.0, 0.0, sz, *rad])
    aspect = width / height
    cam = make_camera([0.0, 3.5, 11.0], [0.0, 0.0, 0.0], [0, 1, 0], 45.0, aspect)
    ndc, _ = calc_render_rect(width, height, float(cam[9] / cam[10]))
    return b.finish(quads, cam, width=width, height=height, spp=spp, max_depth=max_depth, min_rr=min_rr,
                    ndc=ndc, name="terrain_spheres")


def terrain_slit(seed=2, width=1920, height=1080, spp=64, max_depth=8, min_rr=3, grid=(1000, 500)):
    """Moderate indirect (guided config 4): the terrain + icosphere scene with its 16 lights moved
    behind a wall (z = -2.5, 6 high) whose only opening is a 0.3-high slot across its whole width at
    height 1.0-1.3, below the lights (2-3 high): the light falls through it onto a strip of terrain
    just in front of the wall, and the rest of the scene is lit mostly by that strip (one bounce)."""
    sc = terrain_spheres(seed=seed, width=width, height=height, spp=spp, max_depth=max_depth, min_rr=min_rr,
                         grid=grid)
    rng = np.random.default_rng(seed + 1000)
    b = _Builder()
    b.v, b.t, b.obj = [sc["vertices"].astype(np.float64)], [sc["indices"].astype(np.int64)], [sc["tri_object"]]
    b.objmat, b.mats, b.nv = list(sc["object_material"]), [list(m) for m in sc["materials"]], sc["vertices"].shape[0]
    wall = b.material(LAMBERT, [0.6, 0.6, 0.6])
    for y0, y1 in ((-1.0, 1.0), (1.3, 5.0)):
        v, t = _box((0.0, -2.5), y0, (21.0, y1 - y0, 0.2), 0.0)
        b.add(v, t, wall)
    quads = []
    for _ in range(16):
        sx, sz = rng.uniform(0.5, 1.0, 2)
        x, z, y = rng.uniform(-8.0, 8.0), rng.uniform(-4.8, -3.2), rng.uniform(2.0, 3.0)
        quads.append([x - sx / 2, y, z - sz / 2, sx, 0.0, 0.0, 0.0, 0.0, sz, *rng.uniform(5.0, 50.0, 3)])
    out = b.finish(quads, sc["camera"], width=width, height=height, spp=spp, max_depth=max_depth, min_rr=min_rr,
                   ndc=sc["ndc"], name="terrain_slit")
    return out


BATCH_RESOLUTIONS = ((256, 256), (512, 512), (1024, 768), (1920, 1080))
BATCH_SPP = (4, 16, 64)


def batch_scene(index, max_depth=8, min_rr=3):
    """Scene `index` of the batched workload (seed 3 + index): a displaced fBm grid plus instanced
    icospheres (about 30% of the triangles), log-uniform in [1e4, 2e6] triangles in all, with the
    terrain scene's materials and 16 lights; resolution drawn from BATCH_RESOLUTIONS, samples
    per pixel from BATCH_SPP, max depth 8."""
    rng = np.random.default_rng(3 + index)
    total = math.exp(rng.uniform(math.log(1e4), math.log(2e6)))
    width, height = BATCH_RESOLUTIONS[int(rng.integers(len(BATCH_RESOLUTIONS)))]
    spp = BATCH_SPP[int(rng.integers(len(BATCH_SPP)))]
    n_spheres = max(1, int(round(0.3 * total / 1280)))
    grid_tris = max(total - 1280 * n_spheres, 64.0)
    cz = max(2, int(round(math.sqrt(grid_tris / 4))))  # 2 cx cz cells' triangles, cx = 2 cz
    cx = max(2, int(grid_tris // (2 * cz)))
    sc = terrain_spheres(seed=int(rng.integers(1 << 31)), width=width, height=height, spp=spp, max_depth=max_depth,
                         min_rr=min_rr, grid=(cx + 1, cz + 1), n_spheres=n_spheres)
    sc["name"] = f"batch_{index}"
    return sc


SMALL_RESOLUTIONS = [(64, 64), (128, 96), (128, 128), (256, 256)]
SMALL_SPP = [1, 4, 16]


def small_batch_scene(index):
    """Scene `index` of the small-scene batch (seed 100 + index): a Cornell box (config 1's layout,
    its boxes and wall colours drawn from the seed) at a resolution from SMALL_RESOLUTIONS and a
    sample count from SMALL_SPP, depth 5; every fourth has plastic boxes, every fourth (offset two)
    two several-primitive lights. Each holds a few thousand to a million paths: far fewer than a
    path pool's capacity, the regime a shared pool is for."""
    rng = np.random.default_rng(100 + index)
    width, height = SMALL_RESOLUTIONS[int(rng.integers(len(SMALL_RESOLUTIONS)))]
    spp = SMALL_SPP[int(rng.integers(len(SMALL_SPP)))]
    sc = cornell(seed=100 + index, width=width, height=height, spp=spp, plastic=index % 4 == 0,
                 multi_light=index % 4 == 2)
    sc["name"] = f"small_{index}"
    return sc


def _grid_quad(origin, u, v, n):
    """A parallelogram tessellated into n x n cells (2 n^2 triangles)."""
    origin, u, v = (np.asarray(a, np.float64) for a in (origin, u, v))
    s = np.linspace(0.0, 1.0, n + 1)
    S, T = np.meshgrid(s, s, indexing="ij")
    verts = origin + S.reshape(-1, 1) * u + T.reshape(-1, 1) * v
    i = (np.arange(n)[:, None] * (n + 1) + np.arange(n)[None, :]).reshape(-1)
    tris = np.concatenate([np.stack([i, i + n + 1, i + 1], -1), np.stack([i + 1, i + n + 1, i + n + 2], -1)])
    return verts, tris


def hard_room(seed=0, width=512, height=512, spp=64, max_depth=12, tess=204):
    """Hard indirect (guided config 5): a closed 2 x 1.6 x 2 room (walls tessellated to ~500k
    triangles, Lambert U[0.3, 0.7]) with one 0.05 x 0.05 light of radiance 500 inside a lamp
    housing. No straight line joins the light and the housing's opening: light leaves only after a
    GGX conductor reflector (roughness 0.1, alpha 0.01) or the housing's inner walls. The opening
    faces the back wall 0.35-0.65 away, so the reflected beam lights a patch there and the rest of
    the room sees the light only through that patch. The camera stands at the front wall, behind
    the closed back of the housing. The room has the Cornell box's scale, so the 7.9 W light
    leaves pixel values of 0.01-0.5 rather than values lost under relMSE's 0.01 floor."""
    rng = np.random.default_rng(100 + seed)
    b = _Builder()
    W, H, D = 1.0, 1.6, 1.0
    walls = [
        ([-W, 0, -D], [2 * W, 0, 0], [0, 0, 2 * D]),    # floor
        ([-W, H, -D], [0, 0, 2 * D], [2 * W, 0, 0]),    # ceiling
        ([-W, 0, -D], [0, H, 0], [2 * W, 0, 0]),        # back
        ([-W, 0, D], [2 * W, 0, 0], [0, H, 0]),         # front (behind the camera)
        ([-W, 0, -D], [0, 0, 2 * D], [0, H, 0]),        # left
        ([W, 0, -D], [0, H, 0], [0, 0, 2 * D]),         # right
    ]
    for o, u, v in walls:
        verts, tris = _grid_quad(o, u, v, tess)
        b.add(verts, tris, b.material(LAMBERT, rng.uniform(0.3, 0.7, 3)))
    # Housing: 0.3 x 0.2 x 0.3 box at a seeded spot; local depth q runs from its closed back
    # (q = -0.15, toward the camera) to its front (q = +0.15, toward the back wall), z = cz - q.
    # The light lies on the floor of the rear half; a shelf over the front half and the closed
    # lower front hide it from every straight line through the opening (the upper front), so it
    # reaches the room only after the 45-degree reflector in the rear top corner or the walls.
    cx, cz = rng.uniform(-0.5, 0.5), rng.uniform(-0.5, -0.2)
    y0 = rng.uniform(0.7, 1.1)
    x0, x1 = cx - 0.15, cx + 0.15

    def at(x, y, q):
        return [x, y, cz - q]

    def along(dx, dy, dq):
        return [dx, dy, -dq]

    housing = b.material(LAMBERT, [0.5, 0.5, 0.5])
    panels = [
        (at(x0, y0, -0.15), along(0.3, 0, 0), along(0, 0, 0.3)),            # bottom
        (at(x0, y0 + 0.2, -0.15), along(0, 0, 0.3), along(0.3, 0, 0)),      # top
        (at(x0, y0, -0.15), along(0, 0.2, 0), along(0.3, 0, 0)),            # back
        (at(x0, y0, -0.15), along(0, 0, 0.3), along(0, 0.2, 0)),            # left
        (at(x1, y0, -0.15), along(0, 0.2, 0), along(0, 0, 0.3)),            # right
        (at(x0, y0, 0.15), along(0.3, 0, 0), along(0, 0.1, 0)),             # lower half of the front
        (at(x0, y0 + 0.1, 0.0), along(0.3, 0, 0), along(0, 0, 0.15)),       # shelf over the front half
    ]
    for o, u, v in panels:
        verts, tris = _grid_quad(o, u, v, 1)
        b.add(verts, tris, housing)
    verts, tris = _grid_quad(at(x0, y0 + 0.075, -0.15), along(0.3, 0, 0), along(0, 0.125, 0.125), 1)
    b.add(verts, tris, b.material(CONDUCTOR, [0.95, 0.95, 0.95], 0.1 * 0.1))
    lo = at(cx - 0.025, y0 + 0.01, -0.125)
    quads = [lo + [0.05, 0.0, 0.0] + along(0, 0, 0.05) + [500.0, 500.0, 500.0]]
    cam = make_camera([0.0, 0.85, 0.95], [0.0, 0.75, -1.0], [0, 1, 0], 60.0, width / height)
    ndc, _ = calc_render_rect(width, height, float(cam[9] / cam[10]))
    return b.finish(quads, cam, width=width, height=height, spp=spp, max_depth=max_depth, min_rr=3, ndc=ndc,
                    name=f"hard_room_{seed}")


HARD_ROOM_RESOLUTIONS = ((512, 512), (1024, 768), (1280, 720), (1920, 1080))


def hard_room_variant(index, scale=1, spp=64, tess=204):
    """Variant `index` of the guided batch workload (16 seeded hard rooms): hard_room(seed=index) at a
    resolution drawn from HARD_ROOM_RESOLUTIONS (512x512 to 1920x1080) by seed 500 + index, divided
    by `scale` in both directions; max depth 12."""
    rng = np.random.default_rng(500 + index)
    width, height = HARD_ROOM_RESOLUTIONS[int(rng.integers(len(HARD_ROOM_RESOLUTIONS)))]
    return hard_room(seed=index, width=width // scale, height=height // scale, spp=spp, max_depth=12, tess=tess)


def hidden_cornell(width=64, height=64, spp=64, max_depth=6):
    """Small hard-indirect test scene: the Cornell box with a plate under the ceiling light, so the
    room is lit only after the light bounces off the ceiling."""
    sc = cornell(width=width, height=height, spp=spp, max_depth=max_depth)
    b = _Builder()
    b.v, b.t, b.obj = [sc["vertices"].astype(np.float64)], [sc["indices"].astype(np.int64)], [sc["tri_object"]]
    b.objmat, b.mats, b.nv = list(sc["object_material"]), [list(m) for m in sc["materials"]], sc["vertices"].shape[0]
    verts, tris = _grid_quad([-0.45, 0.7, -0.45], [0.9, 0, 0], [0, 0, 0.9], 1)
    b.add(verts, tris, b.material(LAMBERT, [0.5, 0.5, 0.5]))
    return b.finish(sc["quads"], sc["camera"], width=width, height=height, spp=spp, max_depth=max_depth,
                    min_rr=max_depth, ndc=sc["ndc"], name="hidden_cornell")


def sphere_furnace(rho=0.7, intensity=(2.0, 2.0, 2.0), radius=1.0, subdiv=4, width=32, height=32, spp=64,
                   max_depth=3):
    """Closed diffuse sphere (an icosphere mesh, albedo rho) with one point light at its centre and
    a black environment; the camera sits inside, off centre. Every surface point receives the
    direct irradiance E0 = I / R^2, and inside a sphere any uniform radiosity B gives every point
    the irradiance B again (the form factor to an area element is dA / (4 pi R^2) from anywhere on
    the sphere), so each bounce multiplies by rho: with max_depth D (D vertices take a light sample,
    PathTracing/Kernel.cpp:148) every camera ray carries (E0 / pi) sum_{k=1..D} rho^k. Russian
    roulette is off (min_rr = max_depth)."""
    v, f = icosphere(subdiv)
    b = _Builder()
    b.add(v * radius, f, b.material(LAMBERT, [rho] * 3))
    cam = make_camera([0.15 * radius, -0.1 * radius, 0.3 * radius], [-0.2 * radius, 0.1 * radius, -radius],
                      [0, 1, 0], 70.0, width / height)
    ndc, _ = calc_render_rect(width, height, float(cam[9] / cam[10]))
    sc = b.finish(np.zeros((0, 12), np.float32), cam, width=width, height=height, spp=spp, max_depth=max_depth,
                  min_rr=max_depth, ndc=ndc, name="sphere_furnace")
    sc["point_lights"] = np.array([[0.0, 0.0, 0.0, *intensity]], np.float32)
    return sc


def sphere_furnace_radiance(rho, intensity, radius, max_depth):
    """Closed-form radiance of every camera ray in sphere_furnace (per channel)."""
    e0 = np.asarray(intensity, np.float64) / (radius * radius)
    return e0 / math.pi * sum(rho ** k for k in range(1, max_depth + 1))


def triangle_targets(seed=0, n_tris=16, rays_per_tri=64):
    """Isolated random triangles (one every 12 units along x) and rays aimed at points of known
    barycentrics: returns (scene, rays (N, 8), expected t (N,), expected (u, v) (N, 2), triangle
    index (N,)). The target point is (1 - u - v) v0 + u v1 + v v2 (Embree's convention: u weighs
    v1 and v weighs v2, Tracer/Embree/Kernel.cpp:176-178); the ray starts a distance t0 before it
    along a random direction scaled by s in {0.5, 1, 2}, so the expected hit distance in ray
    parameter units is t0 / s (the expected values are solved again in FP64 for the rays as rounded
    to float32). Directions of every sign on every dominant axis occur."""
    rng = np.random.default_rng(seed)
    b = _Builder()
    tris = []
    for i in range(n_tris):
        c = np.array([12.0 * i, 0.0, 0.0])
        tv = c + rng.uniform(-1.5, 1.5, (3, 3))
        while np.linalg.norm(np.cross(tv[1] - tv[0], tv[2] - tv[0])) < 0.5:
            tv = c + rng.uniform(-1.5, 1.5, (3, 3))
        b.add(tv, [[0, 1, 2]], b.material(LAMBERT, [0.5, 0.5, 0.5]))
        tris.append(tv.astype(np.float32).astype(np.float64))
    rays, ts, uvs, idx = [], [], [], []
    for i, tv in enumerate(tris):
        nrm = np.cross(tv[1] - tv[0], tv[2] - tv[0])
        nrm /= np.linalg.norm(nrm)
        for _ in range(rays_per_tri):
            while True:
                u, v = rng.uniform(0.05, 0.9, 2)
                if u + v < 0.95:
                    break
            p = (1 - u - v) * tv[0] + u * tv[1] + v * tv[2]
            while True:
                d = rng.normal(size=3)
                d /= np.linalg.norm(d)
                if abs(np.dot(d, nrm)) > 0.2:
                    break
            t0 = rng.uniform(0.5, 5.0)
            s = [0.5, 1.0, 2.0][int(rng.integers(3))]
            o = p - d * t0
            rays.append([*o, *(d * s), 1e-3, 1e5])
            ts.append(t0 / s)
            uvs.append([u, v])
            idx.append(i)
    rays = np.array(rays, np.float32)
    # the float32 rays hit slightly off the designed points: solve o + t d = (1-u-v) v0 + u v1 + v v2
    # exactly (FP64) for the stored rays
    for k, i in enumerate(idx):
        tv = tris[i]
        o, d = rays[k, :3].astype(np.float64), rays[k, 3:6].astype(np.float64)
        a = np.stack([d, tv[0] - tv[1], tv[0] - tv[2]], 1)
        ts[k], uvs[k][0], uvs[k][1] = np.linalg.solve(a, tv[0] - o)
    cam = make_camera([0, 0, 20], [0, 0, 0], [0, 1, 0], 40.0, 1.0)
    sc = b.finish(np.zeros((0, 12), np.float32), cam, width=8, height=8, spp=1, max_depth=1, min_rr=1,
                  ndc=(0.0, 0.0, 1.0, 1.0), name="triangle_targets")
    return sc, rays, np.array(ts), np.array(uvs), np.array(idx, np.uint32)


def srt(scale=(1.0, 1.0, 1.0), quat=(1.0, 0.0, 0.0, 0.0), trans=(0.0, 0.0, 0.0), skew=(0.0, 0.0, 0.0),
        shift=(0.0, 0.0, 0.0)):
    """A TransformSRT key as 16 floats (Kernel/Transform.hpp:390-393): scale xyz, skew xy xz yz,
    shift xyz, quaternion w x y z, translation xyz."""
    return (*map(float, scale), *map(float, skew), *map(float, shift), *map(float, quat), *map(float, trans))


def motion_occluder(kind="translate", width=48, height=32, spp=16, t_range=(0.0, 1.0), half=0.6, travel=3.0,
                    depth=5.0):
    """A moving black occluder in front of a white environment, seen by a pinhole camera at the
    origin looking down -z: every path's radiance is 1 or 0, depending only on whether its camera
    ray meets the occluder at the path's time (closed form: occluder_blocked). The occluder is the
    scene's one moving instance (no static geometry), two keys at times t_range:
      "translate": a square of half-width `half` at z = -depth whose centre moves along x from
                   -travel/2 to +travel/2;
      "rotate":    a 2 x 0.3 plank at z = -depth turning a quarter turn about the view axis."""
    b = _Builder()
    blk = b.material(LAMBERT, [0.0, 0.0, 0.0])
    if kind == "translate":
        verts, tris = _quad_tris([-half, -half, 0], [half, -half, 0], [half, half, 0], [-half, half, 0])
        keys = [srt(trans=(-travel / 2, 0.0, -depth)), srt(trans=(travel / 2, 0.0, -depth))]
    elif kind == "rotate":
        verts, tris = _quad_tris([-1.0, -0.15, 0], [1.0, -0.15, 0], [1.0, 0.15, 0], [-1.0, 0.15, 0])
        h = math.sqrt(0.5)
        keys = [srt(trans=(0.0, 0.0, -depth)), srt(quat=(h, 0.0, 0.0, h), trans=(0.0, 0.0, -depth))]
    else:
        raise ValueError(kind)
    b.add(verts, tris, blk)
    cam = make_camera([0.0, 0.0, 0.0], [0.0, 0.0, -depth], [0.0, 1.0, 0.0], 40.0, width / height)
    ndc, _ = calc_render_rect(width, height, float(cam[9] / cam[10]))
    sc = b.finish(np.zeros((0, 12), np.float32), cam, width=width, height=height, spp=spp, max_depth=2, min_rr=2,
                  ndc=ndc, name="motion_occluder_" + kind)
    sc["env"] = np.ones(3, np.float32)
    sc["instance_first"] = np.array([0], np.uint32)
    sc["motions"] = [dict(keys=keys, t0=t_range[0], t1=t_range[1])]
    sc["occluder"] = dict(kind=kind, half=half, travel=travel, depth=depth, t_range=t_range)
    return sc


def occluder_blocked(occ, rays, times):
    """Closed form of motion_occluder, in float64: for camera rays (N, 8) and times
    (N,), whether the ray meets the occluder at its time, and the distance of its plane point from
    the occluder's outline (to leave out rays within rounding of an edge)."""
    rays = np.asarray(rays, np.float64)
    t = np.asarray(times, np.float64)
    o, d = rays[:, 0:3], rays[:, 3:6]
    s = (-occ["depth"] - o[:, 2]) / d[:, 2]
    X, Y = o[:, 0] + d[:, 0] * s, o[:, 1] + d[:, 1] * s
    t0, t1 = occ["t_range"]
    present = (t >= t0) & (t <= t1)
    f = np.clip((t - t0) / (t1 - t0), 0.0, 1.0)
    if occ["kind"] == "translate":
        cx = -occ["travel"] / 2 + occ["travel"] * f
        a, b = np.abs(X - cx), np.abs(Y)
        hx = hy = occ["half"]
    else:
        th = 0.5 * math.pi * f  # slerp: the rotation angle grows linearly in the fraction
        a = np.abs(X * np.cos(th) + Y * np.sin(th))
        b = np.abs(-X * np.sin(th) + Y * np.cos(th))
        hx, hy = 1.0, 0.15
    inside = (a <= hx) & (b <= hy)
    margin = np.minimum(np.abs(a - hx), np.abs(b - hy))
    margin = np.where(np.abs(t - t0) < 1e-6, 0.0, margin)
    margin = np.where(np.abs(t - t1) < 1e-6, 0.0, margin)
    return present & inside, margin


def moving_cornell(width=64, height=64, spp=8, max_depth=5):
    """The Cornell box (static) with a moving instance: a tall box that slides, turns about +y and
    grows while the shutter is open (three keys), lit by the ceiling light: shadows, normals and
    MIS rays all see the box at the path's time."""
    sc = cornell(width=width, height=height, spp=spp, max_depth=max_depth)
    b = _Builder()
    b.v, b.t, b.obj = [sc["vertices"].astype(np.float64)], [sc["indices"].astype(np.int64)], [sc["tri_object"]]
    b.objmat, b.mats, b.nv = list(sc["object_material"]), [list(m) for m in sc["materials"]], sc["vertices"].shape[0]
    first = int(sc["indices"].shape[0])
    verts, tris = _box((0.0, 0.0), 0.0, (0.3, 0.5, 0.3), 0.0)
    b.add(verts, tris, b.material(LAMBERT, [0.7, 0.6, 0.3]))
    out = b.finish(sc["quads"], sc["camera"], width=width, height=height, spp=spp, max_depth=max_depth,
                   min_rr=max_depth, ndc=sc["ndc"], name="moving_cornell")
    c, s_ = math.cos(math.radians(30)), math.sin(math.radians(30))
    c2, s2 = math.cos(math.radians(60)), math.sin(math.radians(60))
    y0 = float(np.min(sc["vertices"][:, 1])) + 0.01  # just above the floor: no coplanar faces
    keys = [srt(trans=(-0.25, y0, 0.1)),
            srt(quat=(c, 0.0, s_, 0.0), trans=(0.0, y0, 0.0)),
            srt(scale=(1.0, 1.3, 1.0), quat=(c2, 0.0, s2, 0.0), trans=(0.25, y0, -0.1))]
    out["instance_first"] = np.array([first], np.uint32)
    out["motions"] = [dict(keys=keys, t0=0.0, t1=1.0)]
    return out


def world_triangles(sc):
    """(T, 3, 3) float32 world-space triangle vertices."""
    return sc["vertices"][sc["indices"]]

"""Scene files in the reference's JSON schema (PiperRenderer.cpp) for the front-end tests, written
from a generated scene dict, plus readers for what piper_render writes (.pgsc exports, EXR frames)."""
import json
import os
import struct
import zlib

import numpy as np

from plugin_io import IDENTITY_SRT  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FRONTEND = os.path.join(ROOT, "piper2_amd", "piper_render")

SURF = "Piper.BuiltinComponent.ClassicSurface."
CONST = "Piper.BuiltinComponent.ProceduralTexture.ConstantTexture"


def _tex(*v):
    return {"ClassID": CONST, "Value": [float(x) for x in v]}


def _rgba(c):
    return _tex(c[0], c[1], c[2], 1.0)


def surface_config(m):
    """One material row [type, r, g, b, alpha, spec rgb] as a reference surface config."""
    t, c, a, s = int(m[0]), m[1:4], float(m[4]), m[5:8]
    if t == 0:
        return {"ClassID": SURF + "Matte", "Diffuse": _rgba(c), "Roughness": _tex(a)}
    if t == 2:
        return {"ClassID": SURF + "Plastic", "Diffuse": _rgba(c), "Specular": _rgba(s), "Roughness": _tex(a)}
    if t == 3:
        return {"ClassID": SURF + "Mirror", "Reflection": _rgba(c)}
    if t == 4:
        return {"ClassID": SURF + "Glass", "Reflection": _rgba(c), "Transmission": _rgba(s), "RoughnessX": _tex(a),
                "RoughnessY": _tex(a)}
    if t == 5:
        return {"ClassID": SURF + "Substrate", "Diffuse": _rgba(c), "Specular": _rgba(s), "RoughnessX": _tex(a),
                "RoughnessY": _tex(a)}
    raise ValueError(f"material type {t} has no reference class")


def write_obj(path, verts, tris, negative=False, uv=None):
    with open(path, "w") as f:
        f.write("# written by the front-end tests\n")
        for v in verts:
            f.write("v %.9g %.9g %.9g\n" % tuple(float(x) for x in v))
        for t in (uv if uv is not None else []):
            f.write("vt %.9g %.9g\n" % tuple(float(x) for x in t))
        n = len(verts)
        for t in tris:
            if uv is not None:  # the texcoord list parallels the vertex list
                f.write("f %d/%d %d/%d %d/%d\n" % tuple(int(i) + 1 for i in t for _ in range(2)))
            elif negative:
                f.write("f %d//1 %d %d//2\n" % tuple(int(i) - n for i in t))  # normals named, not read
            else:
                f.write("f %d %d %d\n" % tuple(int(i) + 1 for i in t))


def write_pam(path, texels):
    h, w, c = texels.shape
    kind = {1: "GRAYSCALE", 2: "GRAYSCALE_ALPHA", 4: "RGB_ALPHA"}[c]
    with open(path, "wb") as f:
        f.write(f"P7\nWIDTH {w}\nHEIGHT {h}\nDEPTH {c}\nMAXVAL 255\nTUPLTYPE {kind}\nENDHDR\n".encode())
        f.write(np.ascontiguousarray(texels, np.uint8).tobytes())


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = np.abs(p - a), np.abs(p - b), np.abs(p - c)
    return np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, b, c))


def write_png(path, texels, level=6, rgb=False):
    """8-bit PNG (PNG spec 1.2): row y uses scanline filter y % 5 (None, Sub, Up, Average, Paeth)
    so every filter is decoded; level 0 makes stored deflate blocks, higher levels Huffman-coded ones.
    rgb=True drops a 4-channel image's alpha (colour type 2)."""
    img = np.ascontiguousarray(texels[..., :3] if rgb else texels, np.uint8).astype(np.int32)
    h, w, c = img.shape
    raw = bytearray()
    for y in range(h):
        cur = img[y].reshape(-1)
        up = img[y - 1].reshape(-1) if y else np.zeros_like(cur)
        left = np.concatenate([np.zeros(c, np.int32), cur[:-c]])
        upleft = np.concatenate([np.zeros(c, np.int32), up[:-c]])
        f = y % 5
        pred = [np.zeros_like(cur), left, up, (left + up) // 2, _paeth(left, up, upleft)][f]
        raw.append(f)
        raw += ((cur - pred) & 0xFF).astype(np.uint8).tobytes()

    def chunk(kind, body):
        return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xFFFFFFFF)

    colour = {1: 0, 2: 4, 3: 2, 4: 6}[c]
    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, colour, 0, 0, 0)))
        data = zlib.compress(bytes(raw), level)
        half = len(data) // 2  # two IDAT chunks: the stream continues across them
        f.write(chunk(b"IDAT", data[:half]) + chunk(b"IDAT", data[half:]) + chunk(b"IEND", b""))


def texture_config(t, folder, name, image_format="pam"):
    """A texture as a reference texture config; image_format "pam" or "png" (level 9 for colour,
    0 for single-channel images, so both stored and Huffman-coded deflate blocks are read)."""
    if t["type"] == 1:
        ch = t["channels"]
        return {"ClassID": "Piper.BuiltinComponent.ProceduralTexture.CheckBoard", "Black": list(t["black"][:ch]),
                "White": list(t["white"][:ch]), "Scale": t["scale"]}
    if image_format == "png":
        name = name.rsplit(".", 1)[0] + ".png"
        texels = t["texels"]
        rgb = texels.shape[2] == 4 and bool((texels[..., 3] == 255).all())
        write_png(os.path.join(folder, name), texels, level=9 if texels.shape[2] > 1 else 0, rgb=rgb)
    else:
        write_pam(os.path.join(folder, name), t["texels"])
    return {"WrapMode": "Repeat" if t["wrap"] == 0 else "Mirror",
            "Image": {"ClassID": "Piper.BuiltinComponent.SampledTexture.Image", "Path": name}}


def write_scene_json(sc, folder, spp=None, filter_cfg=None, fit="Fill", transforms=None, image_format="pam"):
    """Write sc as scene.json + one OBJ per object + options.json under folder; returns the options
    path. transforms: optional {object index: Transform config} (the OBJ then holds the object's
    vertices as given, the transform places them)."""
    os.makedirs(folder, exist_ok=True)
    children = []
    for o, mat in enumerate(sc["object_material"]):
        tri = sc["indices"][sc["tri_object"] == o]
        vid, inv = np.unique(tri, return_inverse=True)
        name = f"object{o}.obj"
        uv = sc["texcoords"][vid] if sc.get("texcoords") is not None else None
        write_obj(os.path.join(folder, name), sc["vertices"][vid], inv.reshape(-1, 3), negative=(o % 2 == 1), uv=uv)
        surf = surface_config(sc["materials"][mat])
        if sc.get("material_textures") is not None:
            for slot, key in enumerate(("albedo", "spec", "rough")):
                t = int(sc["material_textures"][mat][slot])
                if t != 0xFFFFFFFF:
                    keys = {"albedo": ("Diffuse", "Reflection"), "spec": ("Specular", "Transmission"),
                            "rough": ("Roughness", "RoughnessX")}[key]
                    for k in keys + (("RoughnessY",) if key == "rough" else ()):
                        if k in surf:
                            surf[k] = texture_config(sc["textures"][t], folder, f"texture{t}.pam", image_format)
                            if k != "RoughnessX":
                                break
        if "RoughnessY" in surf:  # anisotropic glass / substrate: RoughnessY of its own
            ry, ryt = sc.get("roughness_y"), sc.get("roughness_y_textures")
            if ryt is not None and int(ryt[mat]) != 0xFFFFFFFF:
                t = int(ryt[mat])
                surf["RoughnessY"] = texture_config(sc["textures"][t], folder, f"texture{t}.pam", image_format)
            elif ry is not None:
                surf["RoughnessY"] = _tex(float(ry[mat]))
        inst = {"Type": "Instance", "Geometry": {"ClassID": "Piper.BuiltinComponent.TriangleMesh.TriangleMesh",
                                                 "Path": name},
                "Surface": surf}
        if transforms and o in transforms:
            inst = {"Type": "Group", "Children": [dict(inst, Transform=transforms[o])]}
        children.append(inst)
    counts = sc.get("area_lights")  # primitives per area light (one each by default)
    counts = [1] * len(sc["quads"]) if counts is None else [int(c) for c in counts]
    first = 0
    for c in counts:
        qs = sc["quads"][first:first + c]
        first += c
        plane = {"ClassID": "Piper.BuiltinComponent.BasicShapes.Plane",
                 "Primitives": [{"Origin": q[0:3].tolist(), "U": q[3:6].tolist(), "V": q[6:9].tolist()} for q in qs]}
        children.append({"Type": "Instance", "ClassID": "Piper.BuiltinComponent.AreaLight.Diffuse", "Geometry": plane,
                         "Radiance": qs[0][9:12].tolist()})
    for p in sc.get("point_lights", np.zeros((0, 6))):
        children.append({"Type": "Group", "Children": [{
            "Type": "Instance", "ClassID": "Piper.BuiltinComponent.BasicLight.Point", "Intensity": p[3:6].tolist(),
            "Transform": {"Translation": p[0:3].tolist()}}]})
    children.append({"Type": "Instance", "ClassID": "Piper.BuiltinComponent.BasicLight.SampledEnvironment",
                     "Radiance": sc["env"].tolist()})
    c = sc["camera"].astype(np.float64)
    fstop = 1e30 if c[12] == 0 else float(c[11] / (2 * c[12]))
    sensor = {"Type": "Instance", "ID": "Camera", "ClassID": "Piper.BuiltinComponent.PerspectiveCamera.Sensor",
              "LookAt": c[3:6].tolist(), "Up": c[6:9].tolist(), "SensorSize": [c[9] * 1000, c[10] * 1000],
              "FocalLength": c[11] * 1000, "FStop": fstop, "Transform": {"Translation": c[0:3].tolist()}}
    children.append({"Type": "Group", "Children": [sensor]})
    scene = {
        "RenderDriver": {"ClassID": "Piper.BuiltinComponent.FixedSampler.Driver",
                         "Filter": filter_cfg or {"ClassID": "Piper.BuiltinComponent.FixedSampler.SimpleFilter",
                                                  "FilterType": "box"}},
        "Actions": [{"ActiveSensor": "Camera", "TimeBegin": 0.0, "ShutterOpenTime": 0.0, "ShutterCloseTime": 0.0,
                     "FrameCount": 1, "FitMode": fit}],
        "Objects": [],
        "Scene": {"Type": "Group", "Children": children},
        "LightSampler": {"ClassID": "Piper.BuiltinComponent.BasicLight.UniformLightSampler"},
        "Integrator": {"ClassID": "Piper.BuiltinComponent.PathTracing.Integrator",
                       "MaxTraceDepth": int(sc["max_depth"]), "MinRussianRouletteDepth": int(sc["min_rr"])},
        "Sampler": {"ClassID": "Piper.BuiltinComponent.ScrambledSobolSampler.Sampler",
                    "SamplesPerPixel": int(spp or sc["spp"]), "Scramble": False},
    }
    with open(os.path.join(folder, "scene.json"), "w") as f:
        json.dump(scene, f, indent=1)
    opts = {"ImageSize": [int(sc["width"]), int(sc["height"])], "SceneDesc": "scene.json",
            "SceneParser": "Piper.Infrastructure.JsonSerializer.JsonSerializer",
            "Tracer": {"ClassID": "Piper.BuiltinComponent.Embree.Tracer"}, "OutputDir": "out"}
    path = os.path.join(folder, "options.json")
    with open(path, "w") as f:
        json.dump(opts, f, indent=1)
    return path


def make_plane_object(options_path, child, origin, u, v):
    """Replace the geometry of scene child `child` (an Instance with a Surface) by a one-primitive
    BasicShapes Plane: the surface is then shaded at the plane's own (u, v) coordinates."""
    path = os.path.join(os.path.dirname(options_path), "scene.json")
    scene = json.load(open(path))
    node = scene["Scene"]["Children"][child]
    assert node["Type"] == "Instance" and "Surface" in node
    node["Geometry"] = {"ClassID": "Piper.BuiltinComponent.BasicShapes.Plane",
                        "Primitives": [{"Origin": list(map(float, origin)), "U": list(map(float, u)),
                                        "V": list(map(float, v))}]}
    json.dump(scene, open(path, "w"), indent=1)


def read_pgsc(path):
    """The scene-file layout of tests/plugin/render_scene.cpp, as a scene dict."""
    with open(path, "rb") as f:
        data = f.read()
    version = struct.unpack("<I", data[4:8])[0]
    assert data[:4] == b"PGSC" and version & 3 in (1, 2) and version < 16
    multi_primitive = bool(version & 4)  # area lights of several primitives: their counts follow the instances
    anisotropic = bool(version & 8)  # RoughnessY per material and its texture end the file
    version &= 3
    head = np.frombuffer(data, np.uint32, 12, 8)
    nv, nt, no, nm, nq, w, h, spp, md, rr, compat, fit = (int(x) for x in head)
    off = 56
    out = {}

    def take(name, dt, n, shape):
        nonlocal off
        a = np.frombuffer(data, dt, n, off).reshape(shape)
        off += a.nbytes
        out[name] = a.copy()

    take("vertices", np.float32, 3 * nv, (nv, 3))
    take("indices", np.uint32, 3 * nt, (nt, 3))
    take("tri_object", np.uint32, nt, (nt,))
    take("object_material", np.uint32, no, (no,))
    take("materials", np.float32, 8 * nm, (nm, 8))
    take("quads", np.float32, 12 * nq, (nq, 12))
    take("camera", np.float32, 14, (14,))
    take("env", np.float32, 3, (3,))
    take("srt", np.float32, 16 * no, (no, 16))
    assert np.all(out["srt"] == IDENTITY_SRT[None])
    has_textures = False
    if off < len(data):
        npt = struct.unpack("<I", data[off:off + 4])[0]
        off += 4
        take("point_lights", np.float32, 6 * npt, (npt, 6))
        has_textures = off < len(data)
    if version == 2 or multi_primitive or anisotropic:  # an explicit texture flag; the rest after the textures
        has_textures = struct.unpack("<I", data[off:off + 4])[0] == 1
        off += 4
    if has_textures:
        take("texcoords", np.float32, 2 * nv, (nv, 2))
        ntex = struct.unpack("<I", data[off:off + 4])[0]
        off += 4
        texs = []
        for _ in range(ntex):
            typ, ch = struct.unpack("<II", data[off:off + 8])
            fl = np.frombuffer(data, np.float32, 9, off + 8)
            wd, ht, wrap = struct.unpack("<III", data[off + 44:off + 56])
            off += 56
            t = {"type": typ, "channels": ch, "black": fl[0:4].tolist(), "white": fl[4:8].tolist(),
                 "scale": float(fl[8]), "width": wd, "height": ht, "wrap": wrap}
            if typ == 2:
                n = wd * ht * ch
                t["texels"] = np.frombuffer(data, np.uint8, n, off).reshape(ht, wd, ch).copy()
                off += n
            texs.append(t)
        out["textures"] = texs
        take("material_textures", np.uint32, 3 * nm, (nm, 3))
    if version == 2:
        ni = struct.unpack("<I", data[off:off + 4])[0]
        off += 4
        take("instance_first", np.uint32, ni, (ni,))
        motions = []
        for _ in range(ni):
            nk, vis = struct.unpack("<II", data[off:off + 8])
            t0, t1 = struct.unpack("<ff", data[off + 8:off + 16])
            off += 16
            keys = np.frombuffer(data, np.float32, 16 * nk, off).reshape(nk, 16).copy()
            off += keys.nbytes
            parent = np.frombuffer(data, np.float32, 12, off).copy()
            off += parent.nbytes
            motions.append(dict(keys=keys, t0=t0, t1=t1, visible=bool(vis), parent=parent))
        out["motions"] = motions
    if multi_primitive:
        nl = struct.unpack("<I", data[off:off + 4])[0]
        off += 4
        take("area_lights", np.uint32, nl, (nl,))
    if anisotropic:
        take("roughness_y", np.float32, nm, (nm,))
        take("roughness_y_textures", np.uint32, nm, (nm,))
    assert off == len(data)
    out.update(width=w, height=h, spp=spp, max_depth=md, min_rr=rr, fit=fit)
    return out


def read_exr(path):
    """Minimal reader for the uncompressed HALF scanline EXR piper_render writes, following the
    OpenEXR file layout; returns (h, w, 3) float16 RGB and the header attributes."""
    with open(path, "rb") as f:
        d = f.read()
    assert struct.unpack("<ii", d[:8]) == (20000630, 2)
    off, attrs = 8, {}
    while d[off] != 0:
        e = d.index(b"\0", off)
        name = d[off:e].decode()
        e2 = d.index(b"\0", e + 1)
        typ = d[e + 1:e2].decode()
        size = struct.unpack("<i", d[e2 + 1:e2 + 5])[0]
        attrs[name] = (typ, d[e2 + 5:e2 + 5 + size])
        off = e2 + 5 + size
    off += 1
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    chans, cl = [], attrs["channels"][1]
    p = 0
    while cl[p] != 0:
        e = cl.index(b"\0", p)
        chans.append((cl[p:e].decode(), struct.unpack("<i", cl[e + 1:e + 5])[0]))
        p = e + 17
    assert attrs["compression"][1] == b"\0" and all(t == 1 for _, t in chans)
    table = np.frombuffer(d, np.uint64, h, off)
    img = np.zeros((h, w, len(chans)), np.float16)
    for y in range(h):
        o = int(table[y])
        ly, size = struct.unpack("<ii", d[o:o + 8])
        planes = np.frombuffer(d, np.float16, w * len(chans), o + 8).reshape(len(chans), w)
        assert size == planes.nbytes
        img[ly - y0] = planes.T
    names = [c for c, _ in chans]
    rgb = np.stack([img[..., names.index(c)] for c in "RGB"], -1)
    return rgb, attrs

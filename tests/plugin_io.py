"""Scene files for the C++ component-pair driver (tests/plugin/render_scene.cpp) and how to run it."""
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "plugin", "render_scene")

IDENTITY_SRT = np.array([1, 1, 1, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0], np.float32)


def write_scene(sc, path, srt=None, compat=0, fit=0):
    """srt: (objects, 16) float32 TransformSRT per object (identity by default)."""
    no = sc["object_material"].shape[0]
    srt = np.tile(IDENTITY_SRT, (no, 1)) if srt is None else np.asarray(srt, np.float32).reshape(no, 16)
    head = np.array([sc["vertices"].shape[0], sc["indices"].shape[0], no, sc["materials"].shape[0],
                     sc["quads"].shape[0], sc["width"], sc["height"], sc["spp"], sc["max_depth"], sc["min_rr"],
                     compat, fit], np.uint32)
    with open(path, "wb") as f:
        f.write(b"PGSC")
        f.write(np.uint32(1).tobytes())
        for a, dt in ((head, np.uint32), (sc["vertices"], np.float32), (sc["indices"], np.uint32),
                      (sc["tri_object"], np.uint32), (sc["object_material"], np.uint32),
                      (sc["materials"], np.float32), (sc["quads"], np.float32), (sc["camera"][:14], np.float32),
                      (sc["env"], np.float32), (srt, np.float32)):
            f.write(np.ascontiguousarray(a, dt).tobytes())
        pts = sc.get("point_lights")
        if pts is not None and len(pts):  # optional tail: count + {position, intensity} per light
            pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 6)
            f.write(np.uint32(pts.shape[0]).tobytes())
            f.write(pts.tobytes())


def run_driver(*args, timeout=120):
    return subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=timeout)

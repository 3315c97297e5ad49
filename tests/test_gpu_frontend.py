"""The scene front-end on the device: a reference-schema JSON scene rendered by piper_render
(HIPTracer pipeline, FixedSampler frames, EXR output) equals pg_render of the scene the front-end
exports, pixel for pixel after the same half conversion."""
import json
import os
import subprocess

import numpy as np
import pytest

import piper2_amd as pg
from oracle_lib import OracleScene
from piper2_amd import scenes
from scene_json import FRONTEND, make_plane_object, read_exr, read_pgsc, write_scene_json

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = pg.Device(0)
    yield d
    d.close()


def direct_rgb(dev, e, film_filter=0, alpha=0.0):
    sc = dict(e)
    cam = sc["camera"]
    assert sc["fit"] == 0
    sc["ndc"], _ = scenes.calc_render_rect(sc["width"], sc["height"], float(cam[9] / cam[10]))
    s = pg.Scene(dev, sc)
    f = pg.Film(dev, sc["width"], sc["height"])
    s.render(pg.render_params(sc, film_filter=film_filter, filter_alpha=alpha), f)
    film = f.download()
    w = film[..., 3:]
    with np.errstate(divide="ignore", invalid="ignore"):
        rgb = np.where(w != 0, film[..., :3] / w, np.float32(0)).astype(np.float32)
    return rgb.astype(np.float16)


@pytest.mark.parametrize("filt,textured,plane,multi", [
    (None, False, False, False), (None, True, False, False), (None, True, True, False), (None, False, False, True),
    (None, True, False, True),
    ({"ClassID": "Piper.BuiltinComponent.FixedSampler.SimpleFilter", "FilterType": "gaussian", "Alpha": 2.0}, False,
     False, False)])
def test_json_scene_renders_like_pg_render(dev, tmp_path, filt, textured, plane, multi):
    """plane=True makes the checkerboard back wall a BasicShapes Plane, shaded at its own (u, v);
    multi=True has two area lights of several Plane primitives each and anisotropic roughness on the
    glass sphere and the substrate floor (the sphere's RoughnessY from an image when textured)."""
    sc = scenes.cornell(width=64, height=48, spp=8, plastic=True, rough_walls=True, specular=True, point_lights=True,
                        textured=textured, multi_light=multi, glass_roughness=0.2 if multi else 0.0, anisotropic=multi)
    opts = write_scene_json(sc, str(tmp_path), filter_cfg=filt)
    if plane:
        make_plane_object(opts, 2, [-1.0, -1.0, -1.0], [0.0, 2.0, 0.0], [2.0, 0.0, 0.0])
    scene_path = tmp_path / "scene.json"
    scene = json.load(open(scene_path))
    scene["Actions"][0]["FrameCount"] = 2
    json.dump(scene, open(scene_path, "w"))
    r = subprocess.run([FRONTEND, opts, "--export", tmp_path / "exp"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    e = read_pgsc(tmp_path / "exp" / "Action0.pgsc")
    print("piper_render", opts, flush=True)
    r = subprocess.run([FRONTEND, opts], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    frames = [read_exr(tmp_path / "out" / "Action0" / f"Frame{k}.exr")[0] for k in range(2)]
    want = direct_rgb(dev, e, *((pg.PG_FILTER_GAUSSIAN, 2.0) if filt else ()))
    assert np.array_equal(frames[0].view(np.uint16), want.view(np.uint16))
    assert np.array_equal(frames[1].view(np.uint16), frames[0].view(np.uint16))  # a static scene
    assert "paths" in r.stdout
    if plane:  # the plane's checkerboard through the whole chain: GPU paths equal the oracle's
        ex = dict(e)
        ex["ndc"], _ = scenes.calc_render_rect(ex["width"], ex["height"], float(ex["camera"][9] / ex["camera"][10]))
        s = pg.Scene(dev, ex)
        _, pl = s.render(pg.render_params(ex), None, path_l=True)
        ref = OracleScene(ex).render(path_l=True, film=False, threads=16)["path_l"]
        close = np.all(np.abs(pl - ref) <= 1e-3 * np.abs(ref) + 1e-5, axis=1)
        assert close.mean() >= 0.99, close.mean()


def test_keyframed_scene_frames_render_like_pg_render(dev, tmp_path):
    """A box on a keyframed transform: piper_render's frames (the launcher sets each frame's keys,
    updateTransformSRT's selection) equal pg_render of the exported scene with that frame's keys,
    whose paths match the oracle's."""
    sc = scenes.cornell(width=48, height=32, spp=8)
    kf = [{"Key": 0, "Transform": {"Translation": [0.0, 0.0, 0.0]}},
          {"Key": 1, "Transform": {"Translation": [0.3, 0.0, -0.1], "Rotation": [1.0, 0.0, 0.3, 0.0]}},
          {"Key": 3, "Transform": {"Translation": [0.5, 0.1, 0.0], "Scale": [1.0, 1.3, 1.0]}}]
    opts = write_scene_json(sc, str(tmp_path), transforms={5: {"Offset": 0.0, "Step": 0.5, "KeyFrames": kf}})
    scene_path = tmp_path / "scene.json"
    scene = json.load(open(scene_path))
    a0 = dict(scene["Actions"][0], TimeBegin=0.6, ShutterOpenTime=0.2, ShutterCloseTime=0.1, FrameCount=2)
    scene["Actions"] = [a0, dict(a0, TimeBegin=0.9, FrameCount=1)]  # action 1 = action 0's second frame
    json.dump(scene, open(scene_path, "w"))
    r = subprocess.run([FRONTEND, opts, "--export", tmp_path / "exp"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    exports = [read_pgsc(tmp_path / "exp" / f"Action{i}.pgsc") for i in range(2)]
    print("piper_render", opts, flush=True)
    r = subprocess.run([FRONTEND, opts], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    frames = [read_exr(tmp_path / "out" / "Action0" / f"Frame{k}.exr")[0] for k in range(2)]
    for frame, e in zip(frames, exports):
        assert np.array_equal(frame.view(np.uint16), direct_rgb(dev, e).view(np.uint16))
    assert not np.array_equal(frames[0].view(np.uint16), frames[1].view(np.uint16))
    ex = dict(exports[0])
    ex["ndc"], _ = scenes.calc_render_rect(ex["width"], ex["height"], float(ex["camera"][9] / ex["camera"][10]))
    s = pg.Scene(dev, ex)
    _, pl = s.render(pg.render_params(ex), None, path_l=True)
    ref = OracleScene(ex).render(path_l=True, film=False, threads=16)["path_l"]
    close = np.all(np.abs(pl - ref) <= 1e-3 * np.abs(ref) + 1e-5, axis=1)
    assert close.mean() >= 0.98, close.mean()

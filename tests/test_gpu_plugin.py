"""The C++ component pair on the device: a frame driven the way Piper2 drives a Tracer plugin
(scene graph -> HIPTracer::buildPipeline -> Pipeline::prepare -> FixedSampler::renderFrame with
8x8 tile launches) must produce the film pg_render produces for the same flattened scene."""
import re

import numpy as np
import pytest

import piper2_amd as pg
from piper2_amd import scenes
from plugin_io import IDENTITY_SRT, run_driver, write_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = pg.Device(0)
    yield d
    d.close()


def direct_film(dev, sc, flags=0, film_filter=0, alpha=0.0):
    s = pg.Scene(dev, sc)
    f = pg.Film(dev, sc["width"], sc["height"])
    s.render(pg.render_params(sc, flags=flags, film_filter=film_filter, filter_alpha=alpha), f)
    return f.download()


def plugin_film(tmp_path, sc, name, filter_args=(), via_module=False, **kw):
    path, out = tmp_path / f"{name}.pgsc", tmp_path / f"{name}.rgbw"
    write_scene(sc, path, **kw)
    r = run_driver(*(["--via-module"] if via_module else []), str(path), str(out), *filter_args)
    assert r.returncode == 0, r.stdout + r.stderr
    film = np.fromfile(out, np.float32).reshape(sc["height"], sc["width"], 4)
    return film, r.stdout


def test_fixed_sampler_frame_matches_pg_render(dev, tmp_path):
    sc = scenes.cornell(width=100, height=68, spp=8)
    film, log = plugin_film(tmp_path, sc, "cornell")
    assert np.array_equal(film, direct_film(dev, sc))
    # the 8x8 tiles of the frame were coalesced into one render call over the whole rect
    assert "1 render calls" in log and f"{100 * 68 * 8} paths" in log, log
    assert re.search(r"rect: 0 0 100 68", log)


def test_module_loaded_pair_matches_pg_render(dev, tmp_path):
    """The pair as a Piper host loads it: dlopen the plugin library, check piperGetProtocol, build
    the module with the host's allocator and ask it for the "Tracer" class (with its "Accelerator"
    config); every component then reaches the device through its reference-layout payload
    (MatteData, DiffuseAreaLightData + CDFData, PCData, SobolData, Data, RDData + GaussianFilterData)
    and the pipeline's render kernel (compileKernel / launchKernel). The film equals pg_render's."""
    sc = scenes.cornell(width=80, height=64, spp=8, plastic=True, specular=True, point_lights=True)
    film, log = plugin_film(tmp_path, sc, "module", filter_args=("gaussian", "1.5"), via_module=True)
    assert "protocol GCC@STD@0" in log, log
    assert np.array_equal(film, direct_film(dev, sc, film_filter=pg.PG_FILTER_GAUSSIAN, alpha=1.5))


@pytest.mark.parametrize("name,code,alpha", [("triangle", pg.PG_FILTER_TRIANGLE, 0.0),
                                             ("gaussian", pg.PG_FILTER_GAUSSIAN, 2.5)])
def test_fixed_sampler_filters(dev, tmp_path, name, code, alpha):
    # the FixedSampler's SimpleFilter (FilterType, Alpha) becomes the film filter of its launches
    sc = scenes.cornell(width=72, height=56, spp=8)
    args = (name,) if code == pg.PG_FILTER_TRIANGLE else (name, str(alpha))
    film, _ = plugin_film(tmp_path, sc, name, filter_args=args)
    assert np.array_equal(film, direct_film(dev, sc, film_filter=code, alpha=alpha))
    assert not np.array_equal(film, direct_film(dev, sc))


def test_plastic_and_rough_matte_surfaces(dev, tmp_path):
    # Surface::plastic (Diffuse, Specular, Roughness) reaches the library as PG_MAT_PLASTIC, and a
    # matte roughness as the Oren-Nayar lobe
    sc = scenes.cornell(width=64, height=48, spp=8, plastic=True, rough_walls=True)
    film, _ = plugin_film(tmp_path, sc, "plastic")
    assert np.array_equal(film, direct_film(dev, sc))


def test_mirror_glass_substrate_surfaces_and_point_lights(dev, tmp_path):
    sc = scenes.cornell(width=64, height=48, spp=8, specular=True, point_lights=True)
    film, _ = plugin_film(tmp_path, sc, "specular")
    assert np.array_equal(film, direct_film(dev, sc))


def test_conductors_compat_flags_and_environment(dev, tmp_path):
    sc = scenes.terrain_spheres(width=64, height=40, spp=4, grid=(80, 40), n_spheres=10)
    sc["env"] = np.array([0.2, 0.3, 0.4], np.float32)
    flags = pg.PG_COMPAT_D1 | pg.PG_COMPAT_D9
    film, _ = plugin_film(tmp_path, sc, "terrain", compat=flags)
    assert np.array_equal(film, direct_film(dev, sc, flags))


def test_instance_transforms(dev, tmp_path):
    """Objects given in local space with a TransformSRT per instance. A translation is applied as
    x + t in FP32, so the film matches pre-translated vertices bit for bit; a rotation and scale
    match a numpy-transformed copy to rounding."""
    sc = scenes.cornell(width=64, height=64, spp=8)
    box_objects = [5, 6]  # the two boxes after the five walls
    t = np.array([0.25, 0.0, -0.125], np.float32)
    srt = np.tile(IDENTITY_SRT, (sc["object_material"].shape[0], 1))
    moved = dict(sc)
    v = sc["vertices"].copy()
    for o in box_objects:
        srt[o, 13:16] = t
        vid = np.unique(sc["indices"][sc["tri_object"] == o])
        v[vid] = v[vid] + t
    moved["vertices"] = np.ascontiguousarray(v)
    film, _ = plugin_film(tmp_path, sc, "moved", srt=srt)
    assert np.array_equal(film, direct_film(dev, moved))
    # rotation by 30 degrees about y and uniform scale 0.9 about the box's own origin
    ang = np.deg2rad(30.0)
    q = np.array([np.cos(ang / 2), 0.0, np.sin(ang / 2), 0.0])
    srt2 = np.tile(IDENTITY_SRT, (sc["object_material"].shape[0], 1))
    rot = dict(sc)
    v2 = sc["vertices"].astype(np.float64).copy()
    c, s = np.cos(ang), np.sin(ang)
    R = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    for o in box_objects:
        srt2[o, 0:3] = 0.9
        srt2[o, 9:13] = q
        vid = np.unique(sc["indices"][sc["tri_object"] == o])
        v2[vid] = (R @ (0.9 * v2[vid]).T).T
    rot["vertices"] = np.ascontiguousarray(v2.astype(np.float32))
    film2, _ = plugin_film(tmp_path, sc, "rotated", srt=srt2)
    ref = direct_film(dev, rot)
    img = lambda f: f[..., :3] / np.maximum(f[..., 3:], 1e-30)  # noqa: E731
    assert np.mean((img(film2) - img(ref)) ** 2 / (img(ref) ** 2 + 0.01)) < 1e-4


def test_error_paths_raise(tmp_path):
    sc = scenes.cornell(width=16, height=16, spp=1)
    path = tmp_path / "small.pgsc"
    write_scene(sc, path)
    r = run_driver("--errors", str(path))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "raised 14 of 14" in r.stdout, r.stdout
    for case in ("foreign format: raised: buildProgram: unsupported linkable format LLVM IR",
                 "kernel of a foreign format: raised: Unsupported linkable format LLVM IR",
                 "undefined kernel entry: raised: compileKernel: undefined symbol piperMain",
                 "texture channels: raised: Mismatched channel. Expect 1, but get 4",
                 "negative glass RoughnessY: raised: glass roughness must be finite and not negative"):
        assert case in r.stdout, case

"""The batched, moderate-indirect and hard-indirect workloads at their stated sizes on the GPU.

- Moderate indirect (`scenes.terrain_slit`: 1.25M triangles, 16 lights behind a slotted wall,
  1920x1080): per-path plain radiance against the oracle on a crop of the full frame, and the
  guided estimator (frozen trained guide, one-sample MIS 0.5) has the plain estimator's mean there.
- Hard indirect (`scenes.hard_room`: 499,408 triangles, depth 12): per-path plain radiance against
  the oracle on a crop; the 16 seeded variants at their own resolutions (512x512 to 1920x1080) in
  one guided batch give films bit-identical to single guided renders.
- Batched (`scenes.batch_scene`: 64 scenes, 26.5M triangles, 256x256 to 1920x1080): one
  pg_render_batch of all 64 full frames (one sample each, at different sample offsets) leaves
  every film bit-identical to the scene's own pg_render; two of the scenes match the oracle path
  for path on crops.
"""
import numpy as np
import pytest

import piper2_amd as pg
from oracle_lib import OracleScene
from piper2_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = pg.Device(0)
    yield d
    d.close()


def _paths_match(sc, dev, rect, spp, frac):
    """Per-path radiance of the GPU against the oracle on `rect` x the first `spp` samples."""
    s = pg.Scene(dev, sc)
    p = pg.render_params(sc, rect=rect, spp_range=(0, spp))
    _, pl_g = s.render(p, None, path_l=True)
    pl_c = OracleScene(sc).render(rect=rect, spp_range=(0, spp), path_l=True, film=False, threads=16)["path_l"]
    close = np.all(np.abs(pl_g - pl_c) <= 1e-3 * np.abs(pl_c) + 1e-5, axis=1)
    assert close.mean() >= frac, close.mean()
    # these scenes are lit indirectly, so most paths carry nothing: the paths that carry light on
    # either side must match on their own
    lit = (pl_c.max(1) > 0) | (pl_g.max(1) > 0)
    assert lit.sum() >= 500, lit.sum()
    assert close[lit].mean() >= frac - 0.03, close[lit].mean()
    s.close()


@pytest.fixture(scope="module")
def slit():
    return scenes.terrain_slit()


def test_moderate_indirect_crop_paths_match_oracle(dev, slit):
    _paths_match(slit, dev, rect=(880, 380, 96, 64), spp=4, frac=0.98)


def test_moderate_indirect_guided_mean_equals_plain(dev, slit):
    """A guide trained on the first 16 samples of a 256x128 crop, then frozen, renders the next 48
    samples; the crop's guided mean equals the plain mean within 4 standard errors of the
    per-pixel differences."""
    rect = (832, 500, 256, 128)
    s = pg.Scene(dev, slit)
    trainer = pg.Guide(dev, seed=3, train_records=1 << 16, train_batch=1 << 13)
    for k in range(4):
        s.render_guided(trainer, pg.render_params(slit, rect=rect, spp_range=(4 * k, 4 * k + 4)))
    frozen = pg.Guide(dev, seed=3, train_records=0)
    frozen.set_params(trainer.get_params())
    films = []
    for guided in (False, True):
        f = pg.Film(dev, slit["width"], slit["height"])
        p = pg.render_params(slit, rect=rect, spp_range=(16, 64), flags=pg.PG_FILM_FP64)
        if guided:
            s.render_guided(frozen, p, f)
        else:
            s.render(p, f)
        img = f.download(fp64=True)[rect[1]:rect[1] + rect[3], rect[0]:rect[0] + rect[2]]
        films.append((img[..., :3] / img[..., 3:]).mean(-1))
        f.close()
    diff = (films[1] - films[0]).reshape(-1)
    z = diff.mean() / (diff.std() / np.sqrt(diff.size))
    assert abs(z) < 4.0, (z, films[0].mean(), films[1].mean())
    assert films[0].mean() > 0
    for o in (trainer, frozen, s):
        o.close()


def test_hard_indirect_crop_paths_match_oracle(dev):
    sc = scenes.hard_room(seed=0)
    assert sc["indices"].shape[0] == 499408
    _paths_match(sc, dev, rect=(256, 128, 64, 64), spp=16, frac=0.98)


def test_hard_indirect_sixteen_variants_guided_batch(dev):
    """The 16 hard rooms at their own resolutions, one frozen guide each, one sample per pixel (the
    first eight rooms at sample 0, the others at sample 1): the guided batch equals the single
    guided renders bit for bit."""
    rooms = [scenes.hard_room_variant(i) for i in range(16)]
    assert {(sc["width"], sc["height"]) for sc in rooms} <= set(scenes.HARD_ROOM_RESOLUTIONS)
    objs = [pg.Scene(dev, sc) for sc in rooms]
    guides = [pg.Guide(dev, seed=40 + i, train_records=0) for i in range(16)]
    params = [pg.render_params(sc, spp_range=(i // 8, i // 8 + 1)) for i, sc in enumerate(rooms)]
    singles = []
    for s, sc, g, p in zip(objs, rooms, guides, params):
        f = pg.Film(dev, sc["width"], sc["height"])
        s.render_guided(g, p, f)
        singles.append(f.download())
        f.close()
    films = [pg.Film(dev, sc["width"], sc["height"]) for sc in rooms]
    stats, lat = pg.render_batch(dev, objs, params, films, guides=guides)
    assert stats["paths"] == sum(sc["width"] * sc["height"] for sc in rooms)
    for k, f in enumerate(films):
        got = f.download()
        assert np.array_equal(got, singles[k]), f"room {k}: guided batch film differs from its own render"
        assert np.any(got[..., :3] > 0)
    assert np.all(lat > 0)
    for o in films + objs + guides:
        o.close()


def test_batched_sixty_four_scenes_match_single_renders(dev):
    scs = [scenes.batch_scene(i) for i in range(64)]
    tris = sum(sc["indices"].shape[0] for sc in scs)
    assert 20e6 < tris < 40e6, tris
    objs = [pg.Scene(dev, sc) for sc in scs]
    params = [pg.render_params(sc, spp_range=(i % sc["spp"], i % sc["spp"] + 1)) for i, sc in enumerate(scs)]
    singles = []
    for s, sc, p in zip(objs, scs, params):
        f = pg.Film(dev, sc["width"], sc["height"])
        s.render(p, f)
        singles.append(f.download())
        f.close()
    films = [pg.Film(dev, sc["width"], sc["height"]) for sc in scs]
    stats, lat = pg.render_batch(dev, objs, params, films)
    assert stats["paths"] == sum(sc["width"] * sc["height"] for sc in scs)
    for k, f in enumerate(films):
        assert np.array_equal(f.download(), singles[k]), f"scene {k}: batched film differs from its own render"
    assert np.all(lat > 0)
    for o in films + objs:
        o.close()


@pytest.mark.parametrize("index", [5, 38])
def test_batched_scene_crop_paths_match_oracle(dev, index):
    sc = scenes.batch_scene(index)
    w, h = sc["width"], sc["height"]
    _paths_match(sc, dev, rect=(w // 2 - 32, h // 2 - 32, 64, 64), spp=4, frac=0.98)

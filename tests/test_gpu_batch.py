"""Batched rendering (pg_render_batch): several scenes' paths share one wavefront pipeline, and
every film must be bit-identical to the film its own pg_render call leaves."""
import numpy as np
import pytest

import piper2_amd as pg
from piper2_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = pg.Device(0)
    yield d
    d.close()


def lit(sc):
    sc = dict(sc)
    sc["env"] = np.array([0.3, 0.4, 0.5], np.float32)  # non-black: exercises the closest-hit MIS queue
    return sc


def mixed_batch():
    """Scenes of different sizes, depths, rects, sample ranges and flags."""
    items = []
    sc = scenes.cornell(width=96, height=64, spp=16)
    items.append((sc, pg.render_params(sc)))
    sc = scenes.hidden_cornell(width=40, height=56, spp=8)
    items.append((sc, pg.render_params(sc, rect=(5, 7, 23, 31), spp_range=(3, 7))))
    sc = scenes.terrain_spheres(width=120, height=72, spp=4, grid=(120, 60), n_spheres=12)
    items.append((sc, pg.render_params(sc, flags=pg.PG_COMPAT_D1 | pg.PG_COMPAT_D9)))
    sc = lit(scenes.cornell(seed=5, width=64, height=64, spp=8, max_depth=7))
    items.append((sc, pg.render_params(sc, flags=pg.PG_FILM_FP64)))
    sc = scenes.batch_scene(0)  # 15.7k triangles, 256x256 at 4 spp
    items.append((sc, pg.render_params(sc)))
    sc = scenes.cornell(seed=9, width=33, height=17, spp=2, max_depth=3)
    items.append((sc, pg.render_params(sc, flags=pg.PG_COMPAT_D3 | pg.PG_COMPAT_D13)))
    # textures, every surface kind, point lights: the batch takes the textured shade variant, which
    # must leave the untextured scenes' films unchanged too
    sc = scenes.cornell(seed=3, width=48, height=40, spp=8, textured=True, plastic=True, specular=True,
                        point_lights=True, glass_roughness=0.2, rough_walls=True)
    items.append((sc, pg.render_params(sc)))
    return items


def download(film, p):
    return film.download(fp64=bool(p.flags & pg.PG_FILM_FP64))


@pytest.mark.parametrize("wavefront", [0, 1 << 13])
def test_batch_matches_individual_renders(dev, wavefront):
    items = mixed_batch()
    objs = [pg.Scene(dev, sc) for sc, _ in items]
    params = []
    for sc, p in items:
        p.wavefront_paths = wavefront
        params.append(p)
    singles = []
    for s, (sc, p) in zip(objs, items):
        f = pg.Film(dev, sc["width"], sc["height"])
        s.render(p, f)
        singles.append(download(f, p))
        f.close()
    films = [pg.Film(dev, sc["width"], sc["height"]) for sc, _ in items]
    st, lat = pg.render_batch(dev, objs, params, films)
    total = sum(p.rect[2] * p.rect[3] * (p.spp_end - p.spp_begin) for p in params)
    assert st["paths"] == total
    for k, (f, p) in enumerate(zip(films, params)):
        got = download(f, p)
        assert np.array_equal(got, singles[k]), f"scene {k}: batched film differs from its own render"
        assert got[..., 3].sum() > 0
    assert np.all(lat > 0) and np.all(np.isfinite(lat))
    assert lat.max() <= st["ms_total"] * 1.05 + 1.0


def test_small_scene_batch_matches_individual_renders(dev):
    # 96 small Cornell boxes (64x64 to 256x256, 1 to 16 spp; plastic and several-primitive lights
    # among them) through one path pool: every film equals its own pg_render call, and the batch's
    # latencies are ordered within its total
    scs = [scenes.small_batch_scene(i) for i in range(96)]
    objs = [pg.Scene(dev, sc) for sc in scs]
    params = [pg.render_params(sc) for sc in scs]
    singles = []
    for s, p, sc in zip(objs, params, scs):
        f = pg.Film(dev, sc["width"], sc["height"])
        s.render(p, f)
        singles.append(f.download())
        f.close()
    films = [pg.Film(dev, sc["width"], sc["height"]) for sc in scs]
    st, lat = pg.render_batch(dev, objs, params, films)
    assert st["paths"] == sum(sc["width"] * sc["height"] * sc["spp"] for sc in scs)
    for f, ref, sc in zip(films, singles, scs):
        assert np.array_equal(f.download(), ref), sc["name"]
    assert np.all(lat > 0) and np.all(lat <= st["ms_total"] * 1.01)
    for f in films:
        f.close()
    for s in objs:
        s.close()


def test_batch_ring_reuse(dev):
    """Many scenes through small waves: the film record ring wraps and slots are reused."""
    objs, params, singles, films = [], [], [], []
    for k in range(10):
        sc = scenes.cornell(seed=20 + k, width=64, height=48 + 8 * (k % 3), spp=16, max_depth=4 + k % 4)
        s = pg.Scene(dev, sc)
        p = pg.render_params(sc, wavefront_paths=1 << 11)
        f = pg.Film(dev, sc["width"], sc["height"])
        s.render(p, f)
        singles.append(f.download())
        f.clear()
        objs.append(s), params.append(p), films.append(f)
    st, lat = pg.render_batch(dev, objs, params, films)
    for k, f in enumerate(films):
        assert np.array_equal(f.download(), singles[k]), k
    # the last scene's film completes last: scenes finish in batch order
    assert lat[-1] == pytest.approx(lat.max())


def test_batch_without_films_and_stats(dev):
    sc = scenes.cornell(width=32, height=32, spp=4)
    s = pg.Scene(dev, sc)
    p = pg.render_params(sc)
    st1 = s.render(p, None)
    st, lat = pg.render_batch(dev, [s, s], [p, p], None)
    assert st["paths"] == 2 * st1["paths"]
    assert st["extension_rays"] == 2 * st1["extension_rays"]
    assert st["shadow_rays"] == 2 * st1["shadow_rays"]
    assert lat.shape == (2,) and lat[0] <= lat[1]


def test_graph_replay_matches_eager(dev):
    """PG_USE_GRAPH captures a group of bounce iterations and replays it (tens of replays per render,
    with host progress reads in between): films must equal the eager launches' bit for bit, for
    single and batched renders, on every one of several consecutive renders."""
    items = mixed_batch()[:4]
    objs = [pg.Scene(dev, sc) for sc, _ in items]
    for (sc, p), s in zip(items, objs):
        p.wavefront_paths = 1 << 12  # small pool: many replays per render
        f = pg.Film(dev, sc["width"], sc["height"])
        s.render(p, f)
        eager = download(f, p)
        g = pg.render_params(sc, rect=tuple(p.rect), spp_range=(p.spp_begin, p.spp_end),
                             flags=p.flags | pg.PG_USE_GRAPH, wavefront_paths=1 << 12)
        for _ in range(3):
            f.clear()
            st = s.render(g, f)
            assert st["extend_launches"] >= 8
            assert np.array_equal(download(f, p), eager)
        f.close()
    params = [p for _, p in items]
    singles = []
    for (sc, p), s in zip(items, objs):
        f = pg.Film(dev, sc["width"], sc["height"])
        s.render(p, f)
        singles.append(download(f, p))
        f.close()
    gparams = [pg.render_params(sc, rect=tuple(p.rect), spp_range=(p.spp_begin, p.spp_end),
                                flags=p.flags | (pg.PG_USE_GRAPH if k == 0 else 0), wavefront_paths=1 << 12)
               for k, (sc, p) in enumerate(items)]
    for _ in range(2):
        films = [pg.Film(dev, sc["width"], sc["height"]) for sc, _ in items]
        pg.render_batch(dev, objs, gparams, films)
        for k, (f, p) in enumerate(zip(films, params)):
            assert np.array_equal(download(f, p), singles[k]), k
            f.close()


def test_batch_rejects_mixed_bvh_width(dev):
    sc = scenes.cornell(width=16, height=16, spp=1)
    a, b = pg.Scene(dev, sc, bvh_width=8), pg.Scene(dev, sc, bvh_width=4)
    p = pg.render_params(sc)
    with pytest.raises(pg.PgError) as e:
        pg.render_batch(dev, [a, b], [p, p])
    assert e.value.code == pg.PG_ERR_INVALID
    with pytest.raises(pg.PgError):
        pg.render_batch(dev, [a], [pg.render_params(sc, flags=pg.PG_COUNT_TRAVERSAL)])

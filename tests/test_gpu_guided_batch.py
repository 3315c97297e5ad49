"""Guided batches (pg_render_batch_guided): scenes with their own guides share one path pool, and
each bounce runs one grouped inference launch (the extension queue grouped by guide, every block
running one guide's weights). With frozen guides every film must equal the film its own
pg_render_guided call leaves, bit for bit; with training guides each guide learns from its own
scenes' paths only."""
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


def guided_items():
    """(scene, params, guide index): rooms of different sizes and depths, two scenes sharing a guide,
    a sub-rect, a sample sub-range and an FP64 film."""
    items = []
    sc = scenes.hard_room(seed=0, width=48, height=40, spp=8, tess=24)
    items.append((sc, pg.render_params(sc), 0))
    sc = scenes.hard_room(seed=1, width=64, height=48, spp=4, max_depth=9, tess=24)
    items.append((sc, pg.render_params(sc, rect=(3, 5, 50, 37)), 1))
    sc = scenes.hidden_cornell(width=40, height=40, spp=8)
    items.append((sc, pg.render_params(sc, spp_range=(2, 7)), 2))
    sc = scenes.cornell(seed=4, width=32, height=24, spp=8)
    items.append((sc, pg.render_params(sc, flags=pg.PG_FILM_FP64), 2))
    return items


def download(film, p):
    return film.download(fp64=bool(p.flags & pg.PG_FILM_FP64))


@pytest.mark.parametrize("wavefront", [0, 1 << 12])
def test_frozen_guided_batch_matches_single_guided_renders(dev, wavefront):
    items = guided_items()
    guides = [pg.Guide(dev, seed=11 + k, train_records=0) for k in range(3)]
    objs = [pg.Scene(dev, sc) for sc, _, _ in items]
    params = []
    for _, p, _ in items:
        p.wavefront_paths = wavefront
        params.append(p)
    singles = []
    for s, (sc, p, gi) in zip(objs, items):
        f = pg.Film(dev, sc["width"], sc["height"])
        s.render_guided(guides[gi], p, f)
        singles.append(download(f, p))
        f.close()
    films = [pg.Film(dev, sc["width"], sc["height"]) for sc, _, _ in items]
    stats, lat = pg.render_batch(dev, objs, params, films, guides=[guides[gi] for _, _, gi in items])
    assert stats["paths"] == sum(p.rect[2] * p.rect[3] * (p.spp_end - p.spp_begin) for p in params)
    assert np.all(lat > 0)
    for k, (f, p) in enumerate(zip(films, params)):
        got = download(f, p)
        assert np.array_equal(got, singles[k]), f"scene {k}: guided batch film differs from its own render"
        assert np.any(got[..., 3] > 0)
    for o in films + objs + guides:
        o.close()


def test_guided_batch_trains_each_guide_on_its_scenes(dev):
    items = []
    sc = scenes.cornell(seed=2, width=64, height=64, spp=16)
    items.append((sc, pg.render_params(sc), 0))
    sc = scenes.hidden_cornell(width=64, height=64, spp=16)
    items.append((sc, pg.render_params(sc), 1))
    sc = scenes.hard_room(seed=0, width=48, height=40, spp=8, tess=24)
    items.append((sc, pg.render_params(sc), 2))
    guides = [pg.Guide(dev, seed=21, train_records=1 << 14, train_batch=1 << 12),
              pg.Guide(dev, seed=22, train_records=1 << 14, train_batch=1 << 12),
              pg.Guide(dev, seed=23, train_records=0)]
    before = [g.get_params() for g in guides]
    objs = [pg.Scene(dev, sc) for sc, _, _ in items]
    stats, _ = pg.render_batch(dev, objs, [p for _, p, _ in items], None,
                               guides=[guides[gi] for _, _, gi in items])
    assert stats["train_records"] > 0
    after = [g.get_params() for g in guides]
    assert not np.array_equal(before[0], after[0]) and not np.array_equal(before[1], after[1])
    assert np.array_equal(before[2], after[2])  # frozen guide: untouched
    assert np.all(np.isfinite(after[0])) and np.all(np.isfinite(after[1]))
    for o in objs + guides:
        o.close()

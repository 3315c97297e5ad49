"""Autotuner (pg_autotune) and launch settings (pg_scene_set_launch): the autotuner must pick one of
the configurations it tried, leave the scene's BVH at that width and the launch settings it chose,
and no BVH width, path-pool capacity or launch setting may change the image (hits do not depend on
the tree or the traversal stack, and the film does not depend on the order paths finish in)."""
import itertools

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


def test_autotune_picks_a_tried_configuration_and_keeps_the_image(dev):
    sc = scenes.terrain_spheres(width=160, height=90, spp=16, grid=(200, 100), n_spheres=20)
    s = pg.Scene(dev, sc, bvh_width=8)
    film = pg.Film(dev, sc["width"], sc["height"])
    p = pg.render_params(sc)
    s.render(p, film)
    before = film.download()
    tune = s.autotune(p, sample_paths=1 << 16)
    assert tune["bvh_width"] in (4, 8)
    assert tune["wavefront_paths"] in (1 << 21, 1 << 22, 1 << 23, 1 << 24)
    la = tune["launch"]
    assert la["lds_stack"] in (6, 8, 10, 12, 16) and la["trav_waves"] in (0, 16, 20, 24)
    assert la["shade_block"] in (128, 256) and la["infer_waves"] in (4, 8, 16)
    assert s.get_launch() == la
    assert tune["trials"] >= 2 + 20 + 1 + 3 and tune["mpaths_per_s"] > 0
    assert tune["sample_paths"] == 160 * 90 * 4  # whole rect, 4 of the 16 samples
    assert s.bvh_info()["width"] == tune["bvh_width"]
    film.clear()
    s.render(pg.render_params(sc, wavefront_paths=tune["wavefront_paths"]), film)
    assert np.array_equal(film.download(), before)


def _film(s, sc, dev, guide=None):
    f = pg.Film(dev, sc["width"], sc["height"])
    if guide is None:
        s.render(pg.render_params(sc), f)
    else:
        s.render_guided(guide, pg.render_params(sc), f)
    out = f.download()
    f.close()
    return out


@pytest.mark.parametrize("width", [4, 8])
def test_every_launch_setting_keeps_the_film(dev, width):
    """Every traversal stack depth (4 spills nearly every deep stack) x traversal waves per CU, and
    both shade blocks, on a scene with every surface class; every inference block on a guided
    render with a frozen guide; a moving-instance scene at the extreme settings."""
    sc = scenes.cornell(width=48, height=40, spp=8, plastic=True, specular=True, point_lights=True)
    s = pg.Scene(dev, sc, bvh_width=width)
    ref = _film(s, sc, dev)
    for lds, waves, blk in itertools.product((4, 6, 10, 16), (0, 8, 24), (128, 256)):
        print("launch", width, lds, waves, blk, flush=True)
        s.set_launch(lds_stack=lds, trav_waves=waves, shade_block=blk)
        assert np.array_equal(_film(s, sc, dev), ref), (lds, waves, blk)
    g = pg.Guide(dev, seed=3, train_records=0)
    s.set_launch()
    gref = _film(s, sc, dev, g)
    for inf in (4, 16):
        s.set_launch(infer_waves=inf)
        assert np.array_equal(_film(s, sc, dev, g), gref), inf
    mc = scenes.moving_cornell(width=40, height=32, spp=4)
    m = pg.Scene(dev, mc, bvh_width=width)
    mref = _film(m, mc, dev)
    m.set_launch(lds_stack=4, trav_waves=8, shade_block=128)
    assert np.array_equal(_film(m, mc, dev), mref)
    with pytest.raises(pg.PgError):
        s.set_launch(lds_stack=3)
    with pytest.raises(pg.PgError):
        s.set_launch(shade_block=64)

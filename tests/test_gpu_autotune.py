"""Autotuner (pg_autotune): it must pick one of the configurations it tried, leave the scene's BVH at
that width, and change nothing in the image (hits do not depend on the tree, and the film does not
depend on the path-pool capacity)."""
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
    assert tune["trials"] >= 2 and tune["mpaths_per_s"] > 0
    assert tune["sample_paths"] == 160 * 90 * 4  # whole rect, 4 of the 16 samples
    assert s.bvh_info()["width"] == tune["bvh_width"]
    film.clear()
    s.render(pg.render_params(sc, wavefront_paths=tune["wavefront_paths"]), film)
    assert np.array_equal(film.download(), before)

"""Traversal work counted twice: by the GPU kernel (PG_COUNT_TRAVERSAL) and by the CPU oracle
walking the GPU's own wide BVH (downloaded with pg_scene_bvh_download) in the kernel's order. The
per-ray node and triangle counts that the roofline's algorithmic bytes rest on must agree, and the
oracle's walk of the wide tree must give the same hits as its own binary tree."""
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


@pytest.fixture(scope="module")
def terrain():
    return scenes.terrain_spheres()


def _rays(rng, n):
    o = np.stack([rng.uniform(-9, 9, n), rng.uniform(0.3, 3.0, n), rng.uniform(-4.5, 4.5, n)], 1)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d, np.full((n, 1), 1e-3), np.full((n, 1), 1e5)], 1).astype(np.float32)


@pytest.mark.parametrize("width", [4, 8])
def test_wide_tree_walk_gives_the_binary_tree_hits(dev, terrain, width):
    s = pg.Scene(dev, terrain, bvh_width=width)
    nodes, tris = s.bvh_download()
    own = OracleScene(terrain)
    wide = OracleScene(terrain)
    wide.use_wide_bvh(width, nodes, tris)
    rays = _rays(np.random.default_rng(8), 20000)
    a, b = own.trace(rays), wide.trace(rays)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert np.array_equal(wide.occluded(rays), own.occluded(rays))
    g = s.trace(rays)
    for x, y in zip(g, b):
        assert np.array_equal(x, y)
    s.close()


@pytest.mark.parametrize("width", [4, 8])
def test_kernel_traversal_counts_match_oracle_walk(dev, terrain, width):
    """Nodes visited and triangles tested by the extension rays of a 96x64 crop at 4 spp of the
    1.25M-triangle frame: the kernel's counters against the oracle walking the same tree for the
    same paths (a rare path that diverges by an ulp may trace a different bounce, so 0.5%)."""
    s = pg.Scene(dev, terrain, bvh_width=width)
    rect = (800, 400, 96, 64)
    st = s.render(pg.render_params(terrain, rect=rect, spp_range=(0, 4), flags=pg.PG_COUNT_TRAVERSAL), None)
    o = OracleScene(terrain)
    o.use_wide_bvh(width, *s.bvh_download())
    r = o.render(rect=rect, spp_range=(0, 4), film=False, threads=16)
    ext = int(r["rays"][0])
    assert abs(int(st["extension_rays"]) - ext) <= 1e-3 * ext
    for key in ("nodes_visited", "tris_tested"):
        g, c = int(st[key]), r[key]
        assert c > ext
        assert abs(g - c) <= 5e-3 * c, (key, g, c)
    s.close()

"""Moving instances on the GPU (BLAS per instance, ray time = Sobol dimension 2) against the CPU
oracle and the moving-occluder closed form, at both BVH widths."""
import numpy as np
import pytest

import piper2_amd as pg
from oracle_lib import OracleScene
from piper2_amd import scenes
from test_motion import path_rays_and_times

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = pg.Device(0)
    yield d
    d.close()


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("kind,t_range", [("translate", (0.25, 0.75)), ("rotate", (-0.5, 1.5))])
def test_moving_occluder_paths_match_closed_form(dev, width, kind, t_range):
    sc = scenes.motion_occluder(kind, t_range=t_range)
    s = pg.Scene(dev, sc, bvh_width=width)
    _, pl = s.render(pg.render_params(sc), None, path_l=True)
    o = OracleScene(sc)
    _, rays, times = path_rays_and_times(o, sc)
    blocked, margin = scenes.occluder_blocked(sc["occluder"], rays, times)
    sure = margin > 1e-4
    expect = np.where(blocked, 0.0, 1.0)
    assert np.all(pl[sure] == expect[sure, None])
    ref = o.render(path_l=True, film=False, threads=16)["path_l"]
    assert np.mean(np.all(pl == ref, axis=1)) > 0.999
    s.close()


@pytest.mark.parametrize("width", [4, 8])
def test_trace_at_matches_oracle(dev, width):
    """Random rays at random times through the Cornell box with its moving box: the same closest
    primitive as the oracle (which maps rays into the instance in double) for all but grazing
    rays, and the same hit distance and barycentrics."""
    sc = scenes.moving_cornell(width=16, height=16, spp=1)
    s = pg.Scene(dev, sc, bvh_width=width)
    o = OracleScene(sc)
    rng = np.random.default_rng(11)
    n = 200000
    org = rng.uniform(-0.95, 0.95, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([org, d, np.full((n, 1), 1e-3), np.full((n, 1), 1e5)], 1).astype(np.float32)
    times = rng.uniform(0.0, 1.0, n).astype(np.float32)
    tg, pgp, uvg = s.trace_at(rays, times)
    to, po, uvo = o.trace_at(rays, times)
    first = int(sc["instance_first"][0])
    inst = (po >= first) & (po < 0x80000000)
    assert inst.sum() > 5000, inst.sum()
    same = pgp == po
    assert same.mean() > 0.9995, same.mean()
    assert same[inst].mean() > 0.999, same[inst].mean()
    assert np.allclose(tg[same], to[same], rtol=1e-4, atol=1e-6)
    assert np.allclose(uvg[same & inst], uvo[same & inst], atol=2e-3)
    # static hits do not depend on the time: pg_trace at time 0 equals pg_trace_at there
    t0g, p0g, _ = s.trace(rays[:5000])
    t0a, p0a, _ = s.trace_at(rays[:5000], np.zeros(5000, np.float32))
    assert np.array_equal(p0g, p0a) and np.array_equal(t0g, t0a)
    s.close()


@pytest.mark.parametrize("width", [4, 8])
def test_moving_cornell_paths_match_oracle(dev, width):
    """Full light transport with the sliding, turning, growing box: NEE shadow rays, MIS rays and
    hit normals all at the path's time."""
    sc = scenes.moving_cornell(width=48, height=48, spp=8)
    s = pg.Scene(dev, sc, bvh_width=width)
    _, pl_g = s.render(pg.render_params(sc), None, path_l=True)
    pl_c = OracleScene(sc).render(path_l=True, film=False, threads=16)["path_l"]
    close = np.all(np.abs(pl_g - pl_c) <= 1e-3 * np.abs(pl_c) + 1e-5, axis=1)
    assert close.mean() >= 0.98, close.mean()
    s.close()


def test_set_motion_between_renders(dev):
    """New keys take effect on the next render; visible = 0 removes the instance; both match the
    oracle given the same keys."""
    sc = scenes.motion_occluder("translate", t_range=(0.0, 1.0))
    s = pg.Scene(dev, sc)
    o = OracleScene(sc)
    keys = [scenes.srt(trans=(1.0, 0.5, -5.0)), scenes.srt(trans=(-1.0, -0.5, -5.0))]
    s.set_motion(0, keys)
    o.set_motion(0, keys)
    _, a = s.render(pg.render_params(sc), None, path_l=True)
    ref = o.render(path_l=True, film=False, threads=16)["path_l"]
    assert np.mean(np.all(a == ref, axis=1)) > 0.999
    assert 0.03 < np.mean(a[:, 0] == 0.0) < 0.6
    s.set_motion(0, keys, visible=False)
    _, b = s.render(pg.render_params(sc), None, path_l=True)
    assert np.all(b == 1.0)
    s.close()


def test_motion_batch_equals_single_renders(dev):
    """A batch of a moving-instance scene and a static one leaves each film equal to its own
    render."""
    mc = scenes.moving_cornell(width=40, height=32, spp=4)
    cb = scenes.cornell(width=32, height=32, spp=4)
    objs = [pg.Scene(dev, mc), pg.Scene(dev, cb)]
    single = []
    for o, sc in zip(objs, (mc, cb)):
        f = pg.Film(dev, sc["width"], sc["height"])
        o.render(pg.render_params(sc), f)
        single.append(f.download())
        f.close()
    films = [pg.Film(dev, sc["width"], sc["height"]) for sc in (mc, cb)]
    pg.render_batch(dev, objs, [pg.render_params(mc), pg.render_params(cb)], films)
    for f, ref in zip(films, single):
        assert np.array_equal(f.download(), ref)
        f.close()


def test_guided_render_of_moving_scene_is_unbiased(dev):
    """The guide's inputs use the moving triangles' world normals at the path's time: a frozen
    untrained guide's estimate has the plain estimate's mean."""
    sc = scenes.moving_cornell(width=32, height=32, spp=64)
    s = pg.Scene(dev, sc)
    g = pg.Guide(dev, seed=5, train_records=0)
    fg, fp = pg.Film(dev, 32, 32), pg.Film(dev, 32, 32)
    s.render_guided(g, pg.render_params(sc), fg)
    s.render(pg.render_params(sc), fp)
    a, b = fg.download()[..., :3].astype(np.float64), fp.download()[..., :3].astype(np.float64)
    assert np.all(np.isfinite(a))
    d = (a - b).reshape(-1)
    assert abs(d.mean()) < 4 * d.std() / np.sqrt(d.size) + 1e-3 * abs(b.mean()), (d.mean(), d.std())

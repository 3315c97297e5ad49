"""Device BVH builder (pg_bvh_params.builder = PG_BVH_DEVICE): the tree must be a valid wide BVH
over every triangle exactly once, give the same hits and the same film as the host-built tree
(hits do not depend on the tree), be deterministic, and cost about as much to traverse."""
import numpy as np
import pytest

import piper2_amd as pg
from piper2_amd import scenes

pytestmark = pytest.mark.gpu

LEAF_BIT = 0x80000000
INVALID = 0xFFFFFFFF


@pytest.fixture(scope="module")
def dev():
    d = pg.Device(0)
    yield d
    d.close()


def check_tree(s, sc):
    """Walk the downloaded tree: child refs in range, leaves cover every leaf-order triangle once,
    leaf triangles inside their (padded) child box, leaf-order triangles a permutation."""
    info = s.bvh_info()
    W = info["width"]
    nodes, tris = s.bvh_download()
    nodes = nodes.reshape(info["nodes"], 7 * W)
    refs = nodes[:, 6 * W:].view(np.uint32)
    boxes = nodes[:, :6 * W].reshape(-1, 6, W)
    n = sc["indices"].shape[0]
    tv = tris.reshape(n, 3, 4)
    prim = tv[:, 0, 3].view(np.uint32)
    assert np.array_equal(np.sort(prim), np.arange(n, dtype=np.uint32))
    world = sc["vertices"][sc["indices"]]
    assert np.array_equal(tv[:, :, :3], world[prim])
    covered = np.zeros(n, np.int32)
    seen = np.zeros(info["nodes"], np.int32)
    stack, depth_max = [(0, 1)], 0
    while stack:
        k, depth = stack.pop()
        seen[k] += 1
        depth_max = max(depth_max, depth)
        for c in range(W):
            r = int(refs[k, c])
            if r == INVALID:
                continue
            lo, hi = boxes[k, 0:3, c], boxes[k, 3:6, c]
            if r & LEAF_BIT:
                cnt = ((r >> 27) & 0x7) + 1  # leaf ref: bit 31 | (count - 1) << 27 | first
                first = r & ((1 << 27) - 1)
                covered[first:first + cnt] += 1
                v = tv[first:first + cnt, :, :3].reshape(-1, 3)
                assert np.all(v >= lo - 1e-6) and np.all(v <= hi + 1e-6)
            else:
                assert 0 < r < info["nodes"]
                stack.append((r, depth + 1))
    assert np.all(covered == 1)
    assert np.all(seen == 1)  # a tree: every node reached exactly once
    assert depth_max == info["max_depth"]
    return info


def random_rays(sc, n, seed):
    rng = np.random.default_rng(seed)
    v = sc["vertices"]
    lo, hi = v.min(0), v.max(0)
    o = rng.uniform(lo - 0.2 * (hi - lo), hi + 0.2 * (hi - lo), (n, 3)).astype(np.float32)
    tgt = v[sc["indices"][rng.integers(0, sc["indices"].shape[0], n)]].mean(1)
    d = tgt - o + rng.normal(0, 0.05, (n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, :3], r[:, 3:6], r[:, 6], r[:, 7] = o, d, 1e-3, 1e5
    return r


def degenerate_scene():
    """2,600 copies of one small triangle (coincident centroids) plus a ground quad and a light:
    the SAH finds no split in the copies, so the builder must split by position."""
    sc = scenes.cornell(width=32, height=32, spp=1)
    v, t = sc["vertices"], sc["indices"]
    tri = np.array([[0.0, -0.5, 0.0], [0.2, -0.5, 0.0], [0.0, -0.3, 0.1]], np.float32)
    extra_v = np.tile(tri, (2600, 1))
    extra_t = (np.arange(2600 * 3, dtype=np.uint32) + v.shape[0]).reshape(-1, 3)
    sc["vertices"] = np.ascontiguousarray(np.concatenate([v, extra_v]).astype(np.float32))
    sc["indices"] = np.ascontiguousarray(np.concatenate([t, extra_t]).astype(np.uint32))
    sc["tri_object"] = np.ascontiguousarray(np.concatenate([sc["tri_object"], np.zeros(2600, np.uint32)]))
    return sc


CASES = {
    "cornell": lambda: scenes.cornell(width=64, height=64, spp=4),
    "terrain_300k": lambda: scenes.terrain_spheres(width=96, height=54, spp=2, grid=(300, 150), n_spheres=150),
    "degenerate": degenerate_scene,
}


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("case", sorted(CASES))
def test_device_bvh_valid_and_same_hits(dev, case, width):
    sc = CASES[case]()
    host = pg.Scene(dev, sc, bvh_width=width, builder=pg.PG_BVH_HOST)
    gpu = pg.Scene(dev, sc, bvh_width=width, builder=pg.PG_BVH_DEVICE)
    info = check_tree(gpu, sc)
    assert info["width"] == width
    rays = random_rays(sc, 20000, seed=len(case) + width)
    th, ph, uh = host.trace(rays)
    tg, pg_, ug = gpu.trace(rays)
    assert np.array_equal(ph, pg_) and np.array_equal(th, tg) and np.array_equal(uh, ug)
    p = pg.render_params(sc)
    fh, fg = pg.Film(dev, sc["width"], sc["height"]), pg.Film(dev, sc["width"], sc["height"])
    host.render(p, fh)
    gpu.render(p, fg)
    assert np.array_equal(fh.download(), fg.download())


def test_device_bvh_deterministic_and_traversal_cost(dev):
    sc = CASES["terrain_300k"]()
    a = pg.Scene(dev, sc, bvh_width=4, builder=pg.PG_BVH_DEVICE)
    b = pg.Scene(dev, sc, bvh_width=4, builder=pg.PG_BVH_DEVICE)
    na, ta = a.bvh_download()
    nb, tb = b.bvh_download()
    assert np.array_equal(na.view(np.uint32), nb.view(np.uint32)) and np.array_equal(ta.view(np.uint32), tb.view(np.uint32))
    host = pg.Scene(dev, sc, bvh_width=4, builder=pg.PG_BVH_HOST)
    cnt = pg.render_params(sc, flags=pg.PG_COUNT_TRAVERSAL)
    sh, sg = host.render(cnt, None), a.render(cnt, None)
    cost = lambda st: st["nodes_visited"] + 0.5 * st["tris_tested"]  # noqa: E731
    assert cost(sg) <= 1.25 * cost(sh), (sg["nodes_visited"], sh["nodes_visited"])


def test_device_bvh_single_triangle_and_errors(dev):
    sc = scenes.cornell(width=16, height=16, spp=1)
    sc["indices"] = np.ascontiguousarray(sc["indices"][:1])
    sc["tri_object"] = np.ascontiguousarray(sc["tri_object"][:1])
    s = pg.Scene(dev, sc, bvh_width=8, builder=pg.PG_BVH_DEVICE)
    info = check_tree(s, sc)
    assert info["nodes"] == 1 and info["leaves"] == 1
    with pytest.raises(pg.PgError) as e:
        s.build_bvh(width=8, builder=7)
    assert e.value.code == pg.PG_ERR_INVALID
    with pytest.raises(pg.PgError):
        s.build_bvh(width=6, builder=pg.PG_BVH_DEVICE)

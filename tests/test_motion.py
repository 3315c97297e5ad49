"""Moving instances (Group children with keyframed TransformSRT, Embree.cpp:706-835) on the CPU
oracle, pinned by closed forms: a black occluder moving in front of a white environment gives each
path radiance 1 or 0 according to where the occluder is at the path's time (Sobol dimension 2,
Embree/Kernel.cpp:77); an instance at a fixed transform traces like its mesh flattened into world
space with T R S computed independently in numpy."""
import math

import numpy as np
import pytest

from oracle_lib import OracleScene, interval_to_index, sobol_sample
from piper2_amd import scenes


def path_rays_and_times(o, sc):
    """Camera ray and time of every path of a render, in the render's path order (pixel rows,
    then samples)."""
    w, h, spp = sc["width"], sc["height"], sc["spp"]
    m = 0
    while (1 << m) < max(w, h):
        m += 1
    pix = np.array([[x, y, s] for y in range(h) for x in range(w) for s in range(spp)], np.uint32)
    rays, _ = o.camera_rays(pix)
    times = np.array([sobol_sample(interval_to_index(m, int(s), int(x), int(y)), 2) for x, y, s in pix], np.float32)
    return pix, rays, times


@pytest.mark.parametrize("kind,t_range", [("translate", (0.0, 1.0)), ("translate", (0.25, 0.75)),
                                          ("rotate", (0.0, 1.0)), ("rotate", (-0.5, 1.5))])
def test_moving_occluder_closed_form(kind, t_range):
    """Every path is 0 exactly where the occluder covers its ray at its time: linear translation,
    constant-speed rotation (spherical interpolation; a normalised lerp of these keys is 0.9
    degrees off at a quarter of the turn), keys reaching past the shutter, and an instance present
    only for t0 <= time <= t1."""
    sc = scenes.motion_occluder(kind, t_range=t_range)
    o = OracleScene(sc)
    pl = o.render(path_l=True, film=False, threads=4)["path_l"]
    _, rays, times = path_rays_and_times(o, sc)
    blocked, margin = scenes.occluder_blocked(sc["occluder"], rays, times)
    sure = margin > 1e-4
    assert sure.mean() > 0.98
    expect = np.where(blocked, 0.0, 1.0)
    assert np.all(pl[sure] == expect[sure, None]), np.argwhere(pl[sure][:, 0] != expect[sure])[:5]
    assert 0.03 < blocked.mean() < 0.6, blocked.mean()
    if t_range == (0.25, 0.75):  # no path outside the key range is ever blocked
        assert not np.any(blocked[(times < 0.25) | (times > 0.75)])
        assert np.all(pl[(times < 0.249) | (times > 0.751)] == 1.0)


def test_rotation_is_spherical_not_normalised_lerp():
    """The occluder closed form under a normalised lerp of the two keys disagrees with the oracle on
    paths near the plank's ends, so the rotation test above does pin slerp."""
    sc = scenes.motion_occluder("rotate", spp=64)
    o = OracleScene(sc)
    pl = o.render(path_l=True, film=False, threads=4)["path_l"][:, 0]
    _, rays, times = path_rays_and_times(o, sc)
    blocked, _ = scenes.occluder_blocked(sc["occluder"], rays, times)
    o, d = np.asarray(rays, np.float64)[:, 0:3], np.asarray(rays, np.float64)[:, 3:6]
    s = (-5.0 - o[:, 2]) / d[:, 2]
    X, Y = o[:, 0] + d[:, 0] * s, o[:, 1] + d[:, 1] * s
    f = np.asarray(times, np.float64)
    h = math.sqrt(0.5)
    th = 2.0 * np.arctan2(f * h, (1.0 - f) + f * h)  # angle of normalize(lerp(q0, q1, f))
    a = np.abs(X * np.cos(th) + Y * np.sin(th))
    b = np.abs(-X * np.sin(th) + Y * np.cos(th))
    nlerp_blocked = (a <= 1.0) & (b <= 0.15)
    wrong = np.sum((pl == 0.0) != nlerp_blocked)
    assert wrong > 40 and wrong > 10 * np.sum((pl == 0.0) != blocked), wrong


def _srt_matrix(k):
    """T R S of a 16-float TransformSRT key in float64 (rotation of the normalised quaternion)."""
    k = np.asarray(k, np.float64)
    w, x, y, z = k[9:13] / np.linalg.norm(k[9:13])
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    S = np.array([[k[0], k[3], k[4], k[6]], [0, k[1], k[5], k[7]], [0, 0, k[2], k[8]]])
    M = R @ S
    M[:, 3] += k[13:16]
    return M


def test_fixed_instance_traces_like_flattened_mesh():
    """One key (a fixed transform with rotation, non-uniform scale, skew and shift) under a parent
    transform: hits equal those of the mesh flattened to world space."""
    sc = scenes.moving_cornell(width=16, height=16, spp=1)
    key = scenes.srt(scale=(1.2, 0.8, 1.1), quat=(0.9, 0.1, 0.3, -0.2), trans=(0.1, -0.6, 0.05),
                     skew=(0.1, 0.0, -0.05), shift=(0.02, 0.0, 0.01))
    parent = np.array([[0.9, 0.0, 0.1, 0.05], [0.0, 1.0, 0.0, -0.1], [-0.1, 0.0, 0.9, 0.0]])
    sc["motions"] = [dict(keys=[key], parent=parent.reshape(-1))]
    first = int(sc["instance_first"][0])
    flat = dict(sc)
    verts = sc["vertices"].astype(np.float64)
    idx = sc["indices"]
    moving = np.unique(idx[first:])
    M = np.vstack([parent, [0, 0, 0, 1]]) @ np.vstack([_srt_matrix(key), [0, 0, 0, 1]])
    v2 = verts.copy()
    v2[moving] = verts[moving] @ M[:3, :3].T + M[:3, 3]
    flat["vertices"] = v2.astype(np.float32)
    del flat["instance_first"], flat["motions"]
    oi, of = OracleScene(sc), OracleScene(flat)
    rng = np.random.default_rng(3)
    n = 20000
    o = rng.uniform(-0.9, 0.9, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, d, np.full((n, 1), 1e-3), np.full((n, 1), 1e5)], 1).astype(np.float32)
    ti, pi, _ = oi.trace_at(rays, rng.uniform(0, 1, n))
    tf, pf, _ = of.trace(rays)
    inst_hits = (pf >= first) & (pf < 0x80000000)
    assert inst_hits.sum() > 500
    same = pi == pf
    assert same.mean() > 0.999, same.mean()
    assert np.allclose(ti[same], tf[same], rtol=2e-5, atol=1e-6)


def test_moving_cornell_paths_are_finite_and_time_dependent():
    """The lit scene with a sliding, turning, growing box: renders are finite, and moving the keys
    changes the image only where the box's sweep reaches (the walls' far corners keep their paths)."""
    sc = scenes.moving_cornell(width=24, height=24, spp=4)
    o = OracleScene(sc)
    a = o.render(path_l=True, film=False, threads=4)["path_l"]
    assert np.all(np.isfinite(a))
    o.set_motion(0, [scenes.srt(trans=(0.0, -1.0, 0.0))])
    b = o.render(path_l=True, film=False, threads=4)["path_l"]
    diff = np.any(a != b, axis=1)
    assert 0.02 < diff.mean() < 0.9, diff.mean()

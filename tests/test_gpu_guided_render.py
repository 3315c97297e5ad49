"""Guided rendering on the GPU: unbiasedness of the one-sample MIS estimator, and what the online
trained network gains over the plain BSDF-sampling tracer and over an untrained network.

relMSE = mean (I - R)^2 / (R^2 + 0.01) against a 4096-spp FP64 plain render R.
"""
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


def img(f):
    return f[..., :3] / np.maximum(f[..., 3:], 1e-30)


def rel_mse(a, b):
    return float(np.mean((a - b) ** 2 / (b * b + 0.01)))


def test_guided_estimator_is_unbiased(dev):
    sc = scenes.hidden_cornell(width=64, height=64, spp=4096)
    s = pg.Scene(dev, sc)
    film = pg.Film(dev, 64, 64)
    s.render(pg.render_params(sc, spp=4096, flags=pg.PG_FILM_FP64), film)
    plain = img(film.download(fp64=True))
    g = pg.Guide(dev, seed=11, train_records=0)  # random network: a poor but valid proposal
    film.clear()
    s.render_guided(g, pg.render_params(sc, spp=4096, flags=pg.PG_FILM_FP64), film)
    guided = img(film.download(fp64=True))
    assert abs(guided.mean() - plain.mean()) / plain.mean() < 0.01
    blocks = lambda a: a.reshape(8, 8, 8, 8, 3).mean(axis=(1, 3))  # noqa: E731
    rel = np.abs(blocks(guided) - blocks(plain)) / blocks(plain)
    assert np.median(rel) < 0.02


def test_learned_selection_estimator_is_unbiased(dev):
    # a random network whose selection row is live: every vertex samples the guide with its own
    # probability alpha in [0.05, 0.95], and the combined density keeps the estimator unbiased
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
    import guide_oracle as go
    sc = scenes.hidden_cornell(width=64, height=64, spp=4096)
    s = pg.Scene(dev, sc)
    film = pg.Film(dev, 64, 64)
    s.render(pg.render_params(sc, spp=4096, flags=pg.PG_FILM_FP64), film)
    plain = img(film.download(fp64=True))
    g = pg.Guide(dev, seed=12, train_records=0, mis_fraction=-1.0)
    th = g.get_params()
    k4 = sum(o * i + o for o, i in go.LAYERS[:-1])
    rng = np.random.default_rng(12)
    th[k4 + go.ALPHA_ROW * go.NHID:k4 + (go.ALPHA_ROW + 1) * go.NHID] = rng.normal(0, 0.5, go.NHID)
    g.set_params(th)
    film.clear()
    s.render_guided(g, pg.render_params(sc, spp=4096, flags=pg.PG_FILM_FP64), film)
    guided = img(film.download(fp64=True))
    assert abs(guided.mean() - plain.mean()) / plain.mean() < 0.01
    blocks = lambda a: a.reshape(8, 8, 8, 8, 3).mean(axis=(1, 3))  # noqa: E731
    rel = np.abs(blocks(guided) - blocks(plain)) / blocks(plain)
    assert np.median(rel) < 0.02


def test_trained_guide_beats_plain_and_untrained(dev):
    sc = scenes.hidden_cornell(width=128, height=128, spp=4096)
    s = pg.Scene(dev, sc)
    film = pg.Film(dev, 128, 128)
    s.render(pg.render_params(sc, spp=4096, flags=pg.PG_FILM_FP64), film)
    ref = img(film.download(fp64=True))

    def rel_at(render, spp=64):
        film.clear()
        render(pg.render_params(sc, spp=4096, spp_range=(0, spp)))
        return rel_mse(img(film.download()), ref)

    plain = rel_at(lambda p: s.render(p, film))
    untrained = pg.Guide(dev, seed=1, train_records=0)
    e_untrained = rel_at(lambda p: s.render_guided(untrained, p, film))
    g = pg.Guide(dev, seed=1, train_records=1 << 16, train_batch=1 << 14)
    for k in range(12):  # online training passes on samples the final render does not use
        st = s.render_guided(g, pg.render_params(sc, spp=4096, spp_range=(2048 + 4 * k, 2052 + 4 * k)), None)
        assert st["train_records"] > 1000
    trained = pg.Guide(dev, seed=1, train_records=0)
    trained.set_params(g.get_params())
    e_trained = rel_at(lambda p: s.render_guided(trained, p, film))
    assert e_trained < 0.8 * plain, (e_trained, plain)
    assert e_trained < 0.8 * e_untrained, (e_trained, e_untrained)

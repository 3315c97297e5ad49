"""GPU parity of the path-guiding network (MFMA BF16 kernels) against the FP64 oracle.

The oracle rounds weights and layer inputs to bfloat16 exactly as the kernels feed the matrix cores
and sums in FP64; the kernels sum in FP32 in hardware order, so a value that lands near a bfloat16
rounding boundary may round the other way. Tolerances: mixture parameters and pdf 2e-2 relative,
gradients cosine similarity > 0.995 and relative norm error < 3e-2 (written per test).
"""
import os
import sys

import numpy as np
import pytest

import piper2_amd as pg

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import guide_oracle as go  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = pg.Device(0)
    yield d
    d.close()


def _unit(rng, n):
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def _inputs(rng, n):
    return rng.uniform(0, 1, (n, 3)).astype(np.float32), _unit(rng, n), _unit(rng, n)


def test_param_count_and_init_layout(dev):
    g = pg.Guide(dev, seed=3)
    th = g.get_params().astype(np.float64)
    assert th.size == go.param_count() == 24880
    layers = go.unpack(th)
    assert np.all(layers[0][0][:, go.PAD_FEATURES] == 0) and np.all(layers[4][0][go.PAD_OUTPUTS] == 0)
    assert np.count_nonzero(layers[0][0]) == 64 * (144 - len(go.PAD_FEATURES))
    assert all(np.all(b == 0) for _, b in layers)


def test_forward_mixture_matches_oracle(dev):
    rng = np.random.default_rng(1)
    g = pg.Guide(dev, seed=1)
    th = g.get_params().astype(np.float64)
    th += rng.normal(0, 0.05, th.shape) * (th != 0)  # non-trivial biases stay zero, weights move
    layers = go.unpack(th)
    for _, b in layers:
        b[:] = rng.normal(0, 0.1, b.shape)
    go.unpack(th)[4][1][go.PAD_OUTPUTS] = 0.0
    g.set_params(th.astype(np.float32))
    th = g.get_params().astype(np.float64)
    n = 2000
    p01, nrm, wo = _inputs(rng, n)
    mix = g.infer(p01, nrm, wo)
    out, _ = go.forward(th, go.encode(p01.astype(np.float64), nrm.astype(np.float64), wo.astype(np.float64)))
    mu, kappa, pi, _ = go.mixture(out)
    gmu = mix[:, :32].reshape(n, 8, 4)[:, :, :3]
    gk = mix[:, :32].reshape(n, 8, 4)[:, :, 3]
    gpi = mix[:, 32:40]
    assert np.allclose(gpi.sum(1), 1.0, atol=1e-5)
    norm = gk / (2 * np.pi * -np.expm1(-2.0 * gk.astype(np.float64)))
    assert np.allclose(mix[:, 40:48], gpi * norm, rtol=1e-5, atol=1e-30)
    assert np.median(np.abs(gk - kappa) / kappa) < 2e-2
    assert np.median(np.abs(gpi - pi)) < 2e-3
    cosmu = (gmu * mu).sum(-1)
    assert np.median(cosmu) > 0.9995
    # mixture pdf at random directions
    w = _unit(rng, n).astype(np.float64)
    p_or = go.pdf(th, go.encode(p01.astype(np.float64), nrm.astype(np.float64), wo.astype(np.float64)), w)
    c = np.einsum("nlk,nk->nl", gmu.astype(np.float64), w)
    p_gpu = (gpi * go.vmf(gk.astype(np.float64), c)).sum(1)
    rel = np.abs(p_gpu - p_or) / p_or
    assert np.median(rel) < 2e-2 and np.quantile(rel, 0.99) < 0.2


def test_inference_is_repeatable_and_independent_of_sample_order(dev):
    """Every sample's mixture depends on its own inputs only: the same inputs give bit-identical
    records in any order and on every call (the kernel gathers the next 64 samples while its matrix
    tiles run, and a sample lands in any column of a 16-sample tile)."""
    rng = np.random.default_rng(6)
    g = pg.Guide(dev, seed=6, train_records=0)
    n = 5000
    p01, nrm, wo = _inputs(rng, n)
    a = g.infer(p01, nrm, wo)
    perm = rng.permutation(n)
    b = np.empty_like(a)
    b[perm] = g.infer(p01[perm], nrm[perm], wo[perm])
    assert np.array_equal(a, b)
    assert np.array_equal(a, g.infer(p01, nrm, wo))
    head = g.infer(p01[:77], nrm[:77], wo[:77])  # ragged: a partial tile and a partial wave
    assert np.array_equal(head, a[:77])


def test_gradients_match_oracle(dev):
    rng = np.random.default_rng(2)
    g = pg.Guide(dev, seed=2)
    th = g.get_params().astype(np.float64)
    n = 4096
    p01, nrm, wo = _inputs(rng, n)
    d = _unit(rng, n)
    c = rng.uniform(0.0, 2.0, n).astype(np.float32)
    c[::7] = 0.0
    loss_g, grad_g = g.gradients(pg.Guide.records(p01, nrm, wo, d, c))
    L, grad_o = go.loss_and_grad(th, go.encode(p01.astype(np.float64), nrm.astype(np.float64), wo.astype(np.float64)),
                                 d.astype(np.float64), c.astype(np.float64))
    assert abs(loss_g - L) < 2e-3 * abs(L)
    cos = np.dot(grad_g, grad_o) / (np.linalg.norm(grad_g) * np.linalg.norm(grad_o))
    assert cos > 0.995, cos
    assert np.linalg.norm(grad_g - grad_o) / np.linalg.norm(grad_o) < 3e-2
    # per layer as well (a transposed operand in one layer would show here)
    k = 0
    for o, i in go.LAYERS:
        sl = slice(k, k + o * i)
        a, b = grad_g[sl], grad_o[sl]
        if np.linalg.norm(b) > 0:
            assert np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30) > 0.99, (o, i)
        k += o * i + o
    # deterministic: the same batch gives bit-identical gradients
    assert np.array_equal(g.gradients(pg.Guide.records(p01, nrm, wo, d, c))[1], grad_g)


def test_learned_selection_gradients_match_oracle(dev):
    # mis_fraction < 0: the loss is that of alpha p_guide + (1 - alpha) p_bsdf with alpha the
    # network's selection output; its gradients (the selection row's included) match the oracle's
    rng = np.random.default_rng(6)
    g = pg.Guide(dev, seed=6, mis_fraction=-1.0)
    th = g.get_params().astype(np.float64)
    k4 = sum(o * i + o for o, i in go.LAYERS[:-1])
    th[k4 + go.ALPHA_ROW * go.NHID:k4 + (go.ALPHA_ROW + 1) * go.NHID] = rng.normal(0, 0.2, go.NHID)
    g.set_params(th.astype(np.float32))
    th = g.get_params().astype(np.float64)
    n = 4096
    p01, nrm, wo = _inputs(rng, n)
    d = _unit(rng, n)
    c = rng.uniform(0.0, 2.0, n).astype(np.float32)
    pb = rng.uniform(0.05, 1.0, n).astype(np.float32)
    loss_g, grad_g = g.gradients(pg.Guide.records(p01, nrm, wo, d, c, bsdf_pdf=pb))
    L, grad_o = go.loss_and_grad(th, go.encode(p01.astype(np.float64), nrm.astype(np.float64), wo.astype(np.float64)),
                                 d.astype(np.float64), c.astype(np.float64), pb=pb.astype(np.float64))
    assert abs(loss_g - L) < 2e-3 * abs(L)
    cos = np.dot(grad_g, grad_o) / (np.linalg.norm(grad_g) * np.linalg.norm(grad_o))
    assert cos > 0.995, cos
    arow = slice(k4 + go.ALPHA_ROW * go.NHID, k4 + (go.ALPHA_ROW + 1) * go.NHID)
    a, b = grad_g[arow], grad_o[arow]
    assert np.linalg.norm(b) > 0 and np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b)) > 0.99
    # the mixtures carry the learned selection probability, inside its bounds
    mix = g.infer(p01, nrm, wo)
    al = go.selection(go.forward(th, go.encode(p01.astype(np.float64), nrm.astype(np.float64),
                                               wo.astype(np.float64)))[0])
    assert np.all((mix[:, 48] >= go.ALO) & (mix[:, 48] <= go.AHI))
    assert np.median(np.abs(mix[:, 48] - al)) < 5e-3
    assert not np.any(mix[:, 49:52])


def test_adam_step_matches_oracle(dev):
    rng = np.random.default_rng(4)
    g = pg.Guide(dev, seed=4, lr=1e-2, train_batch=1024)
    th0 = g.get_params().astype(np.float64)
    n = 1024
    p01, nrm, wo = _inputs(rng, n)
    rec = pg.Guide.records(p01, nrm, wo, _unit(rng, n), rng.uniform(0.1, 1.0, n).astype(np.float32))
    _, grad = g.gradients(rec)
    g.train(rec)
    th1 = g.get_params().astype(np.float64)
    ref, _, _ = go.adam(th0, grad.astype(np.float64), np.zeros_like(th0), np.zeros_like(th0), 1, lr=1e-2, b1=0.9,
                        b2=0.99, eps=1e-8)
    assert np.allclose(th1, ref, atol=2e-6, rtol=1e-5)


def test_adam_steps_carry_moments_and_bias_correction(dev):
    """Five Adam steps on different batches: the GPU parameters follow the oracle's Adam applied to
    the GPU's own gradients with the moments carried over and the bias correction of each step
    (a missing correction is invisible at step 1 but moves step 2 by 35%)."""
    rng = np.random.default_rng(14)
    g = pg.Guide(dev, seed=14, lr=1e-2, train_batch=1024)
    th = g.get_params().astype(np.float64)
    m, v = np.zeros_like(th), np.zeros_like(th)
    n = 1024
    for t in range(1, 6):
        p01, nrm, wo = _inputs(rng, n)
        rec = pg.Guide.records(p01, nrm, wo, _unit(rng, n), rng.uniform(0.1, 1.0, n).astype(np.float32))
        _, grad = g.gradients(rec)
        g.train(rec)
        th, m, v = go.adam(th, grad.astype(np.float64), m, v, t, lr=1e-2, b1=0.9, b2=0.99, eps=1e-8)
        got = g.get_params().astype(np.float64)
        assert np.allclose(got, th, atol=1e-5, rtol=1e-5), (t, np.abs(got - th).max())
        th = got  # continue from the GPU's FP32 parameters


def test_training_learns_a_directional_target(dev):
    """Radiance arriving from one position-dependent direction: the fitted mixture must put far
    more density there than a uniform sphere (1 / 4 pi) after training."""
    rng = np.random.default_rng(5)
    g = pg.Guide(dev, seed=5, lr=1e-2, train_batch=4096)

    def target_dir(p):
        v = np.stack([np.cos(3 * p[:, 0]), np.ones(len(p)) * 0.5, np.sin(3 * p[:, 0])], 1)
        return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)

    def batch(n):
        p01, nrm, wo = _inputs(rng, n)
        d = target_dir(p01)
        d = d + 0.05 * rng.normal(size=d.shape).astype(np.float32)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        return pg.Guide.records(p01, nrm, wo, d, np.ones(n, np.float32)), p01, nrm, wo

    losses = []
    for _ in range(60):
        rec, *_ = batch(4096)
        losses.append(g.train(rec))
    assert losses[-1] < losses[0] - 2.0, (losses[0], losses[-1])
    _, p01, nrm, wo = batch(512)
    mix = g.infer(p01, nrm, wo)
    d = target_dir(p01).astype(np.float64)
    lob = mix[:, :32].reshape(-1, 8, 4).astype(np.float64)
    p = (mix[:, 32:40] * go.vmf(lob[:, :, 3], np.einsum("nlk,nk->nl", lob[:, :, :3], d))).sum(1)
    assert np.median(p) > 10.0 / (4 * np.pi)

"""CPU checks that pin the FP64 guiding oracle (oracle/guide_oracle.py) before the GPU kernels are
compared with it: the input layout covers every feature once, the encodings satisfy their closed
forms, the vMF mixture density integrates to one, and the analytic gradient of the loss agrees with
central finite differences."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import guide_oracle as go  # noqa: E402


def _unit(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def test_feature_layout_is_a_permutation():
    idx = [go.feature_index(g, t) for g in range(4) for t in range(36)]
    assert sorted(idx) == list(range(go.NIN))
    rng = np.random.default_rng(0)
    f = go.encode(rng.uniform(0, 1, (64, 3)), _unit(rng, 64), _unit(rng, 64))
    assert np.all(f[:, go.PAD_FEATURES] == 0)
    live = np.setdiff1d(np.arange(go.NIN), go.PAD_FEATURES)
    assert np.all(np.any(f[:, live] != 0, axis=0))


def test_frequency_and_blob_closed_forms():
    rng = np.random.default_rng(1)
    p = rng.uniform(0, 1, (32, 3))
    nrm, wo = _unit(rng, 32), _unit(rng, 32)
    f = go.encode(p, nrm, wo)
    # each (coordinate, frequency) item is a (sin, cos) pair on the unit circle, and doubling the
    # frequency doubles the angle: sin(2x) = 2 sin x cos x
    pairs = {}
    for g in range(4):
        for t in range(0, 18, 2):
            P = 9 * g + t // 2
            pairs[P] = (f[:, go.feature_index(g, t)], f[:, go.feature_index(g, t + 1)])
    assert len(pairs) == 36
    for P, (s, c) in pairs.items():
        assert np.allclose(s * s + c * c, 1.0)
        if P % 12 < 11:
            s2, _ = pairs[P + 1]
            assert np.allclose(s2, 2 * s * c, atol=1e-12)
        if P % 12 == 0:
            assert np.allclose(s, np.sin(np.pi * p[:, P // 12]))
    assert np.allclose(f[:, [go.feature_index(0, 34), go.feature_index(0, 35), go.feature_index(1, 34)]], nrm)
    # one-blob: a direction at the centre of theta bin b and phi bin c peaks there with value 1
    b, c = 9, 27
    th, ph = (b + 0.5) / 32 * np.pi, (c + 0.5) / 32 * 2 * np.pi
    w = np.array([[np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)]])
    f = go.encode(np.zeros((1, 3)), np.array([[0, 0, 1.0]]), w)
    tb = [f[0, go.feature_index(g, t)] for g in (0, 1) for t in range(18, 34)]
    pb = [f[0, go.feature_index(g, t)] for g in (2, 3) for t in range(18, 34)]
    assert np.argmax(tb) == b and np.argmax(pb) == c
    assert np.isclose(tb[b], 1.0) and np.isclose(pb[c], 1.0)
    assert np.isclose(tb[b + 1], np.exp(-0.5)) and np.isclose(pb[c - 2], np.exp(-2.0))


def test_vmf_mixture_integrates_to_one():
    rng = np.random.default_rng(2)
    n_t, n_p = 2000, 4000
    t = (np.arange(n_t) + 0.5) / n_t * np.pi
    ph = (np.arange(n_p) + 0.5) / n_p * 2 * np.pi
    T, Ph = np.meshgrid(t, ph, indexing="ij")
    w = np.stack([np.sin(T) * np.cos(Ph), np.sin(T) * np.sin(Ph), np.cos(T)], -1).reshape(-1, 3)
    dA = (np.pi / n_t) * (2 * np.pi / n_p) * np.sin(T).reshape(-1)
    for kappa in (0.02, 1.0, 30.0, 500.0):
        mu = _unit(rng, 1)[0]
        total = (go.vmf(np.array(kappa), w @ mu) * dA).sum()
        assert abs(total - 1.0) < 2e-3, (kappa, total)


def test_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    theta = go.init_params(seed=3)
    layers = go.unpack(theta)
    for _, b in layers:
        b[:] = rng.normal(0, 0.1, b.shape)
    layers[4][1][go.NOUT:] = 0.0
    n = 64
    feats = go.encode(rng.uniform(0, 1, (n, 3)), _unit(rng, n), _unit(rng, n))
    w = _unit(rng, n)
    c = rng.uniform(0.0, 2.0, n)
    _, grad = go.loss_and_grad(theta, feats, w, c, quant=False)
    # parameters with a live gradient: every layer's weights and biases, sampled
    live = np.flatnonzero(grad != 0)
    pick = rng.choice(live, 60, replace=False)
    h = 1e-6
    for k in pick:
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd = (go.loss(tp, feats, w, c, quant=False) - go.loss(tm, feats, w, c, quant=False)) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-4 * max(abs(fd), 1e-3), (k, fd, grad[k])


def test_learned_selection_gradient_matches_finite_differences():
    """With a learned selection probability the loss is that of alpha p_guide + (1 - alpha) p_bsdf:
    its gradient (lobes and the selection logit row alike) matches central differences, and the
    logit's gradient has the sign the closed form gives (alpha grows where the guide's density
    exceeds the BSDF's at the sampled directions)."""
    rng = np.random.default_rng(5)
    theta = go.init_params(seed=5)
    layers = go.unpack(theta)
    for _, b in layers:
        b[:] = rng.normal(0, 0.1, b.shape)
    layers[4][1][go.NOUT:] = 0.0
    W4, b4 = layers[4]
    W4[go.ALPHA_ROW] = rng.normal(0, 0.2, W4.shape[1])  # a live selection row
    b4[go.ALPHA_ROW] = 0.3
    n = 64
    feats = go.encode(rng.uniform(0, 1, (n, 3)), _unit(rng, n), _unit(rng, n))
    w = _unit(rng, n)
    c = rng.uniform(0.0, 2.0, n)
    pb = rng.uniform(0.05, 1.0, n)
    _, grad = go.loss_and_grad(theta, feats, w, c, quant=False, pb=pb)
    k4 = sum(o * i + o for o, i in go.LAYERS[:-1])
    arow = np.arange(k4 + go.ALPHA_ROW * go.NHID, k4 + (go.ALPHA_ROW + 1) * go.NHID)
    abias = k4 + go.NOUT_PAD * go.NHID + go.ALPHA_ROW
    assert np.any(grad[arow] != 0) and grad[abias] != 0
    live = np.flatnonzero(grad != 0)
    pick = np.concatenate([rng.choice(live, 50, replace=False), rng.choice(arow, 8, replace=False), [abias]])
    h = 1e-6
    for k in pick:
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd = (go.loss(tp, feats, w, c, quant=False, pb=pb) - go.loss(tm, feats, w, c, quant=False, pb=pb)) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-4 * max(abs(fd), 1e-3), (k, fd, grad[k])
    # the logit bias moves alpha toward the better density: dL/dz = -sum c (p - pb) alpha' / q / sum c
    out, _ = go.forward(theta, feats, quant=False)
    mu, kappa, pi, _ = go.mixture(out)
    pgd = (pi * go.vmf(kappa, np.einsum("nlk,nk->nl", mu, w))).sum(1)
    sg = 1 / (1 + np.exp(-out[:, go.ALPHA_ROW]))
    a = go.ALO + (go.AHI - go.ALO) * sg
    closed = (-c * (pgd - pb) * (go.AHI - go.ALO) * sg * (1 - sg) / (a * pgd + (1 - a) * pb)).sum() / c.sum()
    assert abs(grad[abias] - closed) <= 1e-9 + 1e-9 * abs(closed)
    # without BSDF densities the selection row carries no gradient (it stays a zero row)
    _, g0 = go.loss_and_grad(theta, feats, w, c, quant=False)
    assert not np.any(g0[arow]) and g0[abias] == 0


def test_adam_constant_gradient_closed_form():
    """With a constant gradient g, bias-corrected Adam moves every parameter by lr g / (|g| + eps)
    per step (m_t / (1 - b1^t) = g and v_t / (1 - b2^t) = g^2 exactly); without the correction the
    step t would be lr (1 - b1^t) / sqrt(1 - b2^t) sign(g), which equals lr at t = 1 (b1 = 0.9,
    b2 = 0.99) but 1.35 lr at t = 2 and 1.62 lr at t = 3, so five steps pin the correction."""
    rng = np.random.default_rng(3)
    g = rng.normal(size=257) * np.logspace(-3, 1, 257)
    theta = rng.normal(size=257)
    m, v = np.zeros_like(g), np.zeros_like(g)
    th = theta.copy()
    lr, eps = 1e-2, 1e-8
    for t in range(1, 6):
        th, m, v = go.adam(th, g, m, v, t, lr=lr, b1=0.9, b2=0.99, eps=eps)
        expect = theta - t * lr * g / (np.abs(g) + eps)
        assert np.allclose(th, expect, rtol=0, atol=1e-12), t
    # the uncorrected sequence differs from step 2 on
    unc = sum((1 - 0.9 ** t) / np.sqrt(1 - 0.99 ** t) for t in range(1, 6))
    assert abs(unc - 5.0) > 1.0

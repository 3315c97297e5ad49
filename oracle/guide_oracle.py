"""FP64 CPU oracle of the neural path-guiding network (test infrastructure only).

The reference has no path guiding; this restates, in plain numpy at FP64, the network the HIP
kernels implement, so that their forward pass, loss and gradients can be checked:

  inputs  : position p (scene box mapped to [0,1]^3) -> 12-frequency sin/cos encoding (72),
            shading normal n (3), outgoing direction wo as (theta, phi) one-blob, 32 bins each (64);
            139 features, zero-padded to 144.
  network : 144 -> 64 -> 64 -> 64 -> 64 (ReLU) -> 40 (linear), padded to 48 rows.
  output  : 8 von Mises-Fisher lobes, each (mu raw 3, log kappa 1, weight logit 1):
            mu = m / |m|, kappa = exp(clamp(r, KLO, KHI)), pi = softmax(logits).
  pdf     : p(w) = sum_i pi_i * kappa_i / (2 pi (1 - exp(-2 kappa_i))) * exp(kappa_i (mu_i . w - 1)).
  loss    : L = -sum_s c_s log p(w_s) / sum_s c_s   (c_s: the sample's radiance / sampling pdf).

Weights and the inputs of every layer are rounded to bfloat16 (round to nearest even) exactly as
the kernels feed the matrix cores; sums, biases and everything after the last layer are FP64.
"""
import numpy as np

NFREQ = 12
NBLOB = 32
NIN = 144
NHID = 64
NOUT = 40
NOUT_PAD = 48
NLOBES = 8
KLO, KHI = -4.0, 9.0
LAYERS = [(NHID, NIN), (NHID, NHID), (NHID, NHID), (NHID, NHID), (NOUT_PAD, NHID)]


def param_count():
    return sum(o * i + o for o, i in LAYERS)


def unpack(theta):
    """Flat parameter vector -> [(W, b)] in the kernels' layout (W row-major [out][in], then b)."""
    out, k = [], 0
    for o, i in LAYERS:
        W = theta[k:k + o * i].reshape(o, i)
        k += o * i
        b = theta[k:k + o]
        k += o
        out.append((W, b))
    return out


def bf16(x):
    """Round float64/float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def encode(pos01, normal, wo):
    """Feature vector (N, 144) for positions in [0,1]^3, unit normals and unit directions."""
    n = pos01.shape[0]
    feats = np.zeros((n, NIN))
    col = 0
    for f in range(NFREQ):
        s = (2.0 ** f) * np.pi * pos01
        feats[:, col:col + 3] = np.sin(s)
        feats[:, col + 3:col + 6] = np.cos(s)
        col += 6
    feats[:, col:col + 3] = normal
    col += 3
    theta = np.arccos(np.clip(wo[:, 2], -1.0, 1.0)) / np.pi
    phi = (np.arctan2(wo[:, 1], wo[:, 0]) / (2 * np.pi)) % 1.0
    centers = (np.arange(NBLOB) + 0.5) / NBLOB
    for v in (theta, phi):
        d = v[:, None] - centers[None, :]
        feats[:, col:col + NBLOB] = np.exp(-0.5 * (d * NBLOB) ** 2)
        col += NBLOB
    assert col == 139
    return feats


def forward(theta, feats, quant=True):
    """Returns (outputs (N, 48), cache of per-layer inputs and pre-activations).
    quant=False skips the bfloat16 rounding (used to pin the gradients by finite differences)."""
    q = bf16 if quant else (lambda a: np.asarray(a, np.float64))
    layers = unpack(theta)
    x = q(feats)
    cache = []
    for li, (W, b) in enumerate(layers):
        z = x @ q(W).T + b
        cache.append((x, z))
        if li < len(layers) - 1:
            x = q(np.maximum(z, 0.0))
    return z, cache


def mixture(out):
    """Lobe parameters from raw outputs: mu (N,8,3), kappa (N,8), pi (N,8), and intermediates."""
    o = out[:, :NOUT].reshape(-1, NLOBES, 5)
    m = o[:, :, 0:3]
    mnorm = np.maximum(np.linalg.norm(m, axis=2), 1e-6)
    mu = m / mnorm[:, :, None]
    r = o[:, :, 3]
    kappa = np.exp(np.clip(r, KLO, KHI))
    a = o[:, :, 4]
    e = np.exp(a - a.max(axis=1, keepdims=True))
    pi = e / e.sum(axis=1, keepdims=True)
    return mu, kappa, pi, (m, mnorm, r)


def vmf(kappa, cosang):
    return kappa / (2.0 * np.pi * (-np.expm1(-2.0 * kappa))) * np.exp(kappa * (cosang - 1.0))


def pdf(theta, feats, w, quant=True):
    out, _ = forward(theta, feats, quant)
    mu, kappa, pi, _ = mixture(out)
    c = np.einsum("nlk,nk->nl", mu, w)
    return (pi * vmf(kappa, c)).sum(axis=1)


def loss(theta, feats, w, c, quant=True):
    p = pdf(theta, feats, w, quant)
    return float(-(c * np.log(p)).sum() / c.sum())


def loss_and_grad(theta, feats, w, c, quant=True):
    """Analytic gradient of loss() with respect to theta (bf16 rounding treated as identity)."""
    q = bf16 if quant else (lambda a: np.asarray(a, np.float64))
    layers = unpack(theta)
    out, cache = forward(theta, feats, quant)
    mu, kappa, pi, (m, mnorm, r) = mixture(out)
    cosang = np.einsum("nlk,nk->nl", mu, w)
    v = vmf(kappa, cosang)
    p = (pi * v).sum(axis=1)
    csum = c.sum()
    L = float(-(c * np.log(p)).sum() / csum)
    dLdp = -(c / csum) / p                            # (N,)
    g = dLdp[:, None] * v                             # dL/dpi_i
    da = pi * (g - (pi * g).sum(axis=1, keepdims=True))
    dv = dLdp[:, None] * pi                           # dL/dv_i
    # dv/dkappa = v (1/kappa - 2/(exp(2 kappa) - 1) + cos - 1); kappa = exp(r) unless clamped
    dlogC = 1.0 / kappa - 2.0 / np.expm1(2.0 * kappa)
    dk = dv * v * (dlogC + cosang - 1.0)
    dr = dk * kappa * ((r > KLO) & (r < KHI))
    # dv/dmu = v kappa w ; mu = m/|m| -> dm = (dmu - mu (mu . dmu)) / |m|
    dmu = (dv * v * kappa)[:, :, None] * w[:, None, :]
    dm = (dmu - mu * np.einsum("nlk,nlk->nl", mu, dmu)[:, :, None]) / mnorm[:, :, None]
    dout = np.zeros_like(out)
    o = dout[:, :NOUT].reshape(-1, NLOBES, 5)
    o[:, :, 0:3] = dm
    o[:, :, 3] = dr
    o[:, :, 4] = da
    dout[:, :NOUT] = o.reshape(-1, NOUT)
    grads = []
    dz = dout
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        x, z = cache[li]
        if li < len(layers) - 1:
            dz = dz * (z > 0.0)
        grads.append((dz.T @ x, dz.sum(axis=0)))
        dz = dz @ q(W)
    grads.reverse()
    flat = np.concatenate([np.concatenate([gW.reshape(-1), gb]) for gW, gb in grads])
    # padding rows of the output layer (40..47) are never trained
    k = sum(o * i + o for o, i in LAYERS[:-1])
    Wlast = flat[k:k + NOUT_PAD * NHID].reshape(NOUT_PAD, NHID)
    Wlast[NOUT:] = 0.0
    flat[k + NOUT_PAD * NHID + NOUT:] = 0.0
    return L, flat


def init_params(seed=0):
    """Deterministic init used by both sides: uniform(+-sqrt(6/(in+out))) weights, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for li, (o, i) in enumerate(LAYERS):
        lim = np.sqrt(6.0 / (i + o))
        W = rng.uniform(-lim, lim, (o, i))
        if li == 0:
            W[:, 139:] = 0.0
        if li == len(LAYERS) - 1:
            W[NOUT:] = 0.0
        parts += [W.reshape(-1), np.zeros(o)]
    return np.concatenate(parts)


def adam(theta, g, m, v, t, lr=1e-2, b1=0.9, b2=0.99, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v

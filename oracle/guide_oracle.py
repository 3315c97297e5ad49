"""FP64 CPU oracle of the neural path-guiding network (test infrastructure only).

The reference has no path guiding; this restates, in plain numpy at FP64, the network the HIP
kernels implement, so that their forward pass, loss and gradients can be checked:

  inputs  : position p (scene box mapped to [0,1]^3) -> 12-frequency sin/cos encoding (72),
            shading normal n (3), outgoing direction wo as (theta, phi) one-blob, 32 bins each (64);
            139 features and 5 zeros, 144 in all, in the order encode() gives (each lane group of
            the kernels computes exactly the features it feeds to the matrix cores).
  network : 144 -> 64 -> 64 -> 64 -> 64 (ReLU) -> 40 (linear) in 48 rows (output_row() gives
            where each lobe parameter sits; PAD_OUTPUTS are zero rows).
  output  : 8 von Mises-Fisher lobes, each (mu raw 3, log kappa 1, weight logit 1):
            mu = m / |m|, kappa = exp(clamp(r, KLO, KHI)), pi = softmax(logits).
  pdf     : p(w) = sum_i pi_i * kappa_i / (2 pi (1 - exp(-2 kappa_i))) * exp(kappa_i (mu_i . w - 1)).
  loss    : L = -sum_s c_s log p(w_s) / sum_s c_s   (c_s: the sample's radiance / sampling pdf).
  learned selection (optional): output row ALPHA_ROW is the logit z of the probability of sampling
            the guide, alpha = ALO + (AHI - ALO) sigmoid(z); the loss is then that of the one-sample
            MIS density q = alpha p + (1 - alpha) p_bsdf, with p_bsdf given per sample.

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
ALO, AHI = 0.05, 0.95
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


def feature_index(g, t):
    """Input feature fed by slot t (0..35) of lane group g (0..3): the MFMA B-operand position
    16 kt + 4 g + j with t = 4 kt + j."""
    return (t // 4) * 16 + 4 * g + t % 4


# Input features that are always zero (their first-layer weights start at zero).
PAD_FEATURES = sorted([feature_index(1, 35), feature_index(2, 34), feature_index(2, 35),
                       feature_index(3, 34), feature_index(3, 35)])


def output_row(l, k):
    """Output row of lobe l (0..7), parameter k (0..4: mu x, mu y, mu z, log kappa, weight logit):
    lane group g = l // 2 holds rows 16 mt + 4 g + j, and its slot u = 4 mt + j = 5 (l % 2) + k."""
    g, u = l // 2, 5 * (l % 2) + k
    return 16 * (u // 4) + 4 * g + u % 4


LOBE_ROWS = np.array([[output_row(l, k) for k in range(5)] for l in range(NLOBES)])  # (8, 5)
# Output rows outside the lobes (slots 10 and 11 of every lane group): zero, except slot 10 of lane
# group 0, the selection logit, when a guide learns its selection probability.
PAD_OUTPUTS = sorted(set(range(NOUT_PAD)) - set(LOBE_ROWS.reshape(-1).tolist()))
ALPHA_ROW = 16 * (10 // 4) + 10 % 4  # lane group 0, slot 10
assert ALPHA_ROW in PAD_OUTPUTS


def encode(pos01, normal, wo):
    """Feature vector (N, 144) for positions in [0,1]^3, unit normals and unit directions.

    Slot t of lane group g (feature feature_index(g, t)) holds
      t = 0..17 : position item P = 9 g + t // 2 (coordinate P // 12, frequency f = P % 12):
                  sin(2^f pi p) for even t, cos(2^f pi p) for odd t;
      t = 18..33: one-blob bin b = 16 (g % 2) + t - 18 of theta = atan2(|wo.xy|, wo.z) / pi
                  (g < 2) or phi = atan2(wo.y, wo.x) / 2 pi mod 1 (g >= 2): exp(-(32 a - b - 0.5)^2 / 2);
      t = 34, 35: n.x, n.y (g = 0); n.z, 0 (g = 1); 0, 0 (g = 2, 3).
    """
    n = pos01.shape[0]
    feats = np.zeros((n, NIN))
    theta = np.arctan2(np.hypot(wo[:, 0], wo[:, 1]), wo[:, 2]) / np.pi
    phi = (np.arctan2(wo[:, 1], wo[:, 0]) / (2 * np.pi)) % 1.0
    for g in range(4):
        for t in range(18):
            P = 9 * g + t // 2
            x = (2.0 ** (P % 12)) * np.pi * pos01[:, P // 12]
            feats[:, feature_index(g, t)] = np.sin(x) if t % 2 == 0 else np.cos(x)
        a = theta if g < 2 else phi
        for t in range(18, 34):
            b = 16 * (g % 2) + t - 18
            feats[:, feature_index(g, t)] = np.exp(-0.5 * (a * NBLOB - b - 0.5) ** 2)
    feats[:, feature_index(0, 34)] = normal[:, 0]
    feats[:, feature_index(0, 35)] = normal[:, 1]
    feats[:, feature_index(1, 34)] = normal[:, 2]
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
    o = out[:, LOBE_ROWS]                                # (N, 8, 5)
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


def selection(out):
    """Learned probability of sampling the guide, from the selection logit row."""
    return ALO + (AHI - ALO) / (1.0 + np.exp(-out[:, ALPHA_ROW]))


def loss(theta, feats, w, c, quant=True, pb=None):
    """pb: the BSDF's density of each w (learned selection: the loss of the combined density)."""
    if pb is None:
        p = pdf(theta, feats, w, quant)
    else:
        out, _ = forward(theta, feats, quant)
        mu, kappa, pi, _ = mixture(out)
        pg = (pi * vmf(kappa, np.einsum("nlk,nk->nl", mu, w))).sum(axis=1)
        a = selection(out)
        p = a * pg + (1.0 - a) * pb
    return float(-(c * np.log(p)).sum() / c.sum())


def loss_and_grad(theta, feats, w, c, quant=True, pb=None):
    """Analytic gradient of loss() with respect to theta (bf16 rounding treated as identity)."""
    q = bf16 if quant else (lambda a: np.asarray(a, np.float64))
    layers = unpack(theta)
    out, cache = forward(theta, feats, quant)
    mu, kappa, pi, (m, mnorm, r) = mixture(out)
    cosang = np.einsum("nlk,nk->nl", mu, w)
    v = vmf(kappa, cosang)
    p = (pi * v).sum(axis=1)
    csum = c.sum()
    if pb is None:
        L = float(-(c * np.log(p)).sum() / csum)
        dLdp = -(c / csum) / p                        # (N,)
    else:  # q = alpha p + (1 - alpha) pb, alpha = ALO + (AHI - ALO) sigmoid(z)
        sg = 1.0 / (1.0 + np.exp(-out[:, ALPHA_ROW]))
        a = ALO + (AHI - ALO) * sg
        qd = a * p + (1.0 - a) * pb
        L = float(-(c * np.log(qd)).sum() / csum)
        dLdq = -(c / csum) / qd
        dLdp = dLdq * a
        dz_alpha = dLdq * (p - pb) * (AHI - ALO) * sg * (1.0 - sg)
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
    dout[:, LOBE_ROWS[:, 0:3]] = dm
    dout[:, LOBE_ROWS[:, 3]] = dr
    dout[:, LOBE_ROWS[:, 4]] = da
    if pb is not None:
        dout[:, ALPHA_ROW] = dz_alpha
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
    # the zero rows of the output layer are never trained
    k = sum(o * i + o for o, i in LAYERS[:-1])
    zero = [o for o in PAD_OUTPUTS if pb is None or o != ALPHA_ROW]
    Wlast = flat[k:k + NOUT_PAD * NHID].reshape(NOUT_PAD, NHID)
    Wlast[zero] = 0.0
    flat[k + NOUT_PAD * NHID + np.array(zero)] = 0.0
    return L, flat


def init_params(seed=0):
    """Deterministic init used by both sides: uniform(+-sqrt(6/(in+out))) weights, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for li, (o, i) in enumerate(LAYERS):
        lim = np.sqrt(6.0 / (i + o))
        W = rng.uniform(-lim, lim, (o, i))
        if li == 0:
            W[:, PAD_FEATURES] = 0.0
        if li == len(LAYERS) - 1:
            W[PAD_OUTPUTS] = 0.0
        parts += [W.reshape(-1), np.zeros(o)]
    return np.concatenate(parts)


def adam(theta, g, m, v, t, lr=1e-2, b1=0.9, b2=0.99, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v

"""CPU restatement of the learned path guide (numpy). TEST INFRASTRUCTURE ONLY.

The reference has no path guiding; this module defines what the HIP guide must
compute, from published methods:
  - von Mises-Fisher lobe density and its stable sampling (W. Jakob, "Numerically
    stable sampling of the von Mises Fisher distribution on S^2", 2012);
  - KL-divergence training of a sampling density on Monte Carlo estimates of
    the target (Mueller et al., "Neural Importance Sampling", 2019);
  - Adam (Kingma, Ba 2015).
Layout (shared with the device through include/spt.h): 32 inputs -> 64 -> 64 ->
64 (ReLU) -> 40 outputs, 8 lobes x (mu_x, mu_y, mu_z, log kappa, logit); the
parameter vector is [W1 64x32 | b1 64 | W2 64x64 | b2 | W3 64x64 | b3 |
W4 48x64 (rows 40..47 zero) | b4 48], 13,552 floats.
"""
from __future__ import annotations

import numpy as np

ENC, HID, OUT, OUTP, LOBES = 32, 64, 40, 48, 8
SHAPES = [("W1", (HID, ENC)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (HID, HID)),
          ("b3", (HID,)), ("W4", (OUTP, HID)), ("b4", (OUTP,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)
KAPPA_MIN, KAPPA_MAX = 1e-2, 1e4


def unpack(theta):
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = theta[o:o + n].reshape(shape)
        o += n
    return out


def pack(p):
    return np.concatenate([p[name].reshape(-1) for name, _ in SHAPES])


def bf16(x):
    """Round float32 to the nearest bfloat16 (ties to even), returned as float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def encode(p, n, wo, bmin, bext):
    """(N,3) position, normal, outgoing direction -> (N,32) features:
    x in [0,1]^3 | sin, cos(2 pi 2^k x_c), k = 0..2, c = x,y,z | n | wo |
    one-blob(4) of (n.wo + 1)/2 | 1."""
    x = np.clip((p - bmin) / bext, 0.0, 1.0)
    f = [x]
    for k in range(3):
        a = 2.0 * np.pi * (2.0 ** k) * x
        f.append(np.stack([np.sin(a[:, 0]), np.cos(a[:, 0]), np.sin(a[:, 1]), np.cos(a[:, 1]),
                           np.sin(a[:, 2]), np.cos(a[:, 2])], 1))
    f += [n, wo]
    c = 0.5 * (np.sum(n * wo, 1) + 1.0)
    centers = (np.arange(4) + 0.5) / 4.0
    f.append(np.exp(-((c[:, None] - centers[None, :]) ** 2) / (2.0 * 0.25 ** 2)))
    f.append(np.ones((len(p), 1)))
    return np.concatenate(f, 1)


def forward(theta, X, emulate_bf16=True):
    """Returns the 48 raw outputs (last 8 zero) and the hidden activations.
    With emulate_bf16, inputs, weights and hidden activations are rounded to
    bfloat16 as the MFMA kernels consume them; sums are float64."""
    p = unpack(np.asarray(theta, np.float64))
    r = (lambda a: bf16(a).astype(np.float64)) if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    h0 = r(X)
    a1 = h0 @ r(p["W1"]).T + p["b1"]
    h1 = r(np.maximum(a1, 0))
    a2 = h1 @ r(p["W2"]).T + p["b2"]
    h2 = r(np.maximum(a2, 0))
    a3 = h2 @ r(p["W3"]).T + p["b3"]
    h3 = r(np.maximum(a3, 0))
    y = h3 @ r(p["W4"]).T + p["b4"]
    return y, (h0, a1, h1, a2, h2, a3, h3)


def mixture(y):
    """raw outputs (N,>=40) -> mu (N,8,3) unit, kappa (N,8), weights (N,8), and
    the clamped log kappa mask."""
    z = y[:, :OUT].reshape(-1, LOBES, 5)
    raw = z[:, :, 0:3]
    nr = np.sqrt(np.sum(raw * raw, -1) + 1e-12)
    mu = raw / nr[..., None]
    s = z[:, :, 3]
    sc = np.clip(s, np.log(KAPPA_MIN), np.log(KAPPA_MAX))
    kappa = np.exp(sc)
    a = z[:, :, 4]
    a = a - a.max(1, keepdims=True)
    w = np.exp(a)
    w /= w.sum(1, keepdims=True)
    return mu, kappa, w, (s == sc), nr


def log_vmf(mu, kappa, d):
    """log of kappa / (2 pi (1 - e^{-2 kappa})) * exp(kappa (mu.d - 1))."""
    c = np.sum(mu * d[:, None, :], -1)
    return np.log(kappa) - np.log(2 * np.pi) - np.log(-np.expm1(-2 * kappa)) + kappa * (c - 1.0)


def mixture_pdf(y, d):
    mu, kappa, w, _, _ = mixture(y)
    return np.sum(w * np.exp(log_vmf(mu, kappa, d)), 1)


def sample_vmf(mu, kappa, u1, u2):
    """Jakob 2012: w = 1 + log(u1 + (1 - u1) e^{-2 kappa}) / kappa, clamped to
    its limit -1 where the argument underflows (u1 = 0, large kappa)."""
    with np.errstate(divide="ignore"):
        wz = np.maximum(1.0 + np.log(u1 + (1.0 - u1) * np.exp(-2.0 * kappa)) / kappa, -1.0)
    st = np.sqrt(np.maximum(0.0, 1.0 - wz * wz))
    ph = 2 * np.pi * u2
    local = np.stack([st * np.cos(ph), st * np.sin(ph), wz], -1)
    t, b = frame(mu)
    return local[..., 0:1] * t + local[..., 1:2] * b + local[..., 2:3] * mu


def frame(n):
    """Orthonormal t, b with t x b = n (Duff et al. 2017, 'Building an
    orthonormal basis, revisited')."""
    s = np.where(n[..., 2] >= 0, 1.0, -1.0)
    a = -1.0 / (s + n[..., 2])
    b = n[..., 0] * n[..., 1] * a
    t = np.stack([1.0 + s * n[..., 0] ** 2 * a, s * b, -s * n[..., 0]], -1)
    bb = np.stack([b, s + n[..., 1] ** 2 * a, -n[..., 1]], -1)
    return t, bb


def loss_grad(y, d, W):
    """Per-sample loss -W log q(d) and its gradient w.r.t. the 48 raw outputs."""
    mu, kappa, w, inside, nr = mixture(y)
    lv = log_vmf(mu, kappa, d)
    lw = np.log(w)
    m = (lw + lv).max(1, keepdims=True)
    logq = (m + np.log(np.sum(np.exp(lw + lv - m), 1, keepdims=True)))[:, 0]
    r = np.exp(lw + lv - logq[:, None])  # posterior lobe responsibility
    g = np.zeros((len(y), OUTP))
    gz = g[:, :OUT].reshape(-1, LOBES, 5)
    c = np.sum(mu * d[:, None, :], -1)
    gz[:, :, 4] = -W[:, None] * (r - w)
    dlv_ds = 1.0 - 2.0 * kappa / np.expm1(2.0 * kappa) + kappa * (c - 1.0)
    gz[:, :, 3] = np.where(inside, -W[:, None] * r * dlv_ds, 0.0)
    gmu = -(W[:, None] * r * kappa)[..., None] * d[:, None, :]
    proj = gmu - np.sum(gmu * mu, -1, keepdims=True) * mu
    gz[:, :, 0:3] = proj / nr[..., None]
    g[:, :OUT] = gz.reshape(-1, OUT)
    return -W * logq, g


def backward(theta, X, dy, emulate_bf16=True):
    """Gradients of sum_n dy[n] . y[n] w.r.t. all parameters (packed)."""
    p = unpack(np.asarray(theta, np.float64))
    r = (lambda a: bf16(a).astype(np.float64)) if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    _, (h0, a1, h1, a2, h2, a3, h3) = forward(theta, X, emulate_bf16)
    g = {}
    g["W4"] = dy.T @ h3
    g["b4"] = dy.sum(0)
    dh3 = (dy @ r(p["W4"])) * (a3 > 0)
    g["W3"] = dh3.T @ h2
    g["b3"] = dh3.sum(0)
    dh2 = (dh3 @ r(p["W3"])) * (a2 > 0)
    g["W2"] = dh2.T @ h1
    g["b2"] = dh2.sum(0)
    dh1 = (dh2 @ r(p["W2"])) * (a1 > 0)
    g["W1"] = dh1.T @ h0
    g["b1"] = dh1.sum(0)
    out = pack(g)
    out[_pad_mask()] = 0.0
    return out


def _pad_mask():
    m = np.zeros(N_PARAMS, bool)
    p = unpack(np.arange(N_PARAMS))
    m[p["W4"][OUT:].reshape(-1)] = True
    m[p["b4"][OUT:]] = True
    return m


def batch_grad(theta, X, d, W, loss_scale=1.0, emulate_bf16=True):
    """The training step's gradient: weights normalised by their batch mean,
    loss averaged over the batch."""
    y, _ = forward(theta, X, emulate_bf16)
    Wn = W / max(W.mean(), 1e-30)
    l, dy = loss_grad(y, d, Wn)
    dy = dy / len(X)
    return l.mean(), backward(theta, X, dy * loss_scale, emulate_bf16) / loss_scale


def adam(theta, m, v, g, t, lr=5e-3, b1=0.9, b2=0.99, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(seed):
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases; the
    output layer is scaled down so the initial mixture is broad."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in SHAPES:
        if name.startswith("W"):
            lim = 1.0 / np.sqrt(shape[1])
            p[name] = rng.uniform(-lim, lim, shape)
        else:
            p[name] = np.zeros(shape)
    p["W4"] *= 0.1
    p["W4"][OUT:] = 0.0
    return pack(p).astype(np.float32)

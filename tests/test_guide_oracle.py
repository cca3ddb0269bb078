"""Pins of the guide's CPU oracle (oracle/guide_ref.py): the analytic loss
gradient against torch autograd (float64, CPU) and central finite
differences; vMF lobe and mixture densities integrate to 1; the sampler
follows the density; Adam matches its closed form on a 1-D quadratic."""
import math

import numpy as np
import pytest

import guide_ref as G


def _batch(n=64, seed=0):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-1, 1, (n, 3))
    nn = rng.normal(size=(n, 3)); nn /= np.linalg.norm(nn, axis=1, keepdims=True)
    wo = rng.normal(size=(n, 3)); wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    d = rng.normal(size=(n, 3)); d /= np.linalg.norm(d, axis=1, keepdims=True)
    W = rng.uniform(0.1, 2.0, n)
    X = G.encode(p, nn, wo, np.array([-1.0] * 3), np.array([2.0] * 3))
    return X, d, W


def test_parameter_layout():
    assert G.N_PARAMS == 13552 and G.encode(*[np.zeros((1, 3))] * 3, np.zeros(3), np.ones(3)).shape == (1, 32)


def test_loss_gradient_matches_torch_autograd():
    torch = pytest.importorskip("torch")
    X, d, W = _batch()
    theta = G.init_params(3).astype(np.float64)
    theta += np.random.default_rng(1).normal(0, 0.05, theta.shape) * (~G._pad_mask())
    loss, g = G.batch_grad(theta, X, d, W, emulate_bf16=False)

    t = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    ps, o = {}, 0
    for name, shape in G.SHAPES:
        n = int(np.prod(shape))
        ps[name] = t[o:o + n].reshape(shape)
        o += n
    x = torch.tensor(X, dtype=torch.float64)
    h = torch.relu(x @ ps["W1"].T + ps["b1"])
    h = torch.relu(h @ ps["W2"].T + ps["b2"])
    h = torch.relu(h @ ps["W3"].T + ps["b3"])
    y = (h @ ps["W4"].T + ps["b4"])[:, :40].reshape(-1, 8, 5)
    mu = y[..., 0:3] / torch.sqrt((y[..., 0:3] ** 2).sum(-1, keepdim=True) + 1e-12)
    kappa = torch.exp(torch.clamp(y[..., 3], math.log(G.KAPPA_MIN), math.log(G.KAPPA_MAX)))
    logw = torch.log_softmax(y[..., 4], 1)
    dd = torch.tensor(d, dtype=torch.float64)
    c = (mu * dd[:, None, :]).sum(-1)
    lv = torch.log(kappa) - math.log(2 * math.pi) - torch.log(-torch.expm1(-2 * kappa)) + kappa * (c - 1)
    logq = torch.logsumexp(logw + lv, 1)
    Wt = torch.tensor(W / W.mean(), dtype=torch.float64)
    (-(Wt * logq).mean()).backward()
    assert abs(float(-(Wt * logq).mean().detach()) - loss) < 1e-10
    np.testing.assert_allclose(g, t.grad.numpy(), rtol=1e-7, atol=1e-10)


def test_loss_gradient_matches_finite_differences():
    X, d, W = _batch(16, 2)
    theta = G.init_params(5).astype(np.float64)
    _, g = G.batch_grad(theta, X, d, W, emulate_bf16=False)
    rng = np.random.default_rng(0)
    idx = rng.choice(np.nonzero(~G._pad_mask())[0], 40, replace=False)
    for i in idx:
        e = np.zeros_like(theta); e[i] = 1e-6
        lp, _ = G.batch_grad(theta + e, X, d, W, emulate_bf16=False)
        lm, _ = G.batch_grad(theta - e, X, d, W, emulate_bf16=False)
        fd = (lp - lm) / 2e-6
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])


@pytest.mark.parametrize("kappa", [0.05, 1.0, 30.0, 800.0])
def test_vmf_lobe_normalised_and_sampled(kappa):
    mu = np.array([[0.3, -0.5, 0.81]]); mu /= np.linalg.norm(mu)
    # density integrates to 1 (quadrature in the lobe frame)
    n_t = 4000
    th = (np.arange(n_t) + 0.5) / n_t * np.pi
    c = np.cos(th)
    lv = math.log(kappa) - math.log(2 * math.pi) - math.log(-math.expm1(-2 * kappa)) + kappa * (c - 1)
    assert abs(np.sum(np.exp(lv) * np.sin(th)) * (np.pi / n_t) * 2 * np.pi - 1.0) < 1e-3
    # sampled cosines follow E[mu.w] = coth(kappa) - 1/kappa
    rng = np.random.default_rng(0)
    u1, u2 = rng.random(200000), rng.random(200000)
    w = G.sample_vmf(np.repeat(mu, len(u1), 0), np.full(len(u1), kappa), u1, u2)
    assert np.allclose(np.linalg.norm(w, axis=1), 1.0, atol=1e-9)
    mean_c = np.mean(w @ mu[0])
    expect = 1.0 / math.tanh(kappa) - 1.0 / kappa
    assert abs(mean_c - expect) < 3e-3


def test_mixture_pdf_integrates_to_one():
    X, _, _ = _batch(4, 3)
    theta = G.init_params(7).astype(np.float64) * 20.0  # sharp lobes
    y, _ = G.forward(theta, X, emulate_bf16=False)
    n_t, n_p = 600, 1200
    th = (np.arange(n_t) + 0.5) / n_t * np.pi
    ph = (np.arange(n_p) + 0.5) / n_p * 2 * np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    dirs = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dA = (np.sin(T) * (np.pi / n_t) * (2 * np.pi / n_p)).reshape(-1)
    for i in range(len(y)):
        q = G.mixture_pdf(np.repeat(y[i:i + 1], len(dirs), 0), dirs)
        assert abs(np.sum(q * dA) - 1.0) < 5e-3


def test_adam_closed_form():
    theta, m, v = np.array([1.0]), np.zeros(1), np.zeros(1)
    g = 2 * theta
    th1, m1, v1 = G.adam(theta, m, v, g, 1, lr=0.1)
    # first step moves by lr * sign(g) (bias-corrected m/sqrt(v) = g/|g|)
    assert abs(th1[0] - 0.9) < 1e-7
    for t in range(2, 400):
        th1, m1, v1 = G.adam(th1, m1, v1, 2 * th1, t, lr=0.05)
    assert abs(th1[0]) < 0.05


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159265, 65504.0], np.float32)
    r = G.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == np.float32(1.0078125)
    assert abs(r[3] + 3.140625) < 1e-6


def test_vmf_sample_limit_at_underflow():
    """For large kappa e^{-2 kappa} underflows; u1 = 0 must give the limit
    direction -mu (w = -1), not a NaN."""
    mu = np.array([[0.0, 0.0, 1.0]])
    for kappa in (50.0, 1e3, 1e4):
        d = G.sample_vmf(mu, np.array([kappa]), np.array([0.0]), np.array([0.3]))
        assert np.all(np.isfinite(d)) and np.allclose(d, [[0.0, 0.0, -1.0]], atol=1e-12)

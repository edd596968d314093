"""Pins the FP64 guiding oracle (oracle/guide.py): analytic gradients against
central finite differences, the vMF normalisation and the sampler's density
against closed forms."""
import math

import numpy as np

from oracle import guide as G


def _batch(rng, B=64):
    p = rng.uniform(-1, 1, (B, 3))
    n = rng.normal(size=(B, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    w = rng.normal(size=(B, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    c = rng.uniform(0.1, 2.0, B)
    return G.features(p, n, [-1, -1, -1], [1, 1, 1]), w, c


def test_feature_layout():
    x = G.features(np.array([[0.5, 0.25, 0.0]]), np.array([[0, 0, 1.0]]), [0, 0, 0], [1, 1, 1])
    assert x.shape == (1, 78) and G.N_IN == 78 and G.n_params() == 15976
    assert np.allclose(x[0, :3], [0.5, 0.25, 0.0])
    assert np.isclose(x[0, 3], math.sin(math.pi * 0.5)) and np.isclose(x[0, 4], math.cos(math.pi * 0.5))
    assert np.allclose(x[0, -3:], [0, 0, 1])


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    theta = G.init_params(1) * 3.0
    x, w, c = _batch(rng, 32)
    loss, grad, _ = G.loss_and_grad(theta, x, w, c)
    idx = rng.choice(len(theta), 300, replace=False)
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (G.loss_and_grad(tp, x, w, c)[0] - G.loss_and_grad(tm, x, w, c)[0]) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-5 * max(1.0, abs(fd)) + 1e-7, (i, fd, grad[i])


def test_vmf_normalisation_and_limits():
    rng = np.random.default_rng(2)
    u = rng.uniform(size=(200000, 2))
    z = 1 - 2 * u[:, 0]
    r = np.sqrt(1 - z * z)
    w = np.stack([r * np.cos(2 * math.pi * u[:, 1]), r * np.sin(2 * math.pi * u[:, 1]), z], 1)
    for k in (0.01, 1.0, 10.0, 200.0):
        mu = np.tile(np.array([[[0, 0, 1.0]]]), (len(w), 1, 1))
        pdf = G.vmf_pdf(w, mu, np.full((len(w), 1), k))[:, 0]
        est = 4 * math.pi * pdf.mean()  # integral over the sphere
        assert abs(est - 1) < (0.02 if k < 100 else 0.2), (k, est)
    # kappa -> 0: uniform density 1/(4 pi)
    assert np.isclose(G.vmf_pdf(w[:1], mu[:1], np.full((1, 1), 1e-9))[0, 0], 1 / (4 * math.pi), rtol=1e-6)


def test_sampler_follows_the_mixture_pdf():
    """Histogram of sampled cos(theta) around lobe 0 against the pdf."""
    rng = np.random.default_rng(3)
    out = np.zeros((1, G.N_OUT))
    for j in range(G.N_LOBES):
        m = rng.normal(size=3)
        out[0, 5 * j] = rng.normal()
        out[0, 5 * j + 1:5 * j + 4] = m / np.linalg.norm(m) * rng.uniform(0.5, 2)
        out[0, 5 * j + 4] = math.log(math.expm1(rng.uniform(0.5, 20) - G.KAPPA_MIN))
    N = 400000
    o = np.repeat(out, N, 0)
    d = G.sample(o, rng.uniform(size=N), rng.uniform(size=N), rng.uniform(size=N))
    assert np.allclose(np.linalg.norm(d, axis=1), 1)
    # E[pdf(d)] under sampling = integral pdf^2; compare with uniform-sphere estimate
    pdf_s = G.pdf(o[:20000], d[:20000])
    u = rng.uniform(size=(N, 2))
    z = 1 - 2 * u[:, 0]
    r = np.sqrt(1 - z * z)
    wu = np.stack([r * np.cos(2 * math.pi * u[:, 1]), r * np.sin(2 * math.pi * u[:, 1]), z], 1)
    pu = G.pdf(o, wu)
    assert abs(pdf_s.mean() - 4 * math.pi * np.mean(pu * pu)) < 0.03 * pdf_s.mean()


def test_adam_reduces_the_loss():
    rng = np.random.default_rng(4)
    theta = G.init_params(5)
    x, _, c = _batch(rng, 256)
    target = np.array([0.0, 0.6, 0.8])
    w = np.tile(target, (256, 1)) + 0.05 * rng.normal(size=(256, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    m = np.zeros_like(theta)
    v = np.zeros_like(theta)
    l0 = G.loss_and_grad(theta, x, w, c)[0]
    for t in range(1, 301):
        loss, g, _ = G.loss_and_grad(theta, x, w, c)
        theta, m, v = G.adam(theta, g, m, v, t, lr=3e-3)
    assert loss < l0 - 1.0


def test_bf16_rounding():
    """Round to nearest even at 8 significant bits, through FP32."""
    one = 1.0
    cases = {one: one, one + 2 ** -8: one, one + 3 * 2 ** -8: one + 2 ** -6,
             one + 2 ** -8 + 2 ** -20: one + 2 ** -7, -2.5: -2.5, 0.0: 0.0,
             3.0e38: float(np.array([0x7F620000], np.uint32).view(np.float32)[0])}  # 0x7F61B1E6 rounds up
    for x, want in cases.items():
        assert G.bf16(np.array([x]))[0] == np.float64(np.float32(want)), x
    # the quantised model is the FP64 model on BF16-representable inputs
    th = G.bf16(G.init_params(1))
    x = G.bf16(np.random.default_rng(0).normal(size=(16, G.N_IN)))
    a, _ = G.forward(th, x)
    b, _ = G.forward(th, x, quant=True)
    np.testing.assert_allclose(a[:, :], b[:, :], rtol=0, atol=0.02)
    h0, acts = G.forward(th, x, quant=True)
    assert np.array_equal(acts[0], x)

"""FP64 restatement of the learned path-guiding model (no reference
counterpart: turner's pathtracer samples the hemisphere uniformly,
lib/sampling.h:20-32).

TEST INFRASTRUCTURE ONLY: the checker for the BF16-MFMA kernels of
turner_amd/csrc/guide.hip.  Its analytic gradients are pinned against
central finite differences (tests/test_guide_oracle.py).

Model
-----
features(p, n) = [u, sin(2^k pi u_c), cos(2^k pi u_c) (c = x,y,z; k = 0..11), n]
                 with u = (p - bmin) / (bmax - bmin)                 -> 78
MLP 78 -> 64 -> 64 -> 64 -> 40, ReLU on the hidden layers
out[5j .. 5j+4] = (logit_j, m_j (3), kraw_j) for lobes j = 0..7
pi = softmax(logit), mu_j = m_j / |m_j|, kappa_j = softplus(kraw_j) + KAPPA_MIN
q(w) = sum_j pi_j vMF(w; mu_j, kappa_j),
vMF(w; mu, k) = k / (2 pi (1 - exp(-2k))) * exp(k (mu . w - 1))
Loss of a batch of (features, w_i, weight c_i): mean_i( c_i / mean(c) * -log q(w_i) )
"""
import math

import numpy as np

N_FREQ = 12
N_IN = 3 + 3 * 2 * N_FREQ + 3  # 78
HIDDEN = 64
N_LOBES = 8
N_OUT = 5 * N_LOBES  # 40
KAPPA_MIN = 0.01
KAPPA_MAX = 1.0e4
LAYERS = [(N_IN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, N_OUT)]


def n_params():
    return sum(o * i + o for i, o in LAYERS)


def bf16(a):
    """Round to BF16 (nearest even) through FP32, as the device does, and
    return the values as float64."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32).astype(np.float64)


def unit_coords_f32(p, bmin, bmax):
    """u = (p - bmin) * (1 / (bmax - bmin)) in FP32 operations, as the device
    computes it before the encoding."""
    p = np.asarray(p, np.float32)
    bmin = np.asarray(bmin, np.float32)
    ext = (np.asarray(bmax, np.float32) - bmin).astype(np.float32)
    inv = (np.float32(1.0) / ext).astype(np.float32)
    return ((p - bmin).astype(np.float32) * inv).astype(np.float32)


def features(p, n, bmin, bmax, u=None):
    """Encoding of points p (n, 3) with normals n; u (optional) gives the unit
    coordinates directly (e.g. from unit_coords_f32)."""
    if u is None:
        p = np.asarray(p, np.float64)
        u = (p - np.asarray(bmin, np.float64)) / (np.asarray(bmax, np.float64) - np.asarray(bmin, np.float64))
    u = np.asarray(u, np.float64)
    cols = [u]
    for c in range(3):
        for k in range(N_FREQ):
            a = (2.0 ** k) * math.pi * u[:, c]
            cols.append(np.sin(a)[:, None])
            cols.append(np.cos(a)[:, None])
    cols.append(np.asarray(n, np.float64))
    return np.concatenate(cols, axis=1)


def unpack(theta):
    """Flat parameter vector -> [(W (out, in), b (out))] per layer."""
    out, k = [], 0
    for i, o in LAYERS:
        W = theta[k:k + o * i].reshape(o, i)
        k += o * i
        b = theta[k:k + o]
        k += o
        out.append((W, b))
    return out


def forward(theta, x, quant=False):
    """MLP forward.  quant=True rounds every product operand (layer inputs and
    weights) to BF16 as the MFMA kernels do; sums stay exact (FP64).
    Returns (outputs, [layer inputs..., outputs]) -- the layer inputs as the
    products saw them."""
    q = bf16 if quant else (lambda a: a)
    layers = unpack(theta)
    acts = []
    h = x
    for li, (W, b) in enumerate(layers):
        hin = q(h)
        acts.append(hin)
        z = hin @ q(W).T + b
        h = np.maximum(z, 0.0) if li < len(layers) - 1 else z
    acts.append(h)
    return h, acts


def mixture(out):
    o = out.reshape(-1, N_LOBES, 5)
    logit = o[..., 0]
    m = o[..., 1:4]
    kraw = o[..., 4]
    pi = np.exp(logit - logit.max(1, keepdims=True))
    pi /= pi.sum(1, keepdims=True)
    mnorm = np.linalg.norm(m, axis=2, keepdims=True)
    mu = m / np.maximum(mnorm, 1e-12)
    kappa = np.minimum(np.logaddexp(0.0, kraw) + KAPPA_MIN, KAPPA_MAX)
    return pi, mu, kappa, mnorm


def vmf_pdf(w, mu, kappa):
    """w (B, 3), mu (B, L, 3), kappa (B, L) -> (B, L)"""
    cos = np.einsum("bc,blc->bl", w, mu)
    norm = kappa / (2 * math.pi * -np.expm1(-2 * kappa))
    return norm * np.exp(kappa * (cos - 1.0))


def pdf(out, w):
    pi, mu, kappa, _ = mixture(out)
    return np.sum(pi * vmf_pdf(w, mu, kappa), axis=1)


def loss_and_grad(theta, x, w, c, quant=False):
    """Returns (loss, dloss/dtheta, dloss/dout).  The per-sample gradients
    carry the weight c_i and the sums are divided by sum(c) at the end
    (mean_i(c_i / mean(c) * l_i) = sum_i c_i l_i / sum(c)).  quant=True
    rounds every product operand to BF16 as the MFMA kernels do: layer
    inputs, weights and the back-propagated deltas (whose BF16 values also
    form the bias gradients)."""
    q = bf16 if quant else (lambda a: a)
    out, acts = forward(theta, x, quant)
    pi, mu, kappa, mnorm = mixture(out)
    v = vmf_pdf(w, mu, kappa)
    qv = np.sum(pi * v, axis=1)
    csum = np.sum(c)
    loss = np.sum(c * -np.log(qv)) / csum
    q_ = qv
    g = c[:, None]  # d (c * -log q) / d(-log q)
    gamma = pi * v / q_[:, None]
    cos = np.einsum("bc,blc->bl", w, mu)
    d_logit = -(gamma - pi) * g
    d_mu = -(gamma * kappa)[..., None] * w[:, None, :] * g[..., None]
    # through mu = m / |m|
    d_m = (d_mu - np.sum(d_mu * mu, axis=2, keepdims=True) * mu) / np.maximum(mnorm, 1e-12)
    o = out.reshape(-1, N_LOBES, 5)
    kraw = o[..., 4]
    kk = np.logaddexp(0.0, kraw) + KAPPA_MIN
    dlogv_dk = 1.0 / kappa - 2.0 / np.expm1(2 * kappa) + (cos - 1.0)
    d_kappa = -gamma * dlogv_dk * g
    sig = 1.0 / (1.0 + np.exp(-kraw))
    d_kraw = d_kappa * sig * (kk < KAPPA_MAX)
    d_out = np.concatenate([d_logit[..., None], d_m, d_kraw[..., None]], axis=2).reshape(len(x), N_OUT)
    # backprop through the MLP
    layers = unpack(theta)
    grads = []
    delta = d_out
    for li in range(len(layers) - 1, -1, -1):
        W, b = layers[li]
        dq = q(delta)
        grads.append((dq.T @ acts[li], dq.sum(0)))
        if li > 0:
            delta = (dq @ q(W)) * (acts[li] > 0)
    grads = grads[::-1]
    flat = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])
    return loss, flat / csum, d_out / csum


def adam(theta, grad, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015); t counts from 1."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v


def sample(out, u0, u1, u2):
    """Draws one direction per row from the mixture (lobe by CDF of pi on u0,
    vMF by the inverse CDF of the cosine on u1, azimuth 2 pi u2)."""
    pi, mu, kappa, _ = mixture(out)
    cdf = np.cumsum(pi, axis=1)
    j = np.minimum((u0[:, None] > cdf).sum(1), N_LOBES - 1)
    r = np.arange(len(out))
    m, k = mu[r, j], kappa[r, j]
    cw = 1.0 + np.log(u1 + (1 - u1) * np.exp(-2 * k)) / k
    sw = np.sqrt(np.maximum(0.0, 1 - cw * cw))
    phi = 2 * math.pi * u2
    # orthonormal frame around m (Duff et al. 2017)
    sgn = np.where(m[:, 2] >= 0, 1.0, -1.0)
    a = -1.0 / (sgn + m[:, 2])
    b = m[:, 0] * m[:, 1] * a
    b1 = np.stack([1 + sgn * m[:, 0] ** 2 * a, sgn * b, -sgn * m[:, 0]], 1)
    b2 = np.stack([b, sgn + m[:, 1] ** 2 * a, -m[:, 1]], 1)
    return (cw[:, None] * m + (sw * np.cos(phi))[:, None] * b1 + (sw * np.sin(phi))[:, None] * b2)


def init_params(seed):
    """Uniform(-1/sqrt(in), 1/sqrt(in)) weights, zero biases (any init works for
    the checks; the product initialises its own)."""
    rng = np.random.default_rng(seed)
    parts = []
    for i, o in LAYERS:
        bound = 1.0 / math.sqrt(i)
        parts.append(rng.uniform(-bound, bound, o * i))
        parts.append(np.zeros(o))
    return np.concatenate(parts)

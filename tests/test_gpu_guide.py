"""GPU checks of the learned guiding kernels (turner_amd/csrc/guide.hip and
the guided shade of render.hip) against the FP64 oracle oracle/guide.py.
No reference counterpart: turner samples the hemisphere uniformly
(lib/sampling.h:20-32).

Bars (the oracle's analytic gradients are pinned by finite differences in
tests/test_guide_oracle.py):
* forward, loss, gradient against the oracle with the kernels' BF16 operand
  rounding emulated (quant=True): outputs within 3e-3 absolute and 99.9%
  within 1e-4; loss within 1e-4 relative; gradient within 5e-3 relative L2
  (FP32 vs FP64 sums, rare BF16 rounding-boundary flips);
* against the unrounded FP64 model (the BF16 model error): outputs within
  0.02, gradient within 0.1 relative L2 and cosine >= 0.995;
* Adam: the step the kernel takes equals numpy's FP64 Adam on the kernel's
  own gradient within 2e-4 relative (plus half an FP32 ulp of the parameter);
* mixture pdf and sampler against the oracle on the kernel's own outputs:
  pdf within 2e-4 relative, directions within 1e-3 for 99.9%; the pdf
  integrates to 1 over the sphere within 1%;
* training records: exact identities of the estimator on a grey open box
  lit by the background, rel 1e-4;
* guided renders estimate the same image as the oracle's cosine renderer
  (3-sigma z-checks), and a trained guide lowers the error at equal samples.
"""
import math

import numpy as np
import pytest

import oracle as O
import oracle.guide as OG
import turner_amd as T
from turner_amd import scenes
from turner_amd._lib import TN_RENDER_AREA_LIGHTS, TN_RENDER_COSINE

pytestmark = pytest.mark.gpu

BMIN = np.array([-1.0, -0.5, -2.0], np.float32)
BMAX = np.array([1.5, 1.0, 2.0], np.float32)


def _torch():
    import torch
    return torch


def _dev(a):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def _unit(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _theta(seed):
    """Oracle initialisation plus non-zero biases and a spread of output
    scales, so every term of the loss is exercised."""
    rng = np.random.default_rng(seed)
    th = OG.init_params(seed)
    k = 0
    for i, o in OG.LAYERS:
        k += o * i
        th[k:k + o] = rng.normal(0.0, 0.1, o)
        k += o
    # output layer: larger weights so kappas and lobe weights vary
    w4 = len(th) - 40 - 40 * 64
    th[w4:w4 + 40 * 64] *= 4.0
    return th.astype(np.float32).astype(np.float64)


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    p = rng.uniform(BMIN, BMAX, (n, 3)).astype(np.float32)
    nrm = _unit(rng, n).astype(np.float32)
    w = _unit(rng, n).astype(np.float32)
    c = rng.lognormal(0.0, 1.0, n).astype(np.float32)
    c[::7] = 0.0
    return p, nrm, w, c


def _features(p, nrm):
    u = OG.unit_coords_f32(p, BMIN, BMAX)
    return OG.features(None, nrm.astype(np.float64), None, None, u=u)


@pytest.fixture(scope="module")
def guide(gpu):
    from turner_amd.guide import Guide
    g = Guide(BMIN, BMAX, seed=3)
    yield g
    g.close()


def test_forward_matches_oracle(guide):
    th = _theta(1)
    guide.set_params(th)
    assert np.array_equal(guide.params(), th.astype(np.float32))
    p, nrm, _, _ = _batch(3000, 2)  # not a multiple of 64: partial last wave
    out = guide.forward(_dev(p), _dev(nrm)).cpu().numpy().astype(np.float64)
    x = _features(p, nrm)
    ref_q, _ = OG.forward(th, x, quant=True)
    ref, _ = OG.forward(th, x)
    err = np.abs(out - ref_q)
    print(f"[guide fwd] max |d| vs bf16 oracle {err.max():.2e}, within 1e-4: {np.mean(err <= 1e-4):.5f}, "
          f"vs fp64 {np.abs(out - ref).max():.2e}, |out| max {np.abs(ref).max():.2f}", flush=True)
    assert err.max() <= 3e-3
    assert np.mean(err <= 1e-4) >= 0.999
    assert np.abs(out - ref).max() <= 0.02


def test_loss_grad_matches_oracle(guide):
    th = _theta(4)
    guide.set_params(th)
    p, nrm, w, c = _batch(4096 + 37, 5)
    loss, grad = guide.loss_grad(_dev(p), _dev(nrm), _dev(w), _dev(c))
    grad = grad.cpu().numpy().astype(np.float64)
    x = _features(p, nrm)
    lq, gq, _ = OG.loss_and_grad(th, x, w.astype(np.float64), c.astype(np.float64), quant=True)
    l64, g64, _ = OG.loss_and_grad(th, x, w.astype(np.float64), c.astype(np.float64))
    rel_q = np.linalg.norm(grad - gq) / np.linalg.norm(gq)
    rel_64 = np.linalg.norm(grad - g64) / np.linalg.norm(g64)
    cos_64 = grad @ g64 / np.linalg.norm(grad) / np.linalg.norm(g64)
    print(f"[guide grad] loss {loss:.6f} (bf16 oracle {lq:.6f}, fp64 {l64:.6f}); rel L2 vs bf16 oracle "
          f"{rel_q:.2e}, vs fp64 {rel_64:.2e} (cos {cos_64:.5f})", flush=True)
    assert abs(loss - lq) <= 1e-4 * abs(lq)
    assert rel_q <= 5e-3
    # per layer, so a wrong operand layout in one product cannot hide
    k = 0
    for li, (i, o) in enumerate(OG.LAYERS):
        for name, size in (("W", o * i), ("b", o)):
            sl = slice(k, k + size)
            r = np.linalg.norm(grad[sl] - gq[sl]) / max(np.linalg.norm(gq[sl]), 1e-30)
            assert r <= 1e-2, f"layer {li + 1} {name}: rel L2 {r:.3e}"
            k += size
    assert abs(loss - l64) <= 1e-3 * abs(l64)
    assert rel_64 <= 0.1 and cos_64 >= 0.995


def test_zero_weights_give_zero_gradient(guide):
    guide.set_params(_theta(6))
    p, nrm, w, c = _batch(256, 7)
    c[:] = 0.0
    c[5] = 2.0  # one sample carries the whole batch
    loss, grad = guide.loss_grad(_dev(p), _dev(nrm), _dev(w), _dev(c))
    x = _features(p[5:6], nrm[5:6])
    lq, gq, _ = OG.loss_and_grad(_theta(6), x, w[5:6].astype(np.float64), np.array([2.0]), quant=True)
    g = grad.cpu().numpy().astype(np.float64)
    assert abs(loss - lq) <= 1e-4 * abs(lq)
    assert np.linalg.norm(g - gq) <= 5e-3 * np.linalg.norm(gq)


def test_adam_step_matches_numpy(guide):
    th = _theta(8)
    guide.set_params(th)
    p, nrm, w, c = _batch(2048, 9)
    args = (_dev(p), _dev(nrm), _dev(w), _dev(c))
    steps = []
    m = np.zeros_like(th)
    v = np.zeros_like(th)
    cur = th.astype(np.float32).astype(np.float64)
    for t in (1, 2, 3):
        _, g = guide.loss_grad(*args)  # gradient at the current parameters
        g = g.cpu().numpy().astype(np.float64)
        guide.train_step(*args)
        new = guide.params().astype(np.float64)
        ref, m, v = OG.adam(cur, g, m, v, t)
        d_gpu, d_ref = new - cur, ref - cur
        # the kernel's FP32 parameters round each update by half an ulp of theta
        bound = 2e-4 * np.abs(d_ref) + 6e-8 * np.abs(ref) + 1e-9
        assert np.all(np.abs(d_gpu - d_ref) <= bound), np.max(np.abs(d_gpu - d_ref) / bound)
        steps.append(np.abs(d_gpu).max())
        cur = new
    assert max(steps) <= 3.2e-3  # |step| <= lr (1 - b1) / sqrt(1 - b2)


def test_pdf_and_sampler_match_oracle(guide):
    guide.set_params(_theta(10))
    p, nrm, w, _ = _batch(4096, 11)
    out = guide.forward(_dev(p), _dev(nrm)).cpu().numpy().astype(np.float64)
    pdf = guide.pdf(_dev(p), _dev(nrm), _dev(w)).cpu().numpy().astype(np.float64)
    ref = OG.pdf(out, w.astype(np.float64))
    np.testing.assert_allclose(pdf, ref, rtol=2e-4, atol=1e-30)
    rng = np.random.default_rng(12)
    u = rng.uniform(0.0, 1.0, (4096, 3)).astype(np.float32)
    ws, ps = guide.sample(_dev(p), _dev(nrm), _dev(u))
    ws = ws.cpu().numpy().astype(np.float64)
    ps = ps.cpu().numpy().astype(np.float64)
    wr = OG.sample(out, u[:, 0].astype(np.float64), u[:, 1].astype(np.float64), u[:, 2].astype(np.float64))
    close = np.linalg.norm(ws - wr, axis=1) <= 1e-3
    print(f"[guide sample] within 1e-3: {close.mean():.5f}", flush=True)
    assert close.mean() >= 0.999
    np.testing.assert_allclose(np.linalg.norm(ws, axis=1), 1.0, atol=1e-5)
    np.testing.assert_allclose(ps, OG.pdf(out, ws), rtol=2e-4)


def test_pdf_integrates_to_one(guide):
    torch = _torch()
    guide.set_params(_theta(13))
    rng = np.random.default_rng(14)
    npts, ndir = 8, 1 << 18
    for k in range(npts):
        p = np.tile(rng.uniform(BMIN, BMAX, 3).astype(np.float32), (ndir, 1))
        nrm = np.tile(_unit(rng, 1).astype(np.float32), (ndir, 1))
        w = _unit(rng, ndir).astype(np.float32)
        q = guide.pdf(_dev(p), _dev(nrm), _dev(w)).double()
        est = (4 * math.pi * q).mean().item()
        sd = (4 * math.pi * q).std().item() / math.sqrt(ndir)
        assert abs(est - 1.0) <= max(0.01, 4 * sd), (k, est, sd)
    torch.cuda.synchronize()


def _vmf_entropy(k):
    logc = math.log(k / (2 * math.pi * -math.expm1(-2 * k)))
    a = 1.0 / math.tanh(k) - 1.0 / k
    return -(logc + k * (a - 1.0))


def test_training_fits_a_known_distribution(gpu):
    """Directions drawn from vMF(mu(p), kappa = 10) with mu depending on the
    position: Adam on the kernels' gradients brings the loss near the
    target's entropy (its minimum)."""
    from turner_amd.guide import Guide
    g = Guide(BMIN, BMAX, seed=21, lr=5e-3, batch=4096)
    rng = np.random.default_rng(22)
    kappa = 10.0
    h = _vmf_entropy(kappa)
    losses = []
    for step in range(600):
        p, nrm, _, _ = _batch(4096, 1000 + step)
        side = (p[:, 0] > 0.25).astype(np.float64)
        mu = np.stack([np.zeros(4096), 1.0 - 2 * side, np.zeros(4096)], 1)  # +y left, -y right
        out = np.zeros((4096, 40))
        out[:, 0] = 50.0  # one lobe
        out[:, 1:4] = mu
        out[:, 4] = math.log(math.expm1(kappa - OG.KAPPA_MIN))
        w = OG.sample(out, rng.uniform(size=4096), rng.uniform(size=4096), rng.uniform(size=4096))
        c = np.ones(4096, np.float32)
        losses.append(g.train_step(_dev(p), _dev(nrm), _dev(w), _dev(c)))
    first, last = np.mean(losses[:10]), np.mean(losses[-50:])
    print(f"[guide train] loss {first:.3f} -> {last:.3f} (entropy {h:.3f})", flush=True)
    assert first > h + 0.8
    assert last < h + 0.15
    g.close()


def _grey_open_box():
    parts = scenes.cornell_walls(L=1.0)
    verts, normals, albedo, emission = scenes.concat(*parts)
    albedo[:, :3] = 0.7
    return {"verts": verts, "normals": normals, "albedo": albedo, "emission": emission,
            "camera": scenes.camera((0.5, 0.5, -1.2), 50.0), "lights": []}


def test_training_records_identities(gpu):
    """Grey open box (albedo 0.7), no lights, background 0.5: a path's radiance
    is w_J * bg where J is the level that escapes.  With one sample per pixel
    every level-0 record carries c_0 = (pi / rho) lum(I_pixel), consecutive
    records of a path satisfy c_k = (rho / pi) (cos/pdf)_k c_{k+1}, and each
    cos/pdf equals cos / (a q(w) + (1 - a) cos / pi) with the guide's q."""
    torch = _torch()
    from turner_amd.guide import Guide, render_guided
    sc = _grey_open_box()
    tree = T.KDTree(sc["verts"], sc["normals"], sc["albedo"], sc["emission"])
    g = Guide.for_tree(tree, seed=5)
    g.set_params(_theta(30))
    w = 48
    conf = T.TracerConfig(width=w, max_recursion_depth=5, num_pixel_samples=1, num_monte_carlo_samples=1,
                          flags=TN_RENDER_COSINE, bg_color=(0.5, 0.5, 0.5, 1.0), seed=3)
    acc = torch.zeros((w, w, 4), dtype=torch.float32, device="cuda")
    render_guided(tree, sc["camera"], [], conf, g, acc, train_samples=0)
    theta_before = g.params()
    acc.zero_()
    render_guided(tree, sc["camera"], [], conf, g, acc, train_samples=1)
    img = acc.cpu().numpy().reshape(-1, 4)
    ts = g.train_set()
    info = g.info()
    assert info["train_samples"] == len(ts["c"]) > 0 and info["steps"] >= 1
    # the render above trained after recording: compare cos/pdf with the
    # parameters it sampled with
    g.set_params(theta_before)
    rho = 0.7
    lvl, path, c = ts["level"], ts["path"], ts["c"].astype(np.float64)
    assert np.all(c > 0) and lvl.max() < 5
    cosn = np.sum(ts["w"].astype(np.float64) * ts["n"], axis=1)
    assert np.all(cosn > 0)
    np.testing.assert_allclose(np.linalg.norm(ts["w"], axis=1), 1.0, atol=1e-5)
    q = g.pdf(_dev(ts["p"]), _dev(ts["n"]), _dev(ts["w"])).cpu().numpy().astype(np.float64)
    cp = cosn / (0.5 * q + 0.5 * cosn / math.pi)
    np.testing.assert_allclose(ts["cos_pdf"], cp, rtol=1e-4)
    l0 = lvl == 0
    lum = img[path[l0], 0] * 0.2126 + img[path[l0], 1] * 0.7152 + img[path[l0], 2] * 0.0722
    np.testing.assert_allclose(c[l0], math.pi / rho * lum, rtol=1e-4)
    key = {(int(a), int(b)): k for k, (a, b) in enumerate(zip(path, lvl))}
    pairs = 0
    for (pp, k), i in key.items():
        j = key.get((pp, k + 1))
        if j is None:
            continue
        np.testing.assert_allclose(c[i], rho / math.pi * ts["cos_pdf"][i] * c[j], rtol=1e-4)
        pairs += 1
    print(f"[records] {len(c)} samples, {l0.sum()} at level 0, {pairs} parent/child pairs", flush=True)
    assert pairs > 100
    g.close()


def _zcheck_passes(passes, cpu_mean, cpu_sq, n_cpu, what):
    """passes: K independent GPU images (per-pixel means of equal sample
    counts).  The guided estimator's variance differs from the cosine
    estimator's, so sigma combines the spread of the GPU passes with the
    oracle's per-sample variance."""
    passes = np.asarray(passes, np.float64)
    k = len(passes)
    gm = passes.mean(0)
    gvar = passes.var(0, ddof=1) / k
    cvar = np.maximum(cpu_sq / n_cpu - cpu_mean.astype(np.float64) ** 2, 0.0) / n_cpu
    sig = np.sqrt(gvar + cvar)[..., :3]
    diff = (gm - cpu_mean)[..., :3]
    ok = sig > 1e-9
    assert np.all(np.abs(diff[~ok]) <= 1e-5 + 1e-5 * np.abs(cpu_mean[..., :3][~ok]))
    z = diff[ok] / sig[ok]
    frac = np.mean(np.abs(z) > 3)
    print(f"[{what}] pixel-channels {z.size}, |z|>3: {frac:.4f}, mean z {z.mean():+.3f}, "
          f"std z {z.std():.3f}", flush=True)
    # K = 16 passes: sigma is itself estimated, so z follows a Student-like law
    assert frac <= 0.02
    assert abs(z.mean()) < 0.25
    assert 0.7 < z.std() < 1.4


def test_guided_render_is_unbiased(gpu):
    """One-sample MIS with the exact mixture pdf: the guided image has the
    expectation of the cosine + area-light estimator (oracle), with an
    untrained and with a trained guide."""
    torch = _torch()
    from turner_amd.guide import Guide, render_guided
    sc = scenes.sphere_in_cornell(level=3)
    tree = T.KDTree(sc["verts"], sc["normals"], sc["albedo"], sc["emission"])
    ot = O.Tree(sc["verts"])
    w, n_cpu = 24, 1024
    flags = TN_RENDER_COSINE | TN_RENDER_AREA_LIGHTS
    ref, sq, _, _ = O.render(sc, ot, w, w, 3, 1, n_cpu, flags=O.FLAG_COSINE | O.FLAG_AREA_LIGHTS, seed=3,
                             want_sq=True)
    g = Guide.for_tree(tree, seed=7, batch=4096)
    passes, per = 16, 64

    def guided_passes(seed):
        out = []
        for k in range(passes):
            conf = T.TracerConfig(width=w, max_recursion_depth=3, num_pixel_samples=per,
                                  num_monte_carlo_samples=1, flags=flags, seed=seed)
            acc = torch.zeros((w, w, 4), dtype=torch.float32, device="cuda")
            render_guided(tree, sc["camera"], sc["lights"], conf, g, acc, sample_offset=k * per)
            out.append(acc.cpu().numpy() / per)
        return out

    _zcheck_passes(guided_passes(17), ref, sq, n_cpu, "guided, untrained")
    # train on a few passes, then render with the trained guide
    tconf = T.TracerConfig(width=64, max_recursion_depth=3, num_pixel_samples=4, num_monte_carlo_samples=1,
                           flags=flags, seed=19)
    tacc = torch.zeros((64, 64, 4), dtype=torch.float32, device="cuda")
    for k in range(8):
        render_guided(tree, sc["camera"], sc["lights"], tconf, g, tacc, train_samples=1 << 16,
                      sample_offset=k * 4)
    assert g.info()["steps"] >= 8
    _zcheck_passes(guided_passes(23), ref, sq, n_cpu, "guided, trained")
    g.close()


def test_guided_reference_furnace(gpu):
    """The reference's furnace (convex sphere, rho = 0.8, lit only by the
    background c = 0.6 of escaping rays): every guided child escapes, so the
    sphere pixels average rho * c for any proposal that covers the
    hemisphere; m = 4 children per hit."""
    torch = _torch()
    from turner_amd.guide import Guide, render_guided
    rho, cbg = 0.8, 0.6
    sc = scenes.reference_furnace(level=3, rho=rho)
    tree = T.KDTree(sc["verts"], sc["normals"], sc["albedo"], sc["emission"])
    g = Guide.for_tree(tree, seed=9)
    g.set_params(_theta(31))
    w, spp = 32, 256
    conf = T.TracerConfig(width=w, max_recursion_depth=1, num_pixel_samples=spp, num_monte_carlo_samples=4,
                          flags=TN_RENDER_COSINE, bg_color=(cbg, cbg, cbg, 1.0), seed=5)
    acc = torch.zeros((w, w, 4), dtype=torch.float32, device="cuda")
    render_guided(tree, sc["camera"], sc["lights"], conf, g, acc)
    img = acc.cpu().numpy() / spp
    from tests.test_gpu_render import _sphere_pixels
    corner = _sphere_pixels(tree, sc["camera"], w)
    inv = T.TN_INVALID_ID
    on = (corner[:-1, :-1] != inv) & (corner[1:, :-1] != inv) & (corner[:-1, 1:] != inv) & (corner[1:, 1:] != inv)
    off = (corner[:-1, :-1] == inv) & (corner[1:, :-1] == inv) & (corner[:-1, 1:] == inv) & (corner[1:, 1:] == inv)
    assert on.sum() > 50 and off.sum() > 50
    np.testing.assert_allclose(img[off][:, :3], cbg, rtol=1e-6)
    v = img[on][:, :3]
    sem = v.std() / np.sqrt(v.size)
    print(f"[guided furnace] sphere mean {v.mean():.5f} (expect {rho * cbg}), sem {sem:.2e}", flush=True)
    assert abs(v.mean() - rho * cbg) < 3 * sem + 1e-4
    g.close()


def test_trained_guide_lowers_error(gpu):
    """Occluded-light Cornell box (emitter behind a baffle; level-5 mesh):
    after online training, guided rendering has a lower relMSE than the
    cosine tracer at the same number of samples per pixel."""
    torch = _torch()
    from turner_amd.guide import Guide, render_guided
    sc = scenes.occluded_light_cornell(level=5)
    tree = T.KDTree(sc["verts"], sc["normals"], sc["albedo"], sc["emission"])
    w, depth = 64, 6
    flags = TN_RENDER_COSINE | TN_RENDER_AREA_LIGHTS
    from turner_amd.relmse import reference_image, relmse, render_plain
    ref = reference_image(tree, sc, w, depth, flags, spp=4096, seed=101)
    g = Guide.for_tree(tree, seed=11, batch=16384)
    tconf = T.TracerConfig(width=w, max_recursion_depth=depth, num_pixel_samples=16, num_monte_carlo_samples=1,
                           flags=flags, seed=12)
    tacc = torch.zeros((w, w, 4), dtype=torch.float32, device="cuda")
    for k in range(30):
        render_guided(tree, sc["camera"], sc["lights"], tconf, g, tacc, train_samples=1 << 16,
                      sample_offset=16 * k)
    spp = 64
    errs_g, errs_p = [], []
    for rep in range(4):
        conf = T.TracerConfig(width=w, max_recursion_depth=depth, num_pixel_samples=spp,
                              num_monte_carlo_samples=1, flags=flags, seed=200 + rep)
        acc = torch.zeros((w, w, 4), dtype=torch.float32, device="cuda")
        render_guided(tree, sc["camera"], sc["lights"], conf, g, acc)
        errs_g.append(relmse(acc / spp, ref))
        errs_p.append(relmse(render_plain(tree, sc, conf) / spp, ref))
    eg, ep = float(np.mean(errs_g)), float(np.mean(errs_p))
    print(f"[guiding gain] relMSE at {spp} spp: guided {eg:.5f}, cosine {ep:.5f} ({ep / eg:.2f}x)", flush=True)
    assert eg < 0.8 * ep
    g.close()


def test_guide_errors(gpu):
    """Invalid options and calls fail loudly with TurnerError."""
    torch = _torch()
    from turner_amd.guide import Guide, render_guided
    g = Guide(BMIN, BMAX, seed=1)
    for kw in ({"batch": 100}, {"batch": 0}, {"lr": 0.0}, {"mis_guided": 1.5}, {"epochs": 0}):
        with pytest.raises(T.TurnerError):
            g.set_options(**{**{"lr": 1e-3, "batch": 16384, "mis_guided": 0.5, "epochs": 1}, **kw})
    with pytest.raises(T.TurnerError):
        g.set_params(np.full(OG.n_params(), np.nan, np.float32))
    empty = torch.zeros((0, 3), dtype=torch.float32, device="cuda")
    with pytest.raises(T.TurnerError):
        g.loss_grad(empty, empty, empty, torch.zeros(0, device="cuda"))
    sc = scenes.cornell_box()
    tree = T.KDTree(sc["verts"], sc["normals"], sc["albedo"], sc["emission"])
    conf = T.TracerConfig(width=8, max_recursion_depth=2, num_pixel_samples=1, num_monte_carlo_samples=2)
    acc = torch.zeros((8, 8, 4), dtype=torch.float32, device="cuda")
    with pytest.raises(T.TurnerError):  # training records need one path per sample
        render_guided(tree, sc["camera"], sc["lights"], conf, g, acc, train_samples=1024)
    render_guided(tree, sc["camera"], sc["lights"], conf, g, acc)  # inference-only guiding works with m > 1
    assert torch.isfinite(acc).all()
    with pytest.raises(T.TurnerError):  # nothing recorded yet
        g.train_recorded(1 << 10)
    g.close()


def test_train_recorded_takes_more_passes(gpu):
    """tn_guide_train_recorded trains another pass (its own subset, the
    configured number of steps) on the records of the last training render."""
    torch = _torch()
    from turner_amd.guide import Guide, render_guided
    sc = scenes.sphere_in_cornell(level=3)
    tree = T.KDTree(sc["verts"], sc["normals"], sc["albedo"], sc["emission"])
    g = Guide.for_tree(tree, seed=3, batch=4096, epochs=2)
    conf = T.TracerConfig(width=32, max_recursion_depth=3, num_pixel_samples=4, num_monte_carlo_samples=1,
                          flags=TN_RENDER_COSINE | TN_RENDER_AREA_LIGHTS)
    acc = torch.zeros((32, 32, 4), dtype=torch.float32, device="cuda")
    render_guided(tree, sc["camera"], sc["lights"], conf, g, acc, train_samples=1 << 12)
    info = g.info()
    assert info["steps"] == 2 and info["train_samples"] > 0  # 2 epochs x 4096 / 4096
    theta = g.params()
    with pytest.raises(T.TurnerError):
        g.train_recorded(0)
    g.train_recorded(1 << 13)
    torch.cuda.synchronize()
    after = g.info()
    assert after["steps"] == info["steps"] + 4 and after["train_samples"] == info["train_samples"]
    assert np.isfinite(after["loss"]) and not np.array_equal(g.params(), theta)
    g.close()


def test_deferred_training(gpu):
    """TN_RENDER_DEFER_TRAINING: a training render leaves its records pending;
    the next deferred call trains them on the guide's stream beside its own
    render, which samples from the weights as they were before that training
    (its image equals a render by an untrained copy of the guide, up to the
    FP32 accumulation order); tn_guide_flush trains what is left; a guide
    trained this way lowers the error like one trained in order."""
    torch = _torch()
    from turner_amd.guide import Guide, render_guided
    sc = scenes.sphere_in_cornell(level=3)
    tree = T.KDTree(sc["verts"], sc["normals"], sc["albedo"], sc["emission"])
    flags = TN_RENDER_COSINE | TN_RENDER_AREA_LIGHTS
    conf = T.TracerConfig(width=32, max_recursion_depth=3, num_pixel_samples=8, num_monte_carlo_samples=1,
                          flags=flags, seed=5)
    g = Guide.for_tree(tree, seed=9, batch=4096, epochs=2)
    theta0 = g.params()
    acc = torch.zeros((32, 32, 4), dtype=torch.float32, device="cuda")
    render_guided(tree, sc["camera"], sc["lights"], conf, g, acc, train_samples=1 << 12, defer=True)
    assert g.info()["steps"] == 0 and np.array_equal(g.params(), theta0)  # pending, not trained
    # an image pass: the pending set trains beside it, the image uses theta0
    img = torch.zeros((32, 32, 4), dtype=torch.float32, device="cuda")
    render_guided(tree, sc["camera"], sc["lights"], conf, g, img, sample_offset=8, defer=True)
    assert g.info()["steps"] == 2 and not np.array_equal(g.params(), theta0)
    g0 = Guide.for_tree(tree, seed=9, batch=4096, epochs=2)
    assert np.array_equal(g0.params(), theta0)
    ref = torch.zeros((32, 32, 4), dtype=torch.float32, device="cuda")
    render_guided(tree, sc["camera"], sc["lights"], conf, g0, ref, sample_offset=8)
    np.testing.assert_allclose(img.cpu().numpy(), ref.cpu().numpy(), rtol=1e-5, atol=1e-6)
    g0.close()
    # two training renders in a row: the first's set trains during the second
    render_guided(tree, sc["camera"], sc["lights"], conf, g, acc, train_samples=1 << 12, defer=True)
    render_guided(tree, sc["camera"], sc["lights"], conf, g, acc, train_samples=1 << 12, defer=True)
    assert g.info()["steps"] == 4
    g.flush()
    torch.cuda.synchronize()
    assert g.info()["steps"] == 6
    g.flush()  # nothing pending
    torch.cuda.synchronize()
    assert g.info()["steps"] == 6 and np.isfinite(g.info()["loss"])
    g.close()
    # the guiding gain with deferred training (the occluded-light case)
    from turner_amd.relmse import reference_image, relmse, render_plain
    sc = scenes.occluded_light_cornell(level=5)
    tree = T.KDTree(sc["verts"], sc["normals"], sc["albedo"], sc["emission"])
    w, depth = 64, 6
    ref = reference_image(tree, sc, w, depth, flags, spp=4096, seed=101)
    g = Guide.for_tree(tree, seed=11, batch=16384)
    tconf = T.TracerConfig(width=w, max_recursion_depth=depth, num_pixel_samples=16, num_monte_carlo_samples=1,
                           flags=flags, seed=12)
    tacc = torch.zeros((w, w, 4), dtype=torch.float32, device="cuda")
    for k in range(30):
        render_guided(tree, sc["camera"], sc["lights"], tconf, g, tacc, train_samples=1 << 16,
                      sample_offset=16 * k, defer=True)
    g.flush()
    spp = 64
    errs_g, errs_p = [], []
    for rep in range(4):
        c = T.TracerConfig(width=w, max_recursion_depth=depth, num_pixel_samples=spp, num_monte_carlo_samples=1,
                           flags=flags, seed=200 + rep)
        a = torch.zeros((w, w, 4), dtype=torch.float32, device="cuda")
        render_guided(tree, sc["camera"], sc["lights"], c, g, a, defer=True)
        errs_g.append(relmse(a / spp, ref))
        errs_p.append(relmse(render_plain(tree, sc, c) / spp, ref))
    eg, ep = float(np.mean(errs_g)), float(np.mean(errs_p))
    print(f"[deferred guiding gain] relMSE at {spp} spp: guided {eg:.5f}, cosine {ep:.5f} ({ep / eg:.2f}x)",
          flush=True)
    assert eg < 0.8 * ep
    g.close()

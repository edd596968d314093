"""GPU parity at the benchmark workloads' own sizes:

* the 36-triangle Cornell box in full (320 x 320, max-depth 3, m 1, 8 pixel
  samples, the reference estimator and pixel jitter) against the oracle's
  trace() with the same jitter: the deterministic depth-0 part within 1e-5,
  the estimator per pixel within 3 sigma over 32 renders each;
* the level-8 displaced icosphere (1.31 M triangles) in the Cornell
  enclosure with the 0.2 x 0.2 emitter, 1920 x 1080, max-depth 3, m 4,
  cosine + area-light NEE: three image rows against the oracle, z-tests on
  40-pixel tiles;
* the 32-scene batch (levels 3..8, widths 128..1024, 16..64 spp): the white
  furnace closed form in scene 0 and batch = single-scene render for four
  scenes;
* the occluded-light scene with the level-8 mesh at 512 x 512, max-depth 6:
  the guided tracer (trained online as the benchmark trains it) estimates
  the plain tracer's image (16 x 16 tile z-tests over 16 repeats each) and has a
  lower relMSE than the plain tracer at equal samples.
"""
import numpy as np
import pytest

import oracle as O
import turner_amd as T
from turner_amd import scenes
from turner_amd._lib import TN_RENDER_AREA_LIGHTS, TN_RENDER_COSINE, TN_RENDER_REF_JITTER

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def ztest(diff, se, what, frac_max=0.015):
    z = (diff / se).ravel()
    frac = float(np.mean(np.abs(z) > 3))
    print(f"[{what}] {z.size} z-values, |z|>3: {frac:.4f}, mean {z.mean():+.3f}, std {z.std():.3f}", flush=True)
    assert frac <= frac_max and abs(z.mean()) < 0.25 and 0.7 < z.std() < 1.4, what


def test_cornell_320_full(gpu):
    """320 x 320, -d 3 -m 1 -p 8, reference jitter: the benchmark's CPU case."""
    import torch
    sc = scenes.cornell_box()
    tree = T.KDTree(sc["verts"], sc["normals"], sc["albedo"], sc["emission"])
    ot = O.Tree(sc["verts"])
    assert np.array_equal(ot.nodes, tree.flat_nodes())
    w, spp, runs = 320, 8, 32
    det = T.TracerConfig(width=w, max_recursion_depth=0, num_pixel_samples=spp, num_monte_carlo_samples=1,
                         flags=TN_RENDER_REF_JITTER)
    from tests.test_gpu_render import _render
    img, st = _render(tree, sc, det)
    ref, _, nr, _ = O.render(sc, ot, w, w, 0, 1, spp, flags=O.FLAG_REF_RNG)
    assert st["rays"] == nr == w * w * spp
    np.testing.assert_allclose(img, ref, rtol=1e-5, atol=1e-6)
    # the full estimator: every render repeats the reference's jitter, so the
    # 32 renders on each side share the primary rays and differ in bounces
    acc = torch.zeros((w, w, 4), dtype=torch.float32, device="cuda")
    conf = T.TracerConfig(width=w, max_recursion_depth=3, num_pixel_samples=spp, num_monte_carlo_samples=1,
                          flags=TN_RENDER_REF_JITTER)
    g_rays = 0
    for k in range(runs):
        conf.seed = 100 + k
        g_rays += T.render_accumulate(tree, sc["camera"], sc["lights"], conf, acc)["rays"]
    gpu = acc.cpu().numpy().astype(np.float64) / (spp * runs)
    s1 = np.zeros((w, w, 4))
    s2 = np.zeros((w, w, 4))
    c_rays = 0
    for k in range(runs):
        m, sq, nr, _ = O.render(sc, ot, w, w, 3, 1, spp, flags=O.FLAG_REF_RNG, seed=k + 1, want_sq=True)
        s1 += m.astype(np.float64) * spp
        s2 += sq
        c_rays += nr
    n = spp * runs
    cpu = s1 / n
    var = np.maximum(s2 / n - cpu**2, 0)[..., :3]
    se = np.sqrt(var * 2 / n)
    live = se > 1e-9
    assert np.all(np.abs(gpu[..., :3] - cpu[..., :3])[~live] <= 1e-5 * (1 + np.abs(cpu[..., :3][~live])))
    ztest((gpu[..., :3] - cpu[..., :3])[live], se[live], "cornell 320 d3 m1 p8")
    # segment counts: same expectation (the open side lets rays escape)
    assert abs(g_rays - c_rays) < 5 * np.sqrt(c_rays)


def test_level8_sphere_1080p_rows(gpu):
    """1920 x 1080, -d 3 -m 4, level-8 mesh, cosine + area-light NEE (rendering
    kd-tree, as benchmarked): rows 100, 540 and 980 against the oracle."""
    import torch
    sc = scenes.sphere_in_cornell(level=8)
    tree = T.KDTree(sc["verts"], sc["normals"], sc["albedo"], sc["emission"], empty_leaves=True)
    w, h, reps, spp_g = 1920, 1080, 16, 4
    flags = TN_RENDER_COSINE | TN_RENDER_AREA_LIGHTS
    rows = [100, 540, 980]
    reps_img = []
    for k in range(reps):
        conf = T.TracerConfig(width=w, aspect=w / h, max_recursion_depth=3, num_pixel_samples=spp_g,
                              num_monte_carlo_samples=4, flags=flags, seed=31)
        acc = torch.zeros((h, w, 4), dtype=torch.float32, device="cuda")
        T.render_accumulate(tree, sc["camera"], sc["lights"], conf, acc, sample_offset=k * spp_g)
        reps_img.append(acc[rows].cpu().numpy().astype(np.float64) / spp_g)
    print("[1080p] GPU renders done; oracle tree build", flush=True)
    ot = O.Tree(sc["verts"])
    spp_c = 4
    # one oracle thread per row: rows 100 + 440 j
    m, sq, _, _ = O.render(sc, ot, w, h, 3, 4, spp_c, flags=O.FLAG_COSINE | O.FLAG_AREA_LIGHTS, seed=5,
                           row_step=440, row_offset=100, want_sq=True, nthreads=3)
    cpu, csq = m[rows].astype(np.float64), sq[rows].astype(np.float64)
    tile = 40
    g = np.stack(reps_img)  # reps x 3 rows x w x 4
    gt = g[..., :3].reshape(reps, 3, w // tile, tile, 3).mean(3)  # per-rep tile means
    ct = cpu[..., :3].reshape(3, w // tile, tile, 3).mean(2)
    # variance of a tile mean: GPU from the spread of its repeats, CPU from the per-pixel sample variances
    g_var = gt.var(0, ddof=1) / reps
    cvar_pix = np.maximum(csq / spp_c - cpu**2, 0)[..., :3].reshape(3, w // tile, tile, 3)
    c_var = cvar_pix.sum(2) / (tile**2 * spp_c)
    ztest(gt.mean(0) - ct, np.sqrt(g_var + c_var), "1080p level-8 rows, 40-px tiles", frac_max=0.03)


def test_batch_32_scenes(gpu):
    """The 32-scene batch of the benchmark: furnace closed form in scene 0,
    and four scenes equal to their single-scene renders (same seeds)."""
    import torch
    from turner_amd.batch import Batch, heterogeneous_workload
    from tests.test_gpu_batch import MASK, STEP, _single
    scs, levels, jobs = heterogeneous_workload(seed=2024)
    trees = [T.KDTree(sc["verts"], sc["normals"], sc["albedo"], sc["emission"], empty_leaves=True) for sc in scs]
    print(f"[batch] {len(trees)} trees, levels {levels}", flush=True)
    b = Batch(trees)
    flags = TN_RENDER_COSINE | TN_RENDER_AREA_LIGHTS
    seed = 2
    imgs, st, _ = b.render(jobs, max_depth=3, mc_samples=1, flags=flags, seed=seed)
    torch.cuda.synchronize()
    # furnace: pixels whose four corner rays hit the sphere see rho L = 0.8, the rest 1
    from tests.test_gpu_render import _sphere_pixels
    f = imgs[0].cpu().numpy()
    corner = _sphere_pixels(trees[0], scs[0]["camera"], jobs[0]["width"])
    nsph = 20 * 4**4
    on = (corner[:-1, :-1] < nsph) & (corner[1:, :-1] < nsph) & (corner[:-1, 1:] < nsph) & (corner[1:, 1:] < nsph)
    off = (corner[:-1, :-1] >= nsph) & (corner[1:, :-1] >= nsph) & (corner[:-1, 1:] >= nsph) & \
        (corner[1:, 1:] >= nsph)
    np.testing.assert_allclose(f[off][:, :3], 1.0, rtol=1e-6)
    v = f[on][:, :3]
    sem = v.std() / np.sqrt(v.size)
    print(f"[batch furnace] sphere mean {v.mean():.5f} (expect 0.8), sem {sem:.2e}", flush=True)
    assert on.sum() > 1000 and abs(v.mean() - 0.8) < 3 * sem + 1e-4
    for s in (9, 14, 24, 26):  # levels 8, 3, 6, 3; widths 319, 191, 198, 293
        ref, _ = _single(trees[s], scs[s], jobs[s], 3, 1, flags, (seed + s * STEP) & MASK)
        got = imgs[s].cpu().numpy() * jobs[s]["spp"]
        np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-5, err_msg=f"scene {s}")
    assert st["primary_rays"] == sum(j["width"] * j["height"] * j["spp"] for j in jobs)
    b.close()


def test_occluded_light_level8_guided(gpu):
    """512 x 512, max-depth 6, level-8 mesh behind-baffle emitter: the guide,
    trained as the benchmark trains it (8 passes of 1 spp, 2^18 samples, 4
    epochs), leaves the image's expectation unchanged and lowers its error."""
    import torch
    from turner_amd.guide import Guide, render_guided
    from turner_amd.relmse import reference_image, relmse
    sc = scenes.occluded_light_cornell(level=8)
    tree = T.KDTree(sc["verts"], sc["normals"], sc["albedo"], sc["emission"], empty_leaves=True)
    w, depth = 512, 6
    flags = TN_RENDER_COSINE | TN_RENDER_AREA_LIGHTS
    g = Guide.for_tree(tree, seed=2, epochs=4)
    tacc = torch.zeros((w, w, 4), dtype=torch.float32, device="cuda")
    tconf = T.TracerConfig(width=w, max_recursion_depth=depth, num_pixel_samples=1, num_monte_carlo_samples=1,
                           flags=flags, seed=12)
    for k in range(8):
        render_guided(tree, sc["camera"], sc["lights"], tconf, g, tacc, train_samples=1 << 18, sample_offset=k)
    reps, spp = 16, 8
    tile = 16  # 2048 samples per tile and repeat: the plain tracer's rare bright paths skew smaller tiles

    def tiles(img):  # (h, w, 4) sums -> per-tile means of RGB
        x = img[..., :3].double() / spp
        return x.reshape(w // tile, tile, w // tile, tile, 3).mean((1, 3))

    tg, tp, eg, ep = [], [], [], []
    ref, ref_var = reference_image(tree, sc, w, depth, flags, spp=1024, seed=777, with_var=True)
    for k in range(reps):
        conf = T.TracerConfig(width=w, max_recursion_depth=depth, num_pixel_samples=spp, num_monte_carlo_samples=1,
                              flags=flags, seed=500 + k)
        acc = torch.zeros((w, w, 4), dtype=torch.float32, device="cuda")
        render_guided(tree, sc["camera"], sc["lights"], conf, g, acc, sample_offset=64 + k * spp)
        pacc = torch.zeros_like(acc)
        T.render_accumulate(tree, sc["camera"], sc["lights"], conf, pacc, sample_offset=64 + k * spp)
        tg.append(tiles(acc).cpu().numpy())
        tp.append(tiles(pacc).cpu().numpy())
        eg.append(relmse(acc / spp, ref, ref_var))
        ep.append(relmse(pacc / spp, ref, ref_var))
    tg, tp = np.stack(tg), np.stack(tp)
    se = np.sqrt(tg.var(0, ddof=1) / reps + tp.var(0, ddof=1) / reps)
    live = se > 1e-12
    ztest((tg.mean(0) - tp.mean(0))[live], se[live], "occluded level-8: guided vs plain, 16x16 tiles", frac_max=0.02)
    mg, mp = float(np.mean(eg)), float(np.mean(ep))
    print(f"[occluded level-8] relMSE at {spp} spp: guided {mg:.4f}, plain {mp:.4f} ({mp / mg:.2f}x)", flush=True)
    assert mg < 0.6 * mp
    g.close()

"""GPU checks of batched heterogeneous rendering (tn_batch_*): many scenes in
one packed kd-tree forest, one launch sequence per chunk of samples, hipGraph
replay and the autotuner.  No reference counterpart.

Bars:
* a batch renders every scene exactly as a single-scene render with the
  seed seed + s * 0x9E3779B97F4A7C15 (same hits, same random streams): per
  pixel within 1e-5 relative + 1e-6 absolute (only the order of the FP32
  atomic accumulation differs);
* graph replay and autotuned launch shapes leave the image unchanged (same
  bar);
* the white furnace as scene 0 of a batch: enclosure pixels exactly L,
  sphere pixels rho * L within 3 sigma.
"""
import numpy as np
import pytest

import turner_amd as T
from turner_amd import scenes
from turner_amd._lib import (TN_RENDER_AREA_LIGHTS, TN_RENDER_AUTOTUNE, TN_RENDER_COSINE, TN_RENDER_GRAPH,
                             TN_RENDER_PERSISTENT)

pytestmark = pytest.mark.gpu

STEP = 0x9E3779B97F4A7C15
MASK = (1 << 64) - 1


def _tree(sc):
    return T.KDTree(sc["verts"], sc["normals"], sc["albedo"], sc["emission"])


def _single(tree, sc, job, depth, mc, flags, seed, chunk=0):
    import torch
    conf = T.TracerConfig(width=job["width"], aspect=job["width"] / job["height"], max_recursion_depth=depth,
                          num_pixel_samples=job["spp"], num_monte_carlo_samples=mc, flags=flags, seed=seed,
                          bg_color=job.get("bg", (0.0, 0.0, 0.0, 1.0)), chunk_paths=chunk)
    acc = torch.zeros((job["height"], job["width"], 4), dtype=torch.float32, device="cuda")
    st = T.render_accumulate(tree, sc["camera"], sc["lights"], conf, acc)
    return acc.cpu().numpy(), st


def _scenes():
    return [scenes.cornell_box(), scenes.sphere_in_cornell(level=3), scenes.furnace(level=3),
            scenes.occluded_light_cornell(level=4)]


def _jobs(scs):
    sizes = [(16, 16, 4), (24, 20, 3), (32, 32, 2), (20, 28, 5)]
    return [{"camera": sc["camera"], "lights": sc["lights"], "width": w, "height": h, "spp": spp,
             "bg": (0.1, 0.2, 0.3, 1.0)} for sc, (w, h, spp) in zip(scs, sizes)]


def test_batch_matches_single_renders(gpu):
    from turner_amd.batch import Batch
    scs = _scenes()
    trees = [_tree(sc) for sc in scs]
    jobs = _jobs(scs)
    b = Batch(trees)
    flags = TN_RENDER_COSINE | TN_RENDER_AREA_LIGHTS
    for depth, mc in ((2, 1), (2, 2)):
        seed = 7
        imgs, st, _ = b.render(jobs, max_depth=depth, mc_samples=mc, flags=flags, seed=seed)
        rays = 0
        for s, (tree, sc, job) in enumerate(zip(trees, scs, jobs)):
            ref, sst = _single(tree, sc, job, depth, mc, flags, (seed + s * STEP) & MASK)
            got = imgs[s].cpu().numpy() * job["spp"]
            np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-6, err_msg=f"scene {s} d{depth} m{mc}")
            rays += sst["rays"]
        assert st["rays"] == rays
    b.close()


def test_batch_of_refined_scenes_matches_single_renders(gpu):
    """Scenes built with TN_SCENE_REFINED: the batch traverses their
    reference trees, their single renders the refined trees; both give the
    reference tree's hits, so the images agree as for any batch."""
    from turner_amd.batch import Batch
    scs = _scenes()
    trees = [T.KDTree(sc["verts"], sc["normals"], sc["albedo"], sc["emission"], refined=True) for sc in scs]
    jobs = _jobs(scs)
    b = Batch(trees)
    flags = TN_RENDER_COSINE | TN_RENDER_AREA_LIGHTS
    imgs, st, _ = b.render(jobs, max_depth=2, mc_samples=2, flags=flags, seed=3)
    rays = 0
    for s, (tree, sc, job) in enumerate(zip(trees, scs, jobs)):
        ref, sst = _single(tree, sc, job, 2, 2, flags, (3 + s * STEP) & MASK)
        np.testing.assert_allclose(imgs[s].cpu().numpy() * job["spp"], ref, rtol=1e-5, atol=1e-6,
                                   err_msg=f"scene {s}")
        rays += sst["rays"]
    assert st["rays"] == rays
    b.close()


def test_persistent_launch_matches_wavefront(gpu):
    """TN_RENDER_PERSISTENT: one cooperative launch walks every chunk and
    level (grid barriers between the stages).  The same paths as the
    launch-per-stage wavefront (equal segment and shadow-ray counts, pixels
    within FP32 accumulation order), for single renders of several chunks
    (m = 1 and m = 2, uniform and cosine sampling) and for a batch."""
    from turner_amd.batch import Batch
    sc = scenes.sphere_in_cornell(level=3)
    tree = _tree(sc)
    job = {"camera": sc["camera"], "lights": sc["lights"], "width": 48, "height": 40, "spp": 6}
    for mc, flags in ((1, TN_RENDER_COSINE | TN_RENDER_AREA_LIGHTS), (2, 0)):
        a, sa = _single(tree, sc, job, 3, mc, flags, 5, chunk=4096)
        b, sb = _single(tree, sc, job, 3, mc, flags | TN_RENDER_PERSISTENT, 5, chunk=4096)
        assert (sa["rays"], sa["shadow_rays"]) == (sb["rays"], sb["shadow_rays"]), mc
        np.testing.assert_allclose(b, a, rtol=1e-5, atol=1e-6, err_msg=f"m {mc}")
    scs = _scenes()
    trees = [_tree(x) for x in scs]
    jobs = _jobs(scs)
    bt = Batch(trees)
    flags = TN_RENDER_COSINE | TN_RENDER_AREA_LIGHTS
    ia, sa, _ = bt.render(jobs, max_depth=2, mc_samples=1, flags=flags, seed=9, chunk_paths=2048)
    ib, sb, _ = bt.render(jobs, max_depth=2, mc_samples=1, flags=flags | TN_RENDER_PERSISTENT, seed=9,
                          chunk_paths=2048)
    assert (sa["rays"], sa["shadow_rays"]) == (sb["rays"], sb["shadow_rays"])
    for s, (x, y) in enumerate(zip(ia, ib)):
        np.testing.assert_allclose(y.cpu().numpy(), x.cpu().numpy(), rtol=1e-5, atol=1e-6, err_msg=f"scene {s}")
    bt.close()


def test_graph_replay_matches_direct(gpu):
    sc = scenes.sphere_in_cornell(level=3)
    tree = _tree(sc)
    job = {"width": 40, "height": 40, "spp": 4}
    flags = TN_RENDER_COSINE | TN_RENDER_AREA_LIGHTS
    a, sa = _single(tree, sc, job, 3, 1, flags, 5, chunk=997)
    g, sg = _single(tree, sc, job, 3, 1, flags | TN_RENDER_GRAPH, 5, chunk=997)
    np.testing.assert_allclose(g, a, rtol=1e-5, atol=1e-6)
    assert sa["rays"] == sg["rays"] and sa["shadow_rays"] == sg["shadow_rays"]


def test_autotuned_batch_matches_untuned(gpu):
    from turner_amd.batch import Batch
    scs = _scenes()
    trees = [_tree(sc) for sc in scs]
    jobs = _jobs(scs)
    for j in jobs:
        j["spp"] *= 64  # enough samples for several chunk sizes
    b = Batch(trees)
    flags = TN_RENDER_COSINE | TN_RENDER_AREA_LIGHTS
    ref, _, _ = b.render(jobs, max_depth=2, mc_samples=1, flags=flags, seed=3)
    ref = [r.cpu().numpy() for r in ref]
    got, st, _ = b.render(jobs, max_depth=2, mc_samples=1, flags=flags | TN_RENDER_AUTOTUNE | TN_RENDER_GRAPH,
                          seed=3)
    tune = b.tuning()
    print(f"[autotune] {tune}", flush=True)
    assert tune["candidates"] >= 3 and tune["chunk_paths"] in (1 << 18, 1 << 20, 1 << 22)
    assert tune["blocks_per_cu"] in (4, 8, 16) and tune["best_rate"] > 0
    for s in range(len(jobs)):
        np.testing.assert_allclose(got[s].cpu().numpy(), ref[s], rtol=1e-5, atol=1e-6)
    # cached: the second call does not probe again
    b.render(jobs, max_depth=2, mc_samples=1, flags=flags | TN_RENDER_AUTOTUNE | TN_RENDER_GRAPH, seed=4)
    assert b.tuning() == tune
    b.close()


def test_furnace_as_scene_zero(gpu):
    from turner_amd.batch import Batch
    from tests.test_gpu_render import _sphere_pixels
    rho, L = 0.8, 1.0
    scs = [scenes.furnace(level=3, rho=rho, radiance=L), scenes.cornell_box(), scenes.sphere_in_cornell(level=3)]
    trees = [_tree(sc) for sc in scs]
    w, spp = 32, 256
    jobs = [{"camera": sc["camera"], "lights": sc["lights"], "width": w, "height": w, "spp": spp} for sc in scs]
    b = Batch(trees)
    imgs, _, _ = b.render(jobs, max_depth=2, mc_samples=1, flags=TN_RENDER_COSINE | TN_RENDER_AREA_LIGHTS, seed=9)
    img = imgs[0].cpu().numpy()
    nsph = 20 * 4 ** 3
    corner = _sphere_pixels(trees[0], scs[0]["camera"], w)
    on = (corner[:-1, :-1] < nsph) & (corner[1:, :-1] < nsph) & (corner[:-1, 1:] < nsph) & (corner[1:, 1:] < nsph)
    off = (corner[:-1, :-1] >= nsph) & (corner[1:, :-1] >= nsph) & (corner[:-1, 1:] >= nsph) & (corner[1:, 1:] >= nsph)
    np.testing.assert_allclose(img[off][:, :3], L, rtol=1e-6)
    v = img[on][:, :3]
    sem = v.std() / np.sqrt(v.size)
    print(f"[batch furnace] sphere mean {v.mean():.5f} (expect {rho * L}), sem {sem:.2e}", flush=True)
    assert abs(v.mean() - rho * L) < 3 * sem + 1e-4
    b.close()


def test_batch_errors(gpu):
    from turner_amd.batch import Batch
    sc = scenes.cornell_box()
    b = Batch([_tree(sc)])
    with pytest.raises(T.TurnerError):
        b.render([{"camera": sc["camera"], "width": 0, "height": 4, "spp": 1}])
    with pytest.raises(ValueError):
        b.render([])
    b.close()

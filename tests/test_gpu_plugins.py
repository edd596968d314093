"""GPU parity of the reference's other two trace() plugins, the raycaster
(raycaster.cpp:7-24) and the raytracer (raytracer.cpp:6-67), run on the same
traversal and queues as the path tracer (TN_RENDERER_*), against the oracle's
recursive restatements (themselves pinned by closed forms in test_oracle.py).

Bars (reference pixel jitter, so primary rays are bit-identical):
* raycaster: bit-exact per pixel at one sample per pixel; ray counts equal;
* raytracer without reflective triangles: bit-exact at one sample per pixel;
* raytracer with mirror chains: within 1e-5 relative + 1e-6 absolute (the
  wavefront sums a chain's contributions level by level, the reference
  nests them), most pixels still bit-exact; ray and shadow-ray counts equal;
* several samples per pixel: within 1e-6 relative (the per-pixel sum is
  accumulated in a different order).
"""
import numpy as np
import pytest

import oracle as O
import turner_amd as T
from tests import helpers as H
from turner_amd import scenes
from turner_amd._lib import TN_RENDER_REF_JITTER, TN_RENDERER_RAYCASTER, TN_RENDERER_RAYTRACER

from tests.test_gpu_render import _render, make

pytestmark = pytest.mark.gpu

BG = (0.1, 0.2, 0.3, 1.0)


def _mirror_box():
    """The benchmark Cornell box with a mirror-like tall box (reflectivity 0.6)
    and a half-reflective back wall, so mirror rays chain between them."""
    sc = scenes.cornell_box()
    n = len(sc["verts"])
    sc["reflective"] = np.tile(np.array([0.95, 0.9, 0.85, 1.0], np.float32), (n, 1))
    sc["reflectivity"] = np.zeros(n, np.float32)
    sc["reflectivity"][4:6] = 0.5  # back wall (walls are 5 quads: floor, ceiling, back, left, right)
    sc["reflectivity"][22:34] = 0.6  # tall box (after 10 wall and 12 short-box triangles)
    return sc


def _conf(renderer, w, spp=1, depth=3, **kw):
    return T.TracerConfig(width=w, max_recursion_depth=depth, num_pixel_samples=spp, num_monte_carlo_samples=1,
                          flags=TN_RENDER_REF_JITTER, renderer=renderer, bg_color=BG, **kw)


def _oracle(sc, ot, renderer, w, spp=1, depth=3, **kw):
    img, _, nr, ns = O.render(sc, ot, w, w, depth, 1, spp, flags=O.FLAG_REF_RNG, bg=BG, nthreads=1,
                              renderer=renderer, **kw)
    return img, nr, ns


def _tree(sc):
    tree, ot = make(sc)
    if "reflectivity" in sc:
        tree.set_reflection(sc["reflective"], sc["reflectivity"])
    return tree, ot


def test_raycaster_bit_exact(gpu):
    sc = scenes.cornell_box()
    tree, ot = _tree(sc)
    w = 64
    for mv in (2000.0, 1500.0):
        img, st = _render(tree, sc, _conf(TN_RENDERER_RAYCASTER, w, max_visibility=mv))
        ref, nr, ns = _oracle(sc, ot, O.RENDERER_RAYCASTER, w, max_visibility=mv)
        assert np.array_equal(img, ref)
        assert st["rays"] == nr and st["shadow_rays"] == ns == 0
        a = img[..., 3]
        assert 0 <= a.min() < a.max() <= 1 and np.mean((a > 0) & (a < 1)) > 0.5  # fades with distance
    img, _ = _render(tree, sc, _conf(TN_RENDERER_RAYCASTER, w, spp=4, max_visibility=2000.0))
    ref, _, _ = _oracle(sc, ot, O.RENDERER_RAYCASTER, w, spp=4, max_visibility=2000.0)
    np.testing.assert_allclose(img, ref, rtol=1e-6, atol=1e-7)


def test_raycaster_misses_are_background(gpu):
    sc = H.wall_scene(z=-3.0)  # behind the camera
    tree, ot = _tree(sc)
    img, st = _render(tree, sc, _conf(TN_RENDERER_RAYCASTER, 16))
    assert np.all(img == np.float32(BG)) and st["rays"] == 0


def test_raytracer_bit_exact_without_reflection(gpu):
    for sc in (scenes.cornell_box(), H.wall_scene(occluder=True)):
        tree, ot = _tree(sc)
        w = 64
        for si in (0.5, 1.0):
            img, st = _render(tree, sc, _conf(TN_RENDERER_RAYTRACER, w, shadow_intensity=si))
            ref, nr, ns = _oracle(sc, ot, O.RENDERER_RAYTRACER, w, shadow_intensity=si)
            assert np.array_equal(img, ref)
            assert st["rays"] == nr and st["shadow_rays"] == ns


def test_raytracer_mirror_chains(gpu):
    sc = _mirror_box()
    tree, ot = _tree(sc)
    w = 96
    for depth in (1, 3, 6):
        img, st = _render(tree, sc, _conf(TN_RENDERER_RAYTRACER, w, depth=depth))
        ref, nr, ns = _oracle(sc, ot, O.RENDERER_RAYTRACER, w, depth=depth)
        np.testing.assert_allclose(img, ref, rtol=1e-5, atol=1e-6)
        exact = np.mean(np.all(img == ref, -1))
        print(f"[raytracer d{depth}] bit-exact pixels {exact:.3f}, rays {st['rays']}, shadow {st['shadow_rays']}",
              flush=True)
        assert exact > 0.5
        assert st["rays"] == nr and st["shadow_rays"] == ns
    # the mirrors matter: the image differs from the one without reflection
    plain = scenes.cornell_box()
    t2, _ = _tree(plain)
    img0, _ = _render(t2, plain, _conf(TN_RENDERER_RAYTRACER, w, depth=3))
    assert np.abs(img0 - img).max() > 0.05


def test_raytracer_fine_mesh(gpu):
    """Level-5 displaced icosphere (reflectivity 0.5) in the Cornell enclosure, point light."""
    sc = scenes.sphere_in_cornell(level=5)
    sc["lights"] = [((0.5, 0.9, 0.3), (1.0, 1.0, 1.0, 1.0))]
    n = len(sc["verts"])
    sc["reflective"] = np.tile(np.array([1.0, 1.0, 1.0, 1.0], np.float32), (n, 1))
    sc["reflectivity"] = np.where(np.arange(n) >= 10, 0.5, 0.0).astype(np.float32)
    tree, ot = _tree(sc)
    w = 64
    img, st = _render(tree, sc, _conf(TN_RENDERER_RAYTRACER, w, spp=2, depth=4))
    ref, nr, ns = _oracle(sc, ot, O.RENDERER_RAYTRACER, w, spp=2, depth=4)
    np.testing.assert_allclose(img, ref, rtol=1e-5, atol=1e-6)
    assert st["rays"] == nr and st["shadow_rays"] == ns


def test_plugin_errors(gpu):
    sc = scenes.cornell_box()
    tree, _ = _tree(sc)
    nolight = dict(sc, lights=[])
    with pytest.raises(T.TurnerError):
        _render(tree, nolight, _conf(TN_RENDERER_RAYTRACER, 8))  # lights.front() (raytracer.cpp:14)
    with pytest.raises(T.TurnerError):
        _render(tree, sc, _conf(7, 8))
    with pytest.raises(ValueError):
        tree.set_reflection(np.zeros((3, 4), np.float32), np.zeros(3, np.float32))
    from turner_amd.guide import Guide, render_guided
    import torch
    g = Guide.for_tree(tree)
    acc = torch.zeros((8, 8, 4), dtype=torch.float32, device="cuda")
    with pytest.raises(T.TurnerError):
        render_guided(tree, sc["camera"], sc["lights"], _conf(TN_RENDERER_RAYCASTER, 8), g, acc)
    g.close()
    from turner_amd.batch import Batch
    b = Batch([tree])
    with pytest.raises(T.TurnerError):
        b.render([{"camera": sc["camera"], "lights": sc["lights"], "width": 8, "height": 8, "spp": 1}],
                 max_depth=1, mc_samples=1, renderer=TN_RENDERER_RAYCASTER)
    b.close()

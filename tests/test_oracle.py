"""CPU tests pinning the oracle (oracle/oracle.c) against the reference's own
known-answer tests, closed forms and independent computations."""
import math
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from tests import helpers as H

INV = O.INVALID_ID
EPS = 1e-5


def test_flat_node_encoding():
    """tests/test_kdtree.cpp:261-337"""
    split = np.float32(3.1415)
    for ax in range(3):
        n = O.flat_inner(ax, split, 42)
        assert not O.is_leaf(n) and O.split_axis(n) == ax
        assert O.split_pos(n) == split and O.right(n) == 42
        # set_right: the builder writes the right index after creating the node
        n0 = O.flat_inner(ax, split, 0)
        n42 = (n0 & 0xFFFFFFFF00000000) | (42 << 2) | ax
        assert n42 == n
    leaf = O.flat_leaf(1)
    assert O.is_leaf(leaf) and O.first_id(leaf) == 1 and not O.has_second(leaf)
    leaf = O.flat_leaf(1, 2)
    assert O.is_leaf(leaf) and O.first_id(leaf) == 1 and O.second_id(leaf) == 2
    assert O.has_second(leaf)


def test_kdtree_shape_kats():
    """tests/test_kdtree.cpp:16-31, 162-186"""
    two = np.stack([H.tri((-1, -1, 0), (1, -1, 0), (1, 1, 0)), H.tri((0, 0, 0), (1, 0, 0), (1, 2, 0))])
    assert O.Tree(two).height() == 0
    t4 = O.Tree(H.four_separated())
    assert t4.height() == 1 and t4.num_nodes() == 5
    tc = O.Tree(H.cube12())
    assert tc.height() == 0 and tc.num_nodes() == 7


def test_four_separated_hits():
    """tests/test_kdtree.cpp:32-62: exact r, s, t"""
    t = O.Tree(H.four_separated())
    o = np.zeros((4, 3), np.float32)
    d = np.array([[0.5, 0.5, 1], [2.5, 0.5, 1], [0.5, 2.5, 1], [2.5, 2.5, 1]], np.float32)
    ids, r, s, tt = O.intersect(t.pre, t.nodes, t.box, o, d)
    assert list(ids) == [0, 1, 2, 3]
    assert list(r) == [1, 1, 1, 1]
    assert list(s) == [0.5, 0, 0, 0] and list(tt) == [0.5, 0.5, 0.5, 0.5]


def test_coplanar_kats():
    """tests/test_kdtree.cpp:188-245"""
    rng = np.random.default_rng(3)
    for ax in range(3):
        t = O.Tree(H.coplanar_stack(ax, rng))
        o = np.zeros((2, 3), np.float32)
        d = np.zeros((2, 3), np.float32)
        o[0, ax], d[0, ax] = -100, 1
        o[1, ax], d[1, ax] = 100, -1
        ids, r, s, tt = O.intersect(t.pre, t.nodes, t.box, o, d)
        assert ids[0] == 0 and int(r[0]) == 100
        assert ids[1] == 9 and int(r[1]) == 91
        assert np.all(np.abs(s - 1 / 3) < EPS) and np.all(np.abs(tt - 1 / 3) < EPS)
        p = O.Tree(H.plane_triangles(ax, 1000, rng))
        dpar = np.zeros(3, np.float32)
        dpar[(ax + 1) % 3], dpar[(ax + 2) % 3] = 0.6, 0.8
        ids, *_ = O.intersect(p.pre, p.nodes, p.box, np.array([[1, 1, 1], [-1, -1, -1]], np.float32),
                              np.array([dpar, [1, 1, 1]], np.float32))
        assert ids[0] == INV and ids[1] != INV


def test_ray_box_kats():
    """tests/test_intersection.cpp:15-36"""
    b1 = [-1, -1, -1, 1, 1, 1]
    assert O.ray_box([0, 0, 0], [1, 1, 1], b1)[0]
    assert O.ray_box([10, 0, 0], [-1, 0, 0], b1)[0]
    assert O.ray_box([0, 10, 0], [0, -1, 0], b1)[0]
    assert O.ray_box([0, 0, 10], [0, 0, -1], b1)[0]
    assert not O.ray_box([0, 0, 0], [1, 0, 0], [-1, -1, 1, 1, 1, 1])[0]
    assert not O.ray_box([-2, -2, -2], [-1, 0, 0], [-1, -1, 1, 1, 1, 1])[0]
    assert not O.ray_box([-1, 0, 0], [-1, 0, 0], [0, 0, 0, 1, 1, 1])[0]


def test_clipping_kats():
    """tests/test_clipping.cpp:55-100"""
    box = [-1, -1, -1, 1, 1, 1]
    bb, _ = O.clip_triangle(H.tri((0, 0, -10), (10, 0, 10), (-10, 0, 10)), box)
    assert list(bb) == [-1, 0, -1, 1, 0, 1]
    bb, _ = O.clip_triangle(H.tri((-1, -1, 0), (1, 1, 0), (1, -1, 0)), [0, 0, 0, 2, 2, 2])
    assert list(bb) == [0, 0, 0, 1, 1, 0]
    bb, _ = O.clip_triangle(H.tri((-1, -1, -1), (2, 2, 2), (2, -1, 2)), box)
    assert list(bb) == box
    bb, _ = O.clip_triangle(H.tri((0, 0, 0), (0.5, 0.5, 0.5), (0.5, 0, 0.5)), box)
    assert list(bb) == [0, 0, 0, 0.5, 0.5, 0.5]
    rng = np.random.default_rng(1)
    for _ in range(2000):
        t = rng.uniform(-10, 10, 9).astype(np.float32)
        bb, n = O.clip_triangle(t, box)
        if n < 2:
            continue
        assert np.all(-EPS < bb[:3] - np.array(box[:3])) and np.all(-EPS < np.array(box[3:]) - bb[3:])


def test_precompute_properties():
    """tests/test_triangle.cpp:7-25: unit normal perpendicular to the edges;
    plus an independent float64 evaluation of the same quantities."""
    rng = np.random.default_rng(2)
    v = rng.uniform(-10, 10, (200, 9)).astype(np.float32)
    p = O.precompute(v)
    n = p[:, 3:6].astype(np.float64)
    assert np.allclose(np.linalg.norm(n, axis=1), 1, atol=1e-6)
    u, w = p[:, 6:9].astype(np.float64), p[:, 9:12].astype(np.float64)
    assert np.allclose(np.sum(n * u, 1) / np.linalg.norm(u, axis=1), 0, atol=1e-5)
    assert np.allclose(np.sum(n * w, 1) / np.linalg.norm(w, axis=1), 0, atol=1e-5)
    v64 = v.reshape(-1, 3, 3).astype(np.float64)
    u64, w64 = v64[:, 1] - v64[:, 0], v64[:, 2] - v64[:, 0]
    nu, nw = np.linalg.norm(u64, axis=1), np.linalg.norm(w64, axis=1)
    # FP32 dot products: error a few ULPs of |u||v| (cancellation allowed)
    assert np.all(np.abs(p[:, 12] - np.sum(u64 * w64, 1)) <= 4e-7 * 3 * nu * nw)
    exact = np.sum(u64 * w64, 1) ** 2 - np.sum(u64 * u64, 1) * np.sum(w64 * w64, 1)
    assert np.all(np.abs(p[:, 15] - exact) <= 4e-6 * (nu * nw) ** 2)


def test_intersect_equals_brute_force():
    """Closest hit of the kd-tree traversal equals the minimum over all
    triangles on random rays (the reference's misses of grazing boundary rays
    are checked separately by the GPU tests against this same oracle)."""
    from turner_amd import scenes
    rng = np.random.default_rng(4)
    for v in (scenes.intersect_scene(level=3)["verts"], H.random_soup(500, rng), H.grid_mesh(10)):
        t = O.Tree(v)
        lo, hi = v.reshape(-1, 3).min(0), v.reshape(-1, 3).max(0)
        o = rng.uniform(lo - 1, hi + 1, (20000, 3)).astype(np.float32)
        d = rng.normal(size=(20000, 3)).astype(np.float32)
        a = O.intersect(t.pre, t.nodes, t.box, o, d)
        b = O.intersect_brute(t.pre, o, d)
        assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
        same_r_diff_id = (a[0] != b[0])
        assert np.all(a[1][same_r_diff_id] == b[1][same_r_diff_id])  # only exact ties


def test_mt64_matches_leaf_distance():
    rng = np.random.default_rng(5)
    v = H.random_soup(300, rng, spread=3)
    t = O.Tree(v)
    o = rng.uniform(-4, 4, (5000, 3)).astype(np.float32)
    d = rng.normal(size=(5000, 3)).astype(np.float32)
    ids, r, s, tt = O.intersect(t.pre, t.nodes, t.box, o, d)
    hit = ids != INV
    r64, b1, b2 = O.mt64(v, ids, o, d)
    assert np.median(np.abs(r[hit] - r64[hit]) / r64[hit]) < 1e-6
    assert np.median(np.abs(s[hit] - b1[hit])) < 1e-5


def test_xorshift64star_closed_form():
    """lib/xorshift.h:36-41 (Vigna 2016) evaluated with Python integers."""
    M = (1 << 64) - 1
    s, ref = 42, []
    for _ in range(50):
        s ^= s >> 12
        s ^= (s << 25) & M
        s ^= s >> 27
        ref.append((s * 2685821657736338717) & M)
    assert O.xorshift64s(42, 50) == ref
    fl = O.xorshift64s_floats(42, 50)
    assert np.array_equal(fl, np.array([(x & 0xFFFFFF) * 2.0 ** -24 for x in ref], np.float32))
    assert np.all((fl >= 0) & (fl < 1))


def test_std_sort_restatement(tmp_path):
    """The oracle's restatement of libstdc++ std::sort yields the same order
    of equal keys as std::sort itself (compiled here with g++)."""
    exe = tmp_path / "stdsort"
    src = os.path.join(os.path.dirname(__file__), "golden", "stdsort_ref.cpp")
    subprocess.run(["g++", "-O2", "-std=c++14", src, "-o", str(exe)], check=True)
    rng = np.random.default_rng(6)
    for n in (0, 1, 5, 16, 17, 33, 100, 1000, 20000):
        for nkeys in (1, 3, max(1, n // 4), max(1, n)):
            hi = rng.integers(0, nkeys, n).astype(np.uint64)
            keys = (hi << np.uint64(32)) | np.arange(n, dtype=np.uint64)
            out = subprocess.run([str(exe)], input=keys.tobytes(), stdout=subprocess.PIPE,
                                 check=True).stdout
            expect = np.frombuffer(out, np.uint64)
            assert np.array_equal(O.introsort_u64(keys), expect), (n, nkeys)
    # adversarial-ish inputs that drive the depth limit toward the heap sort
    for pattern in (np.arange(4096)[::-1], np.tile(np.arange(8), 512), np.zeros(3000)):
        keys = (pattern.astype(np.uint64) << np.uint64(32)) | np.arange(len(pattern), dtype=np.uint64)
        out = subprocess.run([str(exe)], input=keys.tobytes(), stdout=subprocess.PIPE,
                             check=True).stdout
        assert np.array_equal(O.introsort_u64(keys), np.frombuffer(out, np.uint64))


def test_effects_kats():
    """tests/test_effects.cpp"""
    assert O.exposure(0.5, 0) == 0
    assert O.exposure(0.5, 1) < 0.5
    assert 0.5 < O.exposure(0.5, 10)
    assert 0.5 < O.gamma(0.5, 1 / 2.2)


def test_ppm_kat():
    """tests/test_raster.cpp:44-57"""
    img = np.zeros((2, 2, 4), np.float32)
    img[1, 0] = 1
    assert O.ppm(img) == "P3\n2 2\n255\n  0   0   0   0   0   0\n255 255 255   0   0   0"


def test_camera_kat():
    """lib/types.h:119-123: the image centre looks along trafo * (0, 0, -1)."""
    from turner_amd import scenes
    cam = scenes.cornell_box()["camera"]
    assert list(O.raster2cam(cam, 160, 160, 320, 320)) == [0, 0, 1]
    c = O.raster2cam(cam, 0, 0, 320, 320)
    assert np.allclose(c, [cam["delta_x"], cam["delta_y"], 1])


def test_reference_furnace_closed_form():
    """furnace_test.blend: escaping rays return the background c, a convex
    sphere of albedo rho shows rho * c (uniform hemisphere, 2 cos weight)."""
    from turner_amd import scenes
    sc = scenes.reference_furnace(level=2, rho=0.8)
    t = O.Tree(sc["verts"])
    img, sq, nr, ns = O.render(sc, t, 16, 16, 1, 4, 256, bg=(0.6, 0.6, 0.6, 1), want_sq=True)
    centre = img[6:10, 6:10, :3]
    sem = np.sqrt((sq[6:10, 6:10, :3] / 256 - centre.astype(np.float64) ** 2).mean() / (256 * centre.size))
    assert abs(centre.mean() - 0.48) < 4 * sem + 1e-4
    assert np.allclose(img[0, 0, :3], 0.6)


def test_extended_furnace_closed_form():
    from turner_amd import scenes
    sc = scenes.furnace(level=2, rho=0.8)
    t = O.Tree(sc["verts"])
    img, sq, nr, _ = O.render(sc, t, 16, 16, 2, 1, 256, flags=O.FLAG_COSINE | O.FLAG_AREA_LIGHTS,
                              want_sq=True)
    assert nr == 16 * 16 * 256 * 3  # closed scene: every segment hits
    centre = img[6:10, 6:10, :3]
    sem = np.sqrt((sq[6:10, 6:10, :3] / 256 - centre.astype(np.float64) ** 2).mean() / (256 * centre.size))
    assert abs(centre.mean() - 0.8) < 4 * sem + 1e-4
    assert np.allclose(img[0, 0, :3], 1.0)


def test_cosine_and_uniform_agree():
    """Both hemisphere estimators are unbiased for the same integral."""
    from turner_amd import scenes
    sc = scenes.sphere_in_cornell(level=2)
    t = O.Tree(sc["verts"])
    a, sa, _, _ = O.render(sc, t, 8, 8, 2, 1, 4096, flags=O.FLAG_COSINE | O.FLAG_AREA_LIGHTS,
                           seed=1, want_sq=True)
    b, sb, _, _ = O.render(sc, t, 8, 8, 2, 1, 4096, flags=O.FLAG_AREA_LIGHTS, seed=2, want_sq=True)
    va = np.maximum(sa / 4096 - a.astype(np.float64) ** 2, 0) / 4096
    vb = np.maximum(sb / 4096 - b.astype(np.float64) ** 2, 0) / 4096
    z = (a - b)[..., :3] / np.sqrt(va + vb + 1e-30)[..., :3]
    assert np.mean(np.abs(z) > 3) < 0.03 and abs(z.mean()) < 0.5


def test_reference_cornell_ray_counts():
    """The reference renders cornell_box.blend at 320x320, -d 3 -m 1 -p 8 with
    819,200 primary rays and 2,632,399 trace() calls (README.md:30-36), 3.21
    per primary.  The procedural box (same dimensions, camera and light) is
    within a few percent of that ratio."""
    from turner_amd import scenes
    sc = scenes.cornell_box()
    t = O.Tree(sc["verts"])
    img, _, nr, ns = O.render(sc, t, 80, 80, 3, 1, 8, flags=O.FLAG_REF_RNG, nthreads=1)
    ratio = nr / (80 * 80 * 8)
    print("segments per primary", ratio)
    assert 2.9 < ratio < 3.5


# ------------------------------------------------------------ raycaster / raytracer
def _plugin(sc, renderer, w=24, depth=3, bg=(0.1, 0.2, 0.3, 1.0), **kw):
    t = O.Tree(sc["verts"])
    img, _, nr, ns = O.render(sc, t, w, w, depth, 1, 1, flags=O.FLAG_REF_RNG, bg=bg, nthreads=1,
                              renderer=renderer, **kw)
    return img, nr, ns


def _primary_dirs(cam, w):
    """main.cpp:201-209 with one pixel sample: per row gen(42), (dx, dy) per pixel."""
    u = O.xorshift64s_floats(42, 2 * w)
    return np.array([[O.raster2cam(cam, x + u[2 * x], y + u[2 * x + 1], w, w) for x in range(w)]
                     for y in range(w)], np.float64)


def _lambert(p, n, albedo, light):
    lp, lc = np.asarray(light[0], np.float64), np.asarray(light[1], np.float64)
    l = lp - p
    l /= np.linalg.norm(l, axis=-1, keepdims=True)
    ca = np.maximum(0.0, (l * n).sum(-1))
    return ca[..., None] * np.asarray(albedo, np.float64) * lc


def test_raycaster_closed_form():
    """raycaster.cpp:7-24: every ray hits the wall at r = 3 exactly, so each
    pixel is the diffuse colour with alpha clamp(1 - 3 / max_visibility)."""
    sc = H.wall_scene()
    for mv in (5.0, 2.0, 3.0):
        img, nr, ns = _plugin(sc, O.RENDERER_RAYCASTER, max_visibility=mv)
        a = np.float32(1) - np.float32(3) / np.float32(mv)
        a = min(max(a, np.float32(0)), np.float32(1))
        assert np.array_equal(img[..., :3], np.broadcast_to(np.float32([0.6, 0.5, 0.4]), img[..., :3].shape))
        assert np.all(img[..., 3] == a)
        assert nr == 24 * 24 and ns == 0  # num_rays counts hits only


def test_raycaster_background_not_counted():
    sc = H.wall_scene(z=-3.0)  # behind the camera: every ray misses
    img, nr, _ = _plugin(sc, O.RENDERER_RAYCASTER)
    assert nr == 0 and np.all(img == np.float32([0.1, 0.2, 0.3, 1.0]))


def test_raytracer_lambert_closed_form():
    """raytracer.cpp:24-34: Lambert term of the point light (lib/lambertian.h:15-19), lit."""
    sc = H.wall_scene()
    img, nr, ns = _plugin(sc, O.RENDERER_RAYTRACER)
    d = _primary_dirs(sc["camera"], 24)
    p = 3.0 * d
    want = _lambert(p, np.array([0, 0, -1.0]), (0.6, 0.5, 0.4, 1.0), sc["lights"][0])
    assert np.allclose(img, want, rtol=2e-6, atol=1e-7)
    assert nr == 24 * 24 and ns == 24 * 24


def test_raytracer_shadow_closed_form():
    """raytracer.cpp:55-64: a wall point whose segment to the light crosses the
    occluder (at z = 2, halfway to the light) is darkened by c -= c * 0.5."""
    sc = H.wall_scene(occluder=True)
    img, _, _ = _plugin(sc, O.RENDERER_RAYTRACER, w=32, shadow_intensity=0.5)
    d = _primary_dirs(sc["camera"], 32)
    light = sc["lights"][0]
    lo, hi, m = 0.3, 0.9, 1e-3
    q = 2.0 * d[..., :2]  # primary ray at the occluder's plane
    on_occ = np.all((q > lo + m) & (q < hi - m), -1)
    off_occ = np.any((q < lo - m) | (q > hi + m), -1)
    p = 3.0 * d
    c = 0.5 * (p[..., :2] + np.asarray(light[0][:2]))  # segment to the light at z = 2
    shadowed = off_occ & np.all((c > lo + m) & (c < hi - m), -1)
    lit = off_occ & np.any((c < lo - m) | (c > hi + m), -1)
    wall = _lambert(p, np.array([0, 0, -1.0]), (0.6, 0.5, 0.4, 1.0), light)
    occ = _lambert(2.0 * d, np.array([0, 0, -1.0]), (0.2, 0.7, 0.3, 1.0), light)
    assert shadowed.sum() > 20 and lit.sum() > 20 and on_occ.sum() > 20
    assert np.allclose(img[lit], wall[lit], rtol=2e-6, atol=1e-7)
    assert np.allclose(img[shadowed], 0.5 * wall[shadowed], rtol=2e-6, atol=1e-7)
    assert np.allclose(img[on_occ], occ[on_occ], rtol=2e-6, atol=1e-7)


def test_raytracer_reflection_closed_form():
    """raytracer.cpp:42-52: the mirror ray d - 2 (n.d) n leaves the scene, so a
    wall of reflectivity 0.5 shows 0.5 direct + 0.5 reflective * background;
    at max_depth 0 the mirror call returns {} (and is still counted)."""
    sc = H.wall_scene(reflectivity=0.5)
    bg = (0.1, 0.2, 0.3, 1.0)
    d = _primary_dirs(sc["camera"], 24)
    direct = _lambert(3.0 * d, np.array([0, 0, -1.0]), (0.6, 0.5, 0.4, 1.0), sc["lights"][0])
    img, nr, ns = _plugin(sc, O.RENDERER_RAYTRACER, depth=1, bg=bg)
    want = 0.5 * direct + 0.5 * np.array([0.9, 0.8, 0.7, 1.0]) * np.array(bg)
    assert np.allclose(img, want, rtol=2e-6, atol=1e-7)
    assert nr == 2 * 24 * 24 and ns == 24 * 24  # the mirror rays miss before their shadow test
    img0, nr0, _ = _plugin(sc, O.RENDERER_RAYTRACER, depth=0, bg=bg)
    assert np.allclose(img0, 0.5 * direct, rtol=2e-6, atol=1e-7)
    assert nr0 == 2 * 24 * 24

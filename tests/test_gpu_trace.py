"""trace() on the GPU (tn_trace, the batched twin of trace.h:23-25) and the C++
host interface (turner_amd/host/turner.hpp) called the way the reference
calls it, checked against closed forms and the oracle:

* the point-light closed forms, the shadow predicate and the depth-1
  background case of tests/test_oracle_pathtracer.py, through tn_trace;
* tn_trace entering at depth >= 1 against the oracle's trace() on the same
  rays (per-ray means over repeated rays, z-tests);
* tests/cpp/host_api_check: KDTreeIntersection::intersect one ray at a time and
  batched (bit-exact against the oracle's own tree and precompute), trace()
  one ray at a time and batched, and the pixel loop of main.cpp:193-216 that
  calls trace() per pixel sample.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import turner_amd as T
from turner_amd import scenes
from tests import helpers as H
from tests.test_oracle_pathtracer import (LC, RHO, point_light_closed_form, wall, wall_rays)

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tests", "cpp", "host_api_check")
FLT_MAX = np.finfo(np.float32).max


def gpu_tree(sc):
    return T.KDTree(sc["verts"], sc["normals"], sc["albedo"], sc["emission"])


def conf(max_depth, mc, flags=0, seed=1, bg=(0, 0, 0, 1)):
    return T.TracerConfig(max_recursion_depth=max(max_depth, 1), num_monte_carlo_samples=mc, flags=flags,
                          seed=seed, bg_color=bg)


def trace(sc, tree, o, d, depth, max_depth, mc, **kw):
    c = conf(max_depth, mc, **kw)
    c.max_recursion_depth = max_depth
    return T.trace(tree, o, d, sc["lights"], depth, c)


def test_point_light_closed_form(gpu):
    sc = wall()
    tree = gpu_tree(sc)
    o, d = wall_rays(4096, 3)
    got, st = trace(sc, tree, o, d, 0, 0, 1)
    assert st["rays"] == len(o) and st["shadow_rays"] == len(o)
    np.testing.assert_allclose(got, point_light_closed_form(o, d, sc["lights"][0][0]), rtol=1e-6, atol=1e-9)
    back = wall(light=(0.5, 0.5, 4.0))
    got, _ = trace(back, gpu_tree(back), o, d, 0, 0, 1)
    assert np.all(got == 0.0)


@pytest.mark.parametrize("blocker_z,lit", [(None, True), (1.5, False), (1.0, False), (1.0 - 2.0**-15, True),
                                           (0.5, True)])
def test_shadow_predicate(gpu, blocker_z, lit):
    full = (np.float32(RHO) * (np.float32(1.0) * np.float32(LC))) * np.float32(1 / np.pi)
    for x, y in ((0.5, 0.5), (0.75, 1.0), (1.0, 0.3)):
        sc = wall(blocker_z=blocker_z, light=(x, y, 1.0))
        got, _ = trace(sc, gpu_tree(sc), np.array([[x, y, 2.0]], np.float32), np.array([[0, 0, 1.0]], np.float32),
                       0, 0, 1)
        if lit:
            np.testing.assert_allclose(got[0], full, rtol=1e-6)
        else:
            assert np.all(got[0] == 0.0), (x, y, got[0])


def test_depth1_background(gpu):
    sc = wall(lights=False)
    tree = gpu_tree(sc)
    o, d = wall_rays(2048, 5)
    bg = (0.25, 0.5, 0.75, 1.0)
    want = np.float32(RHO) * np.float32(bg)
    got, st = trace(sc, tree, o, d, 0, 1, 1, flags=T._lib.TN_RENDER_COSINE, bg=bg)
    assert np.all(got == want) and st["rays"] == 2 * len(o)
    got, _ = trace(sc, tree, o, d, 0, 1, 4, flags=T._lib.TN_RENDER_COSINE, bg=bg)
    np.testing.assert_allclose(got, np.broadcast_to(want, got.shape), rtol=1e-6)
    # uniform hemisphere (the reference's): rho c in expectation, variance (rho c)^2 / (3 m)
    n, mc = 20000, 2
    o, d = wall_rays(n, 6)
    got, _ = trace(sc, tree, o, d, 0, 1, mc, bg=bg)
    sigma = np.asarray(want, np.float64) / np.sqrt(3 * mc * n)
    assert np.all(np.abs(got.mean(0) - want) < 4 * sigma + 1e-9)


def cornell_rays(k, reps, seed=0):
    """k camera rays through random points of the Cornell image, each repeated reps times."""
    sc = scenes.cornell_box()
    cam = sc["camera"]
    rng = np.random.default_rng(seed)
    px, py = rng.uniform(0, 1, k), rng.uniform(0, 1, k)
    cx = -cam["delta_x"] * (1 - 2 * px)
    cy = cam["delta_y"] * (1 - 2 * py)
    d = (np.stack([cx, cy, -np.ones_like(cx)], -1) @ np.asarray(cam["trafo"], np.float64).T).astype(np.float32)
    o = np.tile(np.asarray(cam["pos"], np.float32), (k, 1))
    return sc, np.repeat(o, reps, 0), np.repeat(d, reps, 0)


def zcheck(a, b, reps, zmax=5.0, what=""):
    """Per-ray means of two samples of reps repetitions each: |z| < zmax per
    ray and channel (RGB), and the mean over all rays within 4 sigma."""
    a, b = a.reshape(-1, reps, 4)[..., :3].astype(np.float64), b.reshape(-1, reps, 4)[..., :3].astype(np.float64)
    ma, mb = a.mean(1), b.mean(1)
    se = np.sqrt(a.var(1, ddof=1) / reps + b.var(1, ddof=1) / reps)
    z = np.where(se > 0, (ma - mb) / np.maximum(se, 1e-30), 0.0)
    assert np.all(np.abs(z) < zmax), f"{what}: max |z| {np.abs(z).max():.2f}"
    both = np.abs(ma - mb) <= 1e-6 * np.maximum(np.abs(ma), np.abs(mb))
    assert np.all((se > 0) | both), what
    za = (ma.mean(0) - mb.mean(0)) / np.sqrt((a.mean(1).var(0) + b.mean(1).var(0)) / ma.shape[0] + 1e-30)
    assert np.all(np.abs(za) < 4.0), f"{what}: image-mean z {za}"


@pytest.mark.parametrize("depth,max_depth,mc,flags", [(1, 3, 1, 0), (2, 3, 2, 0), (1, 2, 1, 1), (0, 2, 2, 0)])
def test_trace_entering_at_depth_matches_oracle(gpu, depth, max_depth, mc, flags):
    """trace() entered at level `depth` (pathtracer.cpp:14-102: levels depth..max_depth
    remain) against the oracle's restatement on the same rays."""
    k, reps = 96, 256
    sc, o, d = cornell_rays(k, reps)
    tree = gpu_tree(sc)
    otree = O.Tree(sc["verts"])
    got, st = trace(sc, tree, o, d, depth, max_depth, mc, flags=flags)
    want, nr, ns = O.trace_rays(sc, otree, o, d, depth=depth, max_depth=max_depth, mc=mc, flags=flags, seed=7)
    zcheck(got, want, reps, what=f"depth {depth}/{max_depth} m {mc}")
    # the box is open on the camera side, so the segment counts are random: a
    # path adds between 1 and K = 1 + m + ... + m^(levels-1) segments, which
    # bounds the spread of the count by sqrt(n) K / 2
    levels = max_depth - depth + 1
    kmax = sum(mc**j for j in range(levels))
    n = len(o)
    assert n <= st["rays"] <= n * kmax and n <= nr <= n * kmax
    assert abs(st["rays"] - nr) <= 5 * np.sqrt(2 * n) * kmax / 2


# ---------------------------------------------------------------- C++ caller


def write_rays(path, o, d):
    with open(path, "wb") as f:
        f.write(np.array([len(o)], np.uint32).tobytes())
        f.write(np.ascontiguousarray(o, np.float32).tobytes())
        f.write(np.ascontiguousarray(d, np.float32).tobytes())


def run_check(*args):
    assert os.path.exists(CHECK), "tests/cpp/host_api_check is not built (__graft_entry__.build())"
    p = subprocess.run([CHECK, *map(str, args)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300)
    assert p.returncode == 0, p.stdout


@pytest.mark.parametrize("scene_flags", [0, 4])
def test_cpp_intersect_matches_oracle(gpu, tmp_path, scene_flags):
    """KDTreeIntersection::intersect from C++ (one ray per call and batched)
    against the oracle's tree, precompute and traversal, bit for bit; with
    scene flags 4 (TN_SCENE_REFINED) the calls walk the refined tree."""
    sc = scenes.intersect_scene(level=4)
    nt = len(sc["verts"])
    sc.update(camera=scenes.camera((0, 0, 0), 60.0), albedo=np.tile(np.float32([0.8, 0.8, 0.8, 1]), (nt, 1)),
              emission=np.zeros((nt, 4), np.float32))
    scenes.save_tns(sc, tmp_path / "s.tns")
    o, d = scenes.random_rays(20000, seed=9)
    ho, hd = H.hard_rays(sc["verts"], 4000, np.random.default_rng(9))
    # rays that miss the root box and rays that enter it and miss every triangle
    mo = np.array([[5, 5, 5], [0, 0, 0], [0.0, 0.0, 0.0]], np.float32)
    md = np.array([[1, 0, 0], [0, 0, 0], [1e-9, 0, 0]], np.float32)
    o, d = np.concatenate([mo, o, ho]), np.concatenate([md, d, hd])
    write_rays(tmp_path / "r.bin", o, d)
    singles = 512
    run_check("intersect", tmp_path / "s.tns", tmp_path / "r.bin", tmp_path / "o.bin", singles, scene_flags)
    raw = np.fromfile(tmp_path / "o.bin", np.uint32)
    n = len(o)
    ids, r, s, t = raw[:n], raw[n:2 * n].view(np.float32), raw[2 * n:3 * n].view(np.float32), \
        raw[3 * n:4 * n].view(np.float32)
    ot = O.Tree(sc["verts"])
    ref = O.intersect(ot.pre, ot.nodes, ot.box, o, d)
    assert np.array_equal(ids, ref[0])
    for x, y in zip((r, s, t), ref[1:]):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    one = raw[4 * n:].reshape(singles, 4)
    sid, sr, sa, sb = one[:, 0], one[:, 1].view(np.float32), one[:, 2].view(np.float32), one[:, 3].view(np.float32)
    assert np.array_equal(sid, ids[:singles])
    hit = sid != T.TN_INVALID_ID
    assert np.array_equal(sr[hit], r[:singles][hit]) and np.array_equal(sa[hit], s[:singles][hit])
    assert np.array_equal(sb[hit], t[:singles][hit])
    assert np.all(np.isnan(sa[~hit])) and np.all(np.isnan(sb[~hit]))
    # on a miss r is FLT_MAX iff the ray entered the root box (lib/kdtree.cpp:535)
    enters = np.array([O.ray_box(o[i], np.where(d[i] == 0, np.float32(1e-5), d[i]), ot.box)[0]
                       for i in np.nonzero(~hit)[0]])
    assert np.all(sr[~hit][enters] == FLT_MAX) and np.all(np.isnan(sr[~hit][~enters]))
    assert (~hit).sum() >= 2 and enters.any() and (~enters).any()


def test_cpp_trace_matches_closed_form_and_oracle(gpu, tmp_path):
    """trace() from C++: one ray per call (as main.cpp:212-214 calls it) and
    batched, against the point-light closed form at depth 0 and the oracle
    at depth >= 1; one-ray calls for the same ray draw fresh samples."""
    sc = wall()
    scenes.save_tns(sc, tmp_path / "w.tns")
    o, d = wall_rays(1024, 3)
    write_rays(tmp_path / "w.bin", o, d)
    run_check("trace", tmp_path / "w.tns", tmp_path / "w.bin", tmp_path / "w.out", 64, 0, 0, 1, 0, 1)
    out = np.fromfile(tmp_path / "w.out", np.float32).reshape(-1, 4)
    want = point_light_closed_form(o, d, sc["lights"][0][0])
    np.testing.assert_allclose(out[:1024], want, rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(out[1024:], want[:64], rtol=1e-6, atol=1e-9)

    k, reps = 64, 256
    cb, co, cd = cornell_rays(k, reps, seed=4)
    scenes.save_tns(cb, tmp_path / "c.tns")
    write_rays(tmp_path / "c.bin", co, cd)
    run_check("trace", tmp_path / "c.tns", tmp_path / "c.bin", tmp_path / "c.out", reps, 1, 3, 1, 0, 5)
    out = np.fromfile(tmp_path / "c.out", np.float32).reshape(-1, 4)
    want, _, _ = O.trace_rays(cb, O.Tree(cb["verts"]), co, cd, depth=1, max_depth=3, mc=1, seed=3)
    zcheck(out[:len(co)], want, reps, what="C++ batched trace")
    singles = out[len(co):]  # reps one-ray calls of the first camera ray
    assert np.unique(singles[:, 0]).size > reps // 2, "one-ray calls repeat the same samples"
    zcheck(np.concatenate([singles, out[:len(co)][reps:]]), want, reps, what="C++ one-ray trace")


def test_cpp_pixel_loop_matches_oracle(gpu, tmp_path):
    """The pixel loop of main.cpp:193-216 calling trace() per pixel sample with
    the reference jitter: at max_depth 0 (no random bounce) equal to the
    oracle's render within 1e-5; at max_depth 2 the image mean within 4 sigma
    of the oracle's renders with the same jitter and 32 bounce seeds."""
    sc = scenes.cornell_box()
    scenes.save_tns(sc, tmp_path / "c.tns")
    otree = O.Tree(sc["verts"])
    w = h = 12
    run_check("pixels", tmp_path / "c.tns", tmp_path / "p0.out", w, h, 2, 0, 1, 0)
    got = np.fromfile(tmp_path / "p0.out", np.float32).reshape(h, w, 4)
    ref, _, _, _ = O.render(sc, otree, w, h, 0, 1, 2, flags=O.FLAG_REF_RNG)
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-6)
    spp, runs = 8, 32
    run_check("pixels", tmp_path / "c.tns", tmp_path / "p2.out", w, h, spp, 2, 1, 0)
    got = np.fromfile(tmp_path / "p2.out", np.float32).reshape(h, w, 4)
    refs, sqs = [], []
    for seed in range(runs):
        ref, sq, _, _ = O.render(sc, otree, w, h, 2, 1, spp, flags=O.FLAG_REF_RNG, seed=seed + 1, want_sq=True)
        refs.append(ref)
        sqs.append(sq)
    ref = np.mean(refs, 0)
    var = np.maximum(np.mean(sqs, 0) / spp - ref**2, 0)[..., :3]  # per-sample variance per pixel
    se = np.sqrt(var.sum((0, 1)) / spp * (1 + 1 / runs)) / (w * h)
    z = (got[..., :3].mean((0, 1)) - ref[..., :3].mean((0, 1))) / se
    assert np.all(np.abs(z) < 4.0), z

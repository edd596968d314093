"""GPU parity of the batched intersection (tn_intersect / tn_occluded) against
the oracle's restatement of KDTreeIntersection::intersect (lib/kdtree.cpp:515-607).

Bar: hit ids, r, s, t bit-exact against the oracle's reference-semantics
traversal (exhaustive, no early exit) for the default and the exhaustive GPU
traversal; r within 1e-5 relative of an FP64 Moeller-Trumbore evaluation.
"""
import numpy as np
import pytest

import oracle as O
import turner_amd as T
from turner_amd import scenes
from turner_amd._lib import (TN_ISECT_COUNT, TN_ISECT_EXHAUSTIVE, TN_ISECT_LEAF, TN_ISECT_MT,
                             TN_ISECT_TICK)
from tests import helpers as H

pytestmark = pytest.mark.gpu
INV = T.TN_INVALID_ID


def gpu_hits(tree, o, d, flags=0):
    return T.KDTreeIntersection(tree).intersect_batch(o, d, flags)


def oracle_hits(tree, o, d):
    return O.intersect(tree.triangle_records(), tree.flat_nodes(), tree.box(), o, d)


def assert_bit_exact(g, ref, what=""):
    gid, gr, gs, gt = g
    oid, orr, os_, ot = ref
    bad = np.nonzero(gid != oid)[0]
    assert len(bad) == 0, f"{what}: {len(bad)} id mismatches, first {bad[:5]}: gpu {gid[bad[:5]]} oracle {oid[bad[:5]]}"
    hit = oid != INV
    for name, a, b in (("r", gr, orr), ("s", gs, os_), ("t", gt, ot)):
        diff = np.nonzero(a[hit].view(np.uint32) != b[hit].view(np.uint32))[0]
        assert len(diff) == 0, f"{what}: {name} differs on {len(diff)} hits"


def test_four_separated_kat(gpu):
    """tests/test_kdtree.cpp:23-63"""
    v = H.four_separated()
    tree = T.KDTree(v)
    assert tree.height() == 1 and tree.num_nodes() == 5
    o = np.zeros((4, 3), np.float32)
    d = np.array([[0.5, 0.5, 1], [2.5, 0.5, 1], [0.5, 2.5, 1], [2.5, 2.5, 1]], np.float32)
    ids, r, s, t = gpu_hits(tree, o, d)
    assert list(ids) == [0, 1, 2, 3]
    assert list(r) == [1, 1, 1, 1]
    assert list(s) == [0.5, 0, 0, 0]
    assert list(t) == [0.5, 0.5, 0.5, 0.5]


def test_cube_and_coplanar_kats(gpu):
    """tests/test_kdtree.cpp:162-186, 188-245"""
    tree = T.KDTree(H.cube12())
    assert tree.height() == 0 and tree.num_nodes() == 7
    rng = np.random.default_rng(3)
    for ax in range(3):
        tree = T.KDTree(H.coplanar_stack(ax, rng))
        o = np.zeros((2, 3), np.float32)
        d = np.zeros((2, 3), np.float32)
        o[0, ax], d[0, ax] = -100, 1
        o[1, ax], d[1, ax] = 100, -1
        ids, r, s, t = gpu_hits(tree, o, d)
        assert ids[0] == 0 and int(r[0]) == 100
        assert ids[1] == 9 and int(r[1]) == 91
        assert np.all(np.abs(s - 1 / 3) < 1e-5) and np.all(np.abs(t - 1 / 3) < 1e-5)
        # all triangles in one plane: parallel ray misses, ray through zero hits
        ptree = T.KDTree(H.plane_triangles(ax, 1000, rng))
        dpar = np.zeros(3, np.float32)
        dpar[(ax + 1) % 3], dpar[(ax + 2) % 3] = 0.6, 0.8
        ids, *_ = gpu_hits(ptree, np.array([[1, 1, 1], [-1, -1, -1]], np.float32),
                           np.array([dpar, [1, 1, 1]], np.float32))
        assert ids[0] == INV and ids[1] != INV


@pytest.mark.parametrize("level", [3, 5])
def test_icosphere_bit_exact(gpu, level):
    sc = scenes.intersect_scene(level=level)
    tree = T.KDTree(sc["verts"])
    rng = np.random.default_rng(level)
    o, d = scenes.random_rays((1 << 17) + 12345, seed=level)  # not a multiple of the chunk
    ho, hd = H.hard_rays(sc["verts"], 1 << 15, rng)
    o = np.concatenate([o, ho])
    d = np.concatenate([d, hd])
    print(f"[icosphere level {level}] {tree.num_triangles()} tris, {len(o)} rays", flush=True)
    ref = oracle_hits(tree, o, d)
    for sched, name in ((0, "cooperative"), (TN_ISECT_LEAF, "leaf"), (TN_ISECT_TICK, "tick")):
        assert_bit_exact(gpu_hits(tree, o, d, sched), ref, name + " early-exit")
        assert_bit_exact(gpu_hits(tree, o, d, sched | TN_ISECT_EXHAUSTIVE), ref, name + " exhaustive")


def test_large_leaves_bit_exact(gpu):
    """The reference tree of the occluded-light scene at level 7 keeps
    clipped-away triangles in every later split: its leaves hold 28
    references on average and up to 75, so the pooled pairs of one leaf can
    fill more than a 64-pair round and most rounds continue a leaf of the
    round before (owner marker at the round's first pair).  Random rays and
    a bundle from one point, bit-exact against the oracle, cooperative and
    per-lane schedules."""
    sc = scenes.occluded_light_cornell(level=7)
    tree = T.KDTree(sc["verts"])
    lo, hi = np.asarray(tree.box()[:3]), np.asarray(tree.box()[3:])
    rng = np.random.default_rng(17)
    n = 1 << 14
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    # a bundle from one point: neighbouring lanes share leaves
    o[: n // 2] = (0.5 * (lo + hi)).astype(np.float32)
    print(f"[large leaves] {tree.num_triangles()} tris, {n} rays", flush=True)
    ref = oracle_hits(tree, o, d)
    for sched in (0, TN_ISECT_LEAF):
        assert_bit_exact(gpu_hits(tree, o, d, sched), ref, f"large leaves sched {sched}")
        assert_bit_exact(gpu_hits(tree, o, d, sched | TN_ISECT_EXHAUSTIVE), ref,
                         f"large leaves sched {sched} exhaustive")


def test_soup_and_grid_ties(gpu):
    rng = np.random.default_rng(11)
    soup = H.random_soup(3000, rng)
    grid = H.grid_mesh(24)
    for v in (soup, grid):
        tree = T.KDTree(v)
        o, d = H.hard_rays(v, 1 << 15, rng)
        # rays straight down onto grid vertices and edges: exact ties
        if v is grid:
            k = rng.integers(0, 25, (4096, 2)).astype(np.float32)
            k[:2048, 0] += 0.5
            o = np.concatenate([o, np.c_[k, np.full(4096, 5, np.float32)]])
            d = np.concatenate([d, np.tile(np.array([[0, 0, -1]], np.float32), (4096, 1))])
        ref = oracle_hits(tree, o, d)
        for sched in (0, TN_ISECT_LEAF, TN_ISECT_TICK):
            assert_bit_exact(gpu_hits(tree, o, d, sched), ref, f"ties sched {sched}")


def _cos_incidence(verts, ids, d):
    v = verts.reshape(-1, 3, 3).astype(np.float64)[ids]
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    dd = d.astype(np.float64) / np.linalg.norm(d, axis=1, keepdims=True)
    return np.abs(np.sum(n * dd, axis=1))


def test_distance_vs_fp64(gpu):
    """r against FP64 Moeller-Trumbore.  Bar: |r - r64| <= 1e-5 * r +
    1e-7 * extent / |cos| for every hit (1e-5 relative, plus a few ULPs of
    the scene coordinates, which the plane-based distance of
    lib/intersection.h:40-49 amplifies by 1/|cos| at grazing incidence and
    which dominate for rays starting next to a triangle), and >= 99.99% of
    hits within 1e-5 relative.  The FP32 Moeller-Trumbore mode meets the same
    bar and agrees on every id whose FP64 barycentric margin exceeds 1e-4."""
    sc = scenes.intersect_scene(level=5)
    tree = T.KDTree(sc["verts"])
    extent = float(np.max(tree.box()[3:] - tree.box()[:3]))
    o, d = scenes.random_rays(1 << 16, seed=99)
    for flags in (0, TN_ISECT_MT):
        ids, r, s, t = gpu_hits(tree, o, d, flags)
        hit = ids != INV
        r64, b1, b2 = O.mt64(sc["verts"], ids, o, d)
        err = np.abs(r[hit] - r64[hit])
        rel = err / np.maximum(np.abs(r64[hit]), 1e-30)
        cos = _cos_incidence(sc["verts"], ids[hit], d[hit])
        bound = 1e-5 * np.abs(r64[hit]) + 1e-7 * extent / np.maximum(cos, 1e-6)
        print(f"[fp64 flags={flags}] within 1e-5 relative: {np.mean(rel < 1e-5):.6f}, "
              f"max err/bound {np.max(err / bound):.3g}", flush=True)
        assert np.all(err <= bound)
        assert np.mean(rel < 1e-5) >= 0.9999
        if flags == 0:
            ref_ids = ids
            margin = np.minimum(np.minimum(b1, b2), 1 - b1 - b2)
            clear = hit & (margin > 1e-4)
        else:
            assert np.array_equal(ids[clear], ref_ids[clear])


def test_occluded_matches_reference_predicate(gpu):
    """tn_occluded == !( !hit || r_closest > tmax ) of pathtracer.cpp:49-53."""
    import torch
    sc = scenes.intersect_scene(level=4)
    tree = T.KDTree(sc["verts"])
    o, d = scenes.random_rays(1 << 15, seed=5)
    ref = oracle_hits(tree, o, d)
    rng = np.random.default_rng(5)
    tmax = np.where(rng.random(len(o)) < 0.5, ref[1] * rng.uniform(0.5, 1.5, len(o)),
                    ref[1]).astype(np.float32)
    tmax[ref[0] == INV] = 10.0
    dev = torch.device("cuda:0")
    to, td, tt = (torch.from_numpy(x).to(dev) for x in (o, d, tmax))
    occ = torch.empty(len(o), dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    T.check(T.lib().tn_occluded(tree.handle, T.ptr(to), T.ptr(td), T.ptr(tt), len(o), T.ptr(occ),
                                0, stream))
    torch.cuda.synchronize()
    expect = (ref[0] != INV) & ~(ref[1] > tmax)
    assert np.array_equal(occ.cpu().numpy().astype(bool), expect)


def test_edge_cases(gpu):
    tree = T.KDTree(H.four_separated())
    # empty batch
    out = gpu_hits(tree, np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32))
    assert len(out[0]) == 0
    # single triangle, zero direction components, rays missing the box,
    # origin exactly on the triangle (r == 0 accepted)
    one = T.KDTree(H.tri((0, 0, 1), (0, 1, 1), (1, 0, 1))[None])
    o = np.array([[0.2, 0.2, 0], [5, 5, 5], [0.2, 0.2, 1], [0.2, 0.2, 2]], np.float32)
    d = np.array([[0, 0, 1], [1, 0, 0], [0, 0, 1], [0, 0, 1]], np.float32)
    ids, r, s, t = gpu_hits(one, o, d)
    ref = O.intersect(one.triangle_records(), one.flat_nodes(), one.box(), o, d)
    assert_bit_exact((ids, r, s, t), ref, "edge")
    assert ids[0] == 0 and r[0] == 1 and ids[1] == INV and ids[2] == 0 and r[2] == 0
    assert ids[3] == INV


def test_torch_device_batch_and_counters(gpu):
    import torch
    sc = scenes.intersect_scene(level=4)
    tree = T.KDTree(sc["verts"])
    o, d = scenes.random_rays(1 << 16, seed=8)
    dev = torch.device("cuda:0")
    isect = T.KDTreeIntersection(tree)
    T.check(T.lib().tn_get_counters(tree.handle, np.zeros(8, np.uint64).ctypes.data, 1))
    ids, r, s, t = isect.intersect_batch(torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev),
                                         TN_ISECT_COUNT)
    torch.cuda.synchronize()
    ref = oracle_hits(tree, o, d)
    assert_bit_exact((ids.cpu().numpy().view(np.uint32), r.cpu().numpy(), s.cpu().numpy(),
                      t.cpu().numpy()), ref, "torch")
    cnt = np.zeros(8, np.uint64)
    T.check(T.lib().tn_get_counters(tree.handle, cnt.ctypes.data, 1))
    assert cnt[4] == len(o)
    # the longest traversal lies between the mean and the total leaf count
    assert cnt[1] <= cnt[7] * cnt[4] and cnt[7] <= cnt[1]
    oc = O.intersect_count(tree.triangle_records(), tree.flat_nodes(), tree.box(), o, d)
    # the oracle counts the same near-first early-exit traversal (with an
    # unbounded stack; restarts of the short stack only add revisits)
    if cnt[5] == 0:
        assert (cnt[0], cnt[1], cnt[3]) == (oc["inner"], oc["leaves"], oc["tests"])
    else:
        assert cnt[3] >= oc["tests"] and cnt[1] >= oc["leaves"]


@pytest.mark.slow
def test_full_config_ray_batch(gpu):
    """1,310,732-triangle scene, 2^24 rays: every ray starts inside the closed
    cube, so every ray must hit; a 2^21 subset is bit-exact against the
    oracle, and the early-exit and exhaustive kernels agree on all rays."""
    import torch
    sc = scenes.intersect_scene(level=8)
    tree = T.KDTree(sc["verts"])
    print(f"[full] {tree.num_triangles()} tris, height {tree.height()}", flush=True)
    o, d = scenes.random_rays(1 << 24, seed=2024)
    dev = torch.device("cuda:0")
    to, td = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
    isect = T.KDTreeIntersection(tree)
    a = isect.intersect_batch(to, td)
    b = isect.intersect_batch(to, td, TN_ISECT_EXHAUSTIVE)
    torch.cuda.synchronize()
    for x, y in zip(a, b):
        assert torch.equal(x, y)
    ids = a[0].cpu().numpy().view(np.uint32)
    assert np.all(ids != INV)
    sub = np.sort(np.random.default_rng(0).choice(len(o), 1 << 21, replace=False))
    ref = O.intersect(tree.triangle_records(), tree.flat_nodes(), tree.box(), o[sub], d[sub], nthreads=16)
    got = tuple(x.cpu().numpy()[sub] for x in a)
    got = (got[0].view(np.uint32), got[1], got[2], got[3])
    assert_bit_exact(got, ref, "full subset")


def test_work_slots_cover_every_ray_once(gpu):
    """Batches of more than 256 rays per resident wave take their chunks from
    the scene's work slots: the rays counter equals n (every ray traversed
    exactly once), back-to-back launches agree, 20 launches alternating over
    two streams (more than the 16 slots, so slots are reused behind their
    events) agree, and a subset is bit-exact against the oracle."""
    import torch
    sc = scenes.intersect_scene(level=5)
    tree = T.KDTree(sc["verts"])
    _, cus = T.device_info()
    # above the work-slot threshold (cus x 28 waves x 256 rays); the chunk is
    # 64 rays there, and n is not a multiple of it, so the last chunk is ragged
    n = max(1 << 21, cus * 28 * 256) + 4321
    assert n % 64 != 0
    print(f"[slots] n={n}", flush=True)
    o, d = scenes.random_rays(n, seed=11)
    dev = torch.device("cuda:0")
    to, td = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
    isect = T.KDTreeIntersection(tree)
    T.check(T.lib().tn_get_counters(tree.handle, np.zeros(8, np.uint64).ctypes.data, 1))
    isect.intersect_batch(to, td, TN_ISECT_COUNT)
    torch.cuda.synchronize()
    cnt = np.zeros(8, np.uint64)
    T.check(T.lib().tn_get_counters(tree.handle, cnt.ctypes.data, 1))
    assert cnt[4] == n
    a = isect.intersect_batch(to, td)
    b = isect.intersect_batch(to, td)
    torch.cuda.synchronize()
    for x, y in zip(a, b):
        assert torch.equal(x, y)
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    outs = []
    for k in range(20):
        with torch.cuda.stream(streams[k % 2]):
            outs.append(isect.intersect_batch(to, td))
    torch.cuda.synchronize()
    for res in outs:
        for x, y in zip(res, a):
            assert torch.equal(x, y)
    ids = a[0].cpu().numpy().view(np.uint32)
    assert np.all(ids != INV)
    # any-hit launch of the same size: occluded iff tmax reaches the closest hit
    tmax = a[1].clone()
    tmax[::2] *= 0.5
    occ = torch.empty(n, dtype=torch.uint8, device=dev)
    T.check(T.lib().tn_occluded(tree.handle, T.ptr(to), T.ptr(td), T.ptr(tmax), n, T.ptr(occ), 0,
                                torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    occ = occ.cpu().numpy()
    r_even = a[1].cpu().numpy()[::2]
    assert np.all(occ[1::2] == 1)
    assert np.all(occ[::2][r_even > 0] == 0)
    sub = np.random.default_rng(1).choice(n, 1 << 14, replace=False)
    ref = O.intersect(tree.triangle_records(), tree.flat_nodes(), tree.box(), o[sub], d[sub])
    got = tuple(x.cpu().numpy()[sub] for x in a)
    got = (got[0].view(np.uint32), got[1], got[2], got[3])
    assert_bit_exact(got, ref, "work slots subset")


def test_large_batch_in_captured_graph(gpu):
    """Under stream capture a batch above the work-slot threshold takes the
    static split: the captured launch, replayed three times with eager work
    in between, gives the eager (work-slot) result every time."""
    import torch
    sc = scenes.intersect_scene(level=4)
    tree = T.KDTree(sc["verts"])
    _, cus = T.device_info()
    n = cus * 28 * 256 + 999
    o, d = scenes.random_rays(n, seed=21)
    dev = torch.device("cuda:0")
    to, td = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
    ref = T.KDTreeIntersection(tree).intersect_batch(to, td)
    ids = torch.empty(n, dtype=torch.int32, device=dev)
    r, s, t = (torch.empty(n, dtype=torch.float32, device=dev) for _ in range(3))
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        T.check(T.lib().tn_intersect(tree.handle, T.ptr(to), T.ptr(td), n, T.ptr(ids), T.ptr(r), T.ptr(s),
                                     T.ptr(t), 0, torch.cuda.current_stream().cuda_stream))
    for k in range(3):
        for x in (ids, r, s, t):
            x.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for x, y in zip((ids, r, s, t), ref):
            assert torch.equal(x, y), f"replay {k}"
    del g


def test_render_tree_gives_the_same_hits(gpu):
    """TN_SCENE_EMPTY_LEAVES (rendering tree): the closest hit distance of
    every ray equals the reference tree's bit for bit; ids may differ only
    between triangles at exactly that distance (shared edges)."""
    for sc in (scenes.occluded_light_cornell(level=5), scenes.intersect_scene(level=4)):
        ref = T.KDTree(sc["verts"])
        ren = T.KDTree(sc["verts"], empty_leaves=True)
        lo, hi = np.asarray(ref.box()[:3]), np.asarray(ref.box()[3:])
        rng = np.random.default_rng(3)
        n = 1 << 18
        o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
        d = rng.normal(size=(n, 3)).astype(np.float32)
        a = T.KDTreeIntersection(ref).intersect_batch(o, d)
        b = T.KDTreeIntersection(ren).intersect_batch(o, d)
        assert np.array_equal(a[1], b[1]), "hit distances differ"
        diff = a[0] != b[0]
        print(f"[render tree] {len(sc['verts'])} tris: ids differ on {diff.sum()} of {n} rays (ties)", flush=True)
        assert diff.mean() <= 1e-3
        ra = a[1][diff]
        assert np.all(np.isfinite(ra)) and np.all(a[0][diff] != T.TN_INVALID_ID)

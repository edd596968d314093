"""CPU tests of the product's host side (no GPU calls): the SAH builder
against the oracle, the tree cache, the PPM writer and the C ABI exports."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as O
import turner_amd as T
from turner_amd import scenes
from tests import helpers as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _scenes():
    rng = np.random.default_rng(7)
    yield "cornell", scenes.cornell_box()["verts"]
    yield "icosphere4", scenes.intersect_scene(level=4)["verts"]
    yield "icosphere6", scenes.intersect_scene(level=6)["verts"]
    yield "sphere_in_cornell", scenes.sphere_in_cornell(level=4)["verts"]
    yield "occluded", scenes.occluded_light_cornell(level=3)["verts"]
    yield "soup", H.random_soup(4000, rng)
    yield "grid", H.grid_mesh(40)
    yield "coplanar", H.coplanar_stack(1, rng)
    yield "plane", H.plane_triangles(2, 500, rng)
    yield "cube", H.cube12()
    yield "single", H.tri((0, 0, 1), (0, 1, 1), (1, 0, 1))[None]


@pytest.mark.parametrize("name,verts", list(_scenes()))
def test_builder_matches_oracle(name, verts):
    """The product builder (C++, std::sort, parallel subtrees) and the oracle
    (C, restated introsort, serial) build bit-identical trees."""
    nodes, box, height = T.build_tree(verts)
    ot = O.Tree(verts)
    assert np.array_equal(nodes, ot.nodes), name
    assert np.array_equal(box, ot.box)
    assert height == ot.height()
    serial, _, _ = T.build_tree(verts, serial=True)
    assert np.array_equal(serial, nodes)


def test_tree_cache_roundtrip_and_validation(tmp_path):
    v = scenes.intersect_scene(level=3)["verts"]
    nodes, box, _ = T.build_tree(v)
    path = str(tmp_path / "tree.cache")
    T.tree_cache_save(path, v, nodes, box)
    n2, b2 = T.tree_cache_load(path, v)
    assert np.array_equal(n2, nodes) and np.array_equal(b2, box)
    # other geometry: the reference would silently use the stale cache (main.cpp:147-153)
    v2 = v.copy()
    v2[0, 0] += 1e-3
    with pytest.raises(T.TurnerError, match="stale"):
        T.tree_cache_load(path, v2)
    with pytest.raises(T.TurnerError):
        T.tree_cache_load(str(tmp_path / "missing"), v)
    raw = open(path, "rb").read()
    open(path, "wb").write(raw[: len(raw) - 16])
    with pytest.raises(T.TurnerError, match="truncated"):
        T.tree_cache_load(path, v)
    bad = bytearray(raw)
    hdr = 4 + 4 + 4 + 8 + 8 + 24
    bad[hdr:hdr + 8] = (0xFFFFFFF0).to_bytes(8, "little")  # root: inner node, right out of range
    open(path, "wb").write(bytes(bad))
    with pytest.raises(T.TurnerError, match="corrupt"):
        T.tree_cache_load(path, v)


def test_ppm_writer_matches_reference_format():
    """lib/raster.h:79-100 and tests/test_raster.cpp:44-57"""
    img = np.zeros((2, 2, 4), np.float32)
    img[1, 0] = 1
    assert T.write_ppm(img) == "P3\n2 2\n255\n  0   0   0   0   0   0\n255 255 255   0   0   0"
    rng = np.random.default_rng(0)
    img = rng.uniform(-0.2, 1.3, (7, 5, 4)).astype(np.float32)
    assert T.write_ppm(img) == O.ppm(img)


def test_every_declared_symbol_is_exported():
    header = open(os.path.join(ROOT, "include", "turner_amd.h")).read()
    declared = set(re.findall(r"^\s*(?:int|long|void|const char\*)\s+(tn_\w+)\(", header, re.M))
    assert len(declared) >= 20
    lib = C.CDLL(os.path.join(ROOT, "turner_amd", "libturner_amd.so"))
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    from turner_amd import _lib
    assert declared <= set(_lib.SIGNATURES)


def test_invalid_arguments_fail_loudly():
    with pytest.raises(T.TurnerError):
        T.build_tree(np.zeros((0, 9), np.float32))
    L = T.lib()
    assert L.tn_scene_destroy(None) == T._lib.TN_ERR_INVALID
    assert L.tn_intersect(None, None, None, 0, None, None, None, None, 0, None) == T._lib.TN_ERR_INVALID
    assert b"NULL" in L.tn_last_error() or b"scene" in L.tn_last_error()


def test_no_device_is_an_error_not_a_fallback():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(T.TurnerError) as e:
        T.KDTree(H.cube12())
    assert e.value.code == T._lib.TN_ERR_NODEVICE


def _leaf_of_point(nodes, p):
    """Leaves (node indices) whose cell holds point p: descend by the split
    planes, both ways for a point on a plane."""
    out, st = [], [0]
    while st:
        i = st.pop()
        x, y = int(nodes[i, 0]), int(nodes[i, 1])
        if (y & 3) == 3:
            out.append(i)
            continue
        ax = y & 3
        split = np.array([x], np.uint32).view(np.float32)[0]
        if p[ax] <= split:
            st.append(i + 1)
        if p[ax] >= split:
            st.append(y >> 2)
    return out


@pytest.mark.parametrize("name", ["soup", "grid", "sphere_walls"])
def test_render_tree_covers_every_triangle(name):
    """The rendering tree (TN_SCENE_EMPTY_LEAVES) drops triangles clipped far
    away and keeps empty leaves; it must still be conservative: every point of
    every triangle lies in a leaf that references the triangle, so a ray
    hitting a triangle visits a leaf that tests it.  Also: valid indices, one
    leaf per cell, and at most 4x the reference tree's references (its splits
    lie between vertex coordinates, so straddling triangles are shared)."""
    rng = np.random.default_rng(5)
    if name == "soup":
        verts = H.random_soup(300, rng)
    elif name == "grid":
        verts = H.grid_mesh(12)
    else:
        v, _ = scenes.icosphere(3, 0.25, 0.05, 1)
        v = (v.reshape(-1, 3) + np.array([0.5, 0.3, 0.5], np.float32)).reshape(-1, 9)
        verts = np.concatenate([scenes.concat(*scenes.unit_cornell_walls())[0], v])
    verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 9)
    nodes, refs, box, h = T.build_device_tree(verts, empty_leaves=True)
    rnodes, rrefs, rbox, _ = T.build_device_tree(verts)
    assert np.array_equal(box, rbox)
    assert len(refs) <= 4 * len(rrefs)
    inner = (nodes[:, 1] & 3) != 3
    assert np.all((nodes[inner, 1] >> 2) < len(nodes))
    leaf = ~inner
    assert np.all(nodes[leaf, 0] + (nodes[leaf, 1] >> 2) <= len(refs))
    assert np.all(refs < len(verts))
    tri = verts.reshape(-1, 3, 3).astype(np.float64)
    for t in range(len(tri)):
        for _ in range(6):
            a, b = rng.uniform(size=2)
            if a + b > 1:
                a, b = 1 - a, 1 - b
            p = (tri[t, 0] + a * (tri[t, 1] - tri[t, 0]) + b * (tri[t, 2] - tri[t, 0])).astype(np.float32)
            cells = _leaf_of_point(nodes, p)
            assert cells, (t, p)
            found = any(t in refs[nodes[c, 0]:nodes[c, 0] + (nodes[c, 1] >> 2)] for c in cells)
            assert found, f"triangle {t} point {p} not referenced by its leaf"


def test_triangle_records_bitwise_equal_to_oracle():
    """The product's precompute (the records tn_scene_create uploads) equals the
    oracle's or_precompute bit for bit (lib/triangle.h:22-39): random soups
    over several magnitudes, the reference test meshes, the displaced icosphere
    and degenerate triangles (repeated vertices, collinear points)."""
    from turner_amd import scenes
    rng = np.random.default_rng(11)
    sets = [rng.uniform(-1, 1, (200000, 9)), rng.uniform(-1e3, 1e3, (50000, 9)),
            rng.uniform(-1e-3, 1e-3, (50000, 9)), H.cube12(), H.four_separated(), H.grid_mesh(8),
            scenes.intersect_scene(level=5)["verts"]]
    deg = rng.uniform(-1, 1, (1000, 9)).astype(np.float32)
    deg[:500, 3:6] = deg[:500, 0:3]                      # v1 == v0
    deg[500:, 6:9] = 2 * deg[500:, 3:6] - deg[500:, 0:3]  # v2 on the line v0 v1
    sets.append(deg)
    for v in sets:
        v = np.ascontiguousarray(v, np.float32).reshape(-1, 9)
        a = T.triangle_records(v)
        b = O.precompute(v)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), len(v)


def test_refined_tree_shape():
    """TN_SCENE_REFINED on the host: the refined tree keeps every inner node of
    the reference device tree in the same depth-first order, and every leaf
    below a refined reference leaf holds only triangles of that leaf, among
    them every triangle of the leaf whose bounding box meets the leaf's box."""
    sc = scenes.occluded_light_cornell(level=4)
    v = sc["verts"].reshape(-1, 3, 3)
    rn, rr, rb, _ = T.build_device_tree(sc["verts"])
    fn, fr, fb, _ = T.build_device_tree(sc["verts"], refined=True)
    assert np.array_equal(rb, fb) and len(fr) < 2 * len(rr)
    stack = [(0, 0, rb[:3].astype(np.float64), rb[3:].astype(np.float64))]
    leaves = 0
    while stack:
        i, j, lo, hi = stack.pop()
        x, y = int(rn[i, 0]), int(rn[i, 1])
        fx, fy = int(fn[j, 0]), int(fn[j, 1])
        if (y & 3) != 3:
            assert (fx, fy & 3) == (x, y & 3), "an inner reference node changed"
            ax, split = y & 3, float(np.array([x], np.uint32).view(np.float32)[0])
            l_hi, r_lo = hi.copy(), lo.copy()
            l_hi[ax] = split
            r_lo[ax] = split
            stack += [(y >> 2, fy >> 2, r_lo, hi), (i + 1, j + 1, lo, l_hi)]
            continue
        leaves += 1
        leaf = set(rr[x:x + (y >> 2)].tolist())
        below, sub = set(), [j]
        while sub:
            k = sub.pop()
            kx, ky = int(fn[k, 0]), int(fn[k, 1])
            if (ky & 3) == 3:
                below |= set(fr[kx:kx + (ky >> 2)].tolist())
            else:
                sub += [ky >> 2, k + 1]
        assert below <= leaf
        touching = {t for t in leaf if np.all(v[t].min(0) <= hi) and np.all(v[t].max(0) >= lo)}
        assert touching <= below, "a triangle meeting the leaf's box was dropped"
    assert leaves > 100

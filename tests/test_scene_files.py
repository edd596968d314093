"""Scene and image files (no reference counterpart for the formats: turner
imports .blend through assimp, which is not available here): the OBJ/MTL
loader and writer, and the portable float map of linear images."""
import numpy as np
import pytest

import turner_amd as T
from turner_amd import scenes


@pytest.mark.parametrize("make", [scenes.cornell_box, lambda: scenes.sphere_in_cornell(level=3)])
def test_obj_round_trip_is_exact(make, tmp_path):
    sc = make()
    path = str(tmp_path / "s.obj")
    scenes.save_obj(sc, path)
    back = scenes.load_obj(path)
    for k in ("verts", "normals", "albedo", "emission"):
        assert np.array_equal(sc[k], back[k]), k


def test_obj_features(tmp_path):
    (tmp_path / "m.mtl").write_text("newmtl red\nKd 0.9 0.1 0.1\nnewmtl lamp\nKd 0 0 0\nKe 5 4 3\n")
    (tmp_path / "s.obj").write_text(
        "# a quad (fan), a triangle with normals and texture references, relative indices\n"
        "mtllib m.mtl\n"
        "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\n"
        "vt 0 0\nvn 0 0 1\n"
        "usemtl red\n"
        "f 1 2 3 4\n"
        "usemtl lamp\n"
        "v 0 0 2\nv 0 2 2\nv 2 0 2\n"
        "f -3/1/1 -2/1/1 -1/1/1\n"
        "usemtl missing\n"
        "f 1 3 4\n")
    sc = scenes.load_obj(str(tmp_path / "s.obj"), albedo=(0.5, 0.5, 0.5))
    v = sc["verts"].reshape(-1, 3, 3)
    assert len(v) == 4
    assert np.array_equal(v[0], [[0, 0, 0], [1, 0, 0], [1, 1, 0]])  # fan: (1 2 3), (1 3 4)
    assert np.array_equal(v[1], [[0, 0, 0], [1, 1, 0], [0, 1, 0]])
    assert np.array_equal(v[2], [[0, 0, 2], [0, 2, 2], [2, 0, 2]])
    n = sc["normals"].reshape(-1, 3, 3)
    assert np.array_equal(n[0], np.tile([0, 0, 1], (3, 1)))  # face normal (no vn)
    assert np.array_equal(n[2], np.tile([0, 0, 1], (3, 1)))  # the given vn
    assert np.allclose(sc["albedo"][0], [0.9, 0.1, 0.1, 1]) and np.allclose(sc["albedo"][3], [0.5, 0.5, 0.5, 1])
    assert np.allclose(sc["emission"][2], [5, 4, 3, 1]) and not sc["emission"][0].any()
    # a node transformation: vertices by T, normals by its 3x3 part (main.cpp:33-61)
    Tm = np.array([[0, -1, 0, 10], [1, 0, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], np.float64)
    st = scenes.load_obj(str(tmp_path / "s.obj"), transform=Tm)
    assert np.allclose(st["verts"].reshape(-1, 3, 3)[0, 1], [10, 1, 0])
    assert np.allclose(st["normals"].reshape(-1, 3, 3)[2, 0], [0, 0, 1])


def test_obj_scene_builds_the_same_tree(tmp_path):
    """The loaded Cornell box gives the reference tree of the procedural one."""
    sc = scenes.cornell_box()
    scenes.save_obj(sc, str(tmp_path / "c.obj"))
    back = scenes.load_obj(str(tmp_path / "c.obj"))
    a, ba, ha = T.build_tree(sc["verts"])
    b, bb, hb = T.build_tree(back["verts"])
    assert np.array_equal(a, b) and np.array_equal(ba, bb) and ha == hb


def test_pfm_round_trip_and_layout(tmp_path):
    rng = np.random.default_rng(1)
    img = rng.random((5, 7, 4), dtype=np.float32)
    path = str(tmp_path / "i.pfm")
    T.write_pfm(path, img)
    data = open(path, "rb").read()
    assert data.startswith(b"PF\n7 5\n-1.0\n")
    body = np.frombuffer(data[len(b"PF\n7 5\n-1.0\n"):], "<f4").reshape(5, 7, 3)
    assert np.array_equal(body[0], img[-1, :, :3])  # bottom row first
    assert np.array_equal(T.read_pfm(path), img[..., :3])
    # big-endian files (positive scale) read too
    (tmp_path / "b.pfm").write_bytes(b"PF\n7 5\n1.0\n" + np.ascontiguousarray(img[::-1, :, :3], ">f4").tobytes())
    assert np.array_equal(T.read_pfm(str(tmp_path / "b.pfm")), img[..., :3])

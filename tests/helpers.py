"""Shared test inputs: the reference's own test geometry and seeded scenes."""
import math

import numpy as np


def tri(a, b, c):
    return np.array([*a, *b, *c], np.float32)


def four_separated():
    """tests/test_kdtree.cpp:23-28"""
    return np.stack([tri((0, 0, 1), (0, 1, 1), (1, 0, 1)), tri((2, 0, 1), (3, 0, 1), (3, 1, 1)),
                     tri((0, 2, 1), (0, 3, 1), (1, 3, 1)), tri((3, 2, 1), (3, 3, 1), (2, 3, 1))])


def cube12():
    """tests/test_kdtree.cpp:162-183"""
    T = [((-1, -1, -1), (1, -1, -1), (1, 1, -1)), ((-1, -1, -1), (-1, 1, -1), (1, 1, -1)),
         ((-1, -1, 1), (1, -1, 1), (1, 1, 1)), ((-1, -1, 1), (-1, 1, 1), (1, 1, 1)),
         ((-1, -1, -1), (-1, -1, 1), (-1, 1, 1)), ((-1, -1, -1), (-1, 1, -1), (-1, 1, 1)),
         ((1, -1, -1), (1, -1, 1), (1, 1, 1)), ((1, -1, -1), (1, 1, -1), (1, 1, 1)),
         ((-1, 1, -1), (-1, 1, 1), (1, 1, 1)), ((-1, 1, -1), (1, 1, -1), (1, 1, 1)),
         ((-1, -1, -1), (-1, -1, 1), (1, -1, 1)), ((-1, -1, -1), (1, -1, -1), (1, -1, 1))]
    return np.stack([tri(*t) for t in T])


def coplanar_stack(ax, rng):
    """tests/test_kdtree.cpp:211-216: 10 regular triangles on the unit circle in
    planes ax = 0..9 (random rotation phi, same for all, as the reference's
    static generator draws one phi per triangle)."""
    out = []
    for pos in range(10):
        phi = rng.uniform(0, 2 * math.pi)
        v = np.zeros((3, 3), np.float32)
        for i in range(3):
            v[i, ax] = pos
            v[i, (ax + 1) % 3] = math.cos(phi + 2 * math.pi / 3 * i)
            v[i, (ax + 2) % 3] = math.sin(phi + 2 * math.pi / 3 * i)
        out.append(v.reshape(9))
    return np.stack(out).astype(np.float32)


def plane_triangles(ax, n, rng):
    """tests/test_kdtree.cpp:188-193: random triangles with vertices on the unit
    circle in the plane ax = 0."""
    out = []
    for _ in range(n):
        v = np.zeros((3, 3), np.float32)
        for i in range(3):
            phi = rng.uniform(0, 2 * math.pi)
            v[i, (ax + 1) % 3] = math.cos(phi)
            v[i, (ax + 2) % 3] = math.sin(phi)
        out.append(v.reshape(9))
    return np.stack(out).astype(np.float32)


def random_soup(n, rng, spread=10.0, size=0.3):
    """tests/test_kdtree.cpp:97-115 stress-test style soup."""
    c = rng.uniform(-spread, spread, (n, 1, 3))
    v = c + rng.uniform(-size, size, (n, 3, 3))
    return v.reshape(n, 9).astype(np.float32)


def grid_mesh(n, z=0.0):
    """Axis-aligned n x n grid of unit quads (2 triangles each) in the plane
    z: many shared edges, equal event positions and exact ties."""
    out = []
    for i in range(n):
        for j in range(n):
            a, b, c, d = (i, j, z), (i + 1, j, z), (i + 1, j + 1, z), (i, j + 1, z)
            out.append(tri(a, b, c))
            out.append(tri(a, c, d))
    return np.stack(out)


def hard_rays(verts, n, rng):
    """Rays aimed at vertices, edge midpoints and split-plane-like positions
    plus axis-aligned and zero-component directions."""
    v = verts.reshape(-1, 3, 3)
    lo, hi = v.reshape(-1, 3).min(0), v.reshape(-1, 3).max(0)
    span = hi - lo
    k = rng.integers(0, len(v), n)
    which = rng.integers(0, 3, n)
    targets = v[k, which]
    mids = 0.5 * (v[k, which] + v[k, (which + 1) % 3])
    sel = rng.random(n) < 0.5
    tgt = np.where(sel[:, None], targets, mids)
    o = lo - 0.5 * span + rng.random((n, 3)) * 2 * span
    d = tgt - o
    # a quarter of the rays axis-aligned through the target
    ax = rng.integers(0, 3, n)
    aa = rng.random(n) < 0.25
    for i in np.nonzero(aa)[0]:
        o[i] = tgt[i]
        o[i, ax[i]] = lo[ax[i]] - span[ax[i]]
        d[i] = 0
        d[i, ax[i]] = 1
    return o.astype(np.float32), d.astype(np.float32)


def wall_scene(z=3.0, occluder=False, reflectivity=None, light=(0.5, 0.5, 1.0)):
    """Closed-form scene for the raycaster and raytracer: a wall at z = `z`
    facing a camera at the origin that looks along +z (every primary ray has
    d.z = 1, so it hits the wall at ray parameter r = z exactly), one point
    light in front of the wall, optionally a small occluder quad at z = 2
    between light and wall, optionally the wall's reflectivity (reflective
    colour (0.9, 0.8, 0.7, 1))."""
    from turner_amd import scenes as S
    wall = S._pack(S._quad([-20, -20, z], [-20, 20, z], [20, 20, z], [20, -20, z]), (0.6, 0.5, 0.4))
    parts = [wall]
    if occluder:
        parts.append(S._pack(S._quad([0.3, 0.3, 2.0], [0.3, 0.9, 2.0], [0.9, 0.9, 2.0], [0.9, 0.3, 2.0]),
                             (0.2, 0.7, 0.3)))
    verts, normals, albedo, emission = S.concat(*parts)
    sc = {"verts": verts, "normals": normals, "albedo": albedo, "emission": emission,
          "camera": S.camera((0.0, 0.0, 0.0), 60.0),
          "lights": [(light, (1.0, 0.9, 0.8, 1.0))]}
    n = len(verts)
    sc["reflective"] = np.tile(np.array([0.9, 0.8, 0.7, 1.0], np.float32), (n, 1))
    sc["reflectivity"] = np.zeros(n, np.float32)
    if reflectivity is not None:
        sc["reflectivity"][:2] = reflectivity
    return sc

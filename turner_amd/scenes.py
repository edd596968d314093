"""Procedural scenes and ray batches for the pathtracer hot path.

The reference imports Blender files through assimp (main.cpp:25-136), which
is not available here; the scenes its benchmarks name are rebuilt
procedurally instead, with its camera and light conventions
(lib/types.h:89-144: ray = trafo * (-dx(1-2px/w), dy(1-2py/h), -1) from the
camera position; lib/types.h:81-84 point lights).

A scene is a dict with float32 arrays
    verts    (N, 9)   v0 v1 v2
    normals  (N, 9)   vertex normals (flat faces repeat the face normal)
    albedo   (N, 4)   RGBA diffuse colour (alpha 1)
    emission (N, 4)   RGBA emitted radiance (zero for non-emitters)
plus "camera" (pos, trafo 3x3 row-major, delta_x, delta_y) and "lights"
(list of (pos, rgba)).
"""
import math

import numpy as np

WHITE = (0.73, 0.73, 0.73)
RED = (0.63, 0.06, 0.05)
GREEN = (0.14, 0.45, 0.09)


# ---------------------------------------------------------------- helpers
def _quad(a, b, c, d):
    """Two triangles (a, b, c), (a, c, d); face normal = normalize((b-a) x (c-a))."""
    return [np.array([a, b, c], np.float32), np.array([a, c, d], np.float32)]


def _face_normals(verts):
    v = verts.reshape(-1, 3, 3).astype(np.float64)
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    return n.astype(np.float32)


def _pack(tris, albedo, emission=(0, 0, 0)):
    verts = np.stack(tris).reshape(-1, 9).astype(np.float32)
    fn = _face_normals(verts)
    normals = np.tile(fn, (1, 3))
    alb = np.tile(np.array([*albedo, 1.0], np.float32), (len(verts), 1))
    emi = np.tile(np.array([*emission, 1.0 if any(emission) else 0.0], np.float32),
                  (len(verts), 1))
    return verts, normals, alb, emi


def concat(*parts):
    return tuple(np.concatenate([p[k] for p in parts]) for k in range(4))


def box_mesh(size, center, angle_deg, albedo):
    """Axis-aligned box (sx, sy, sz) on the floor at center (cx, y0, cz),
    rotated about the vertical axis; outward-facing triangles (12)."""
    sx, sy, sz = size
    x0, x1, y0, y1, z0, z1 = -sx / 2, sx / 2, 0.0, sy, -sz / 2, sz / 2
    P = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float64)
    a = math.radians(angle_deg)
    R = np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])
    P = P @ R.T + np.array(center, np.float64)
    P = P.astype(np.float32)
    faces = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (3, 7, 6, 2), (0, 4, 7, 3), (1, 2, 6, 5)]
    tris = []
    for f in faces:
        tris += _quad(P[f[0]], P[f[1]], P[f[2]], P[f[3]])
    return _pack(tris, albedo)


def camera(pos, fov_deg, aspect=1.0, look=(0.0, 0.0, 1.0)):
    """Pinhole camera looking along +z (world), up +y: trafo = diag(-1, 1, -1)
    maps camera space (looking down -z, x right) to world.  delta_x is
    tan(fov/2): assimp's mHorizontalFOV is the half angle (lib/types.h:101)."""
    assert tuple(look) == (0.0, 0.0, 1.0)
    dx = math.tan(math.radians(fov_deg) / 2.0)
    return {"pos": np.array(pos, np.float32),
            "trafo": np.array([[-1, 0, 0], [0, 1, 0], [0, 0, -1]], np.float32),
            "delta_x": np.float32(dx), "delta_y": np.float32(dx / aspect)}


# ---------------------------------------------------------------- Cornell box
def cornell_walls(L=555.0, light_quad=True):
    """Floor, ceiling and back wall white, left (x = L) red, right (x = 0)
    green, open towards the camera (z < 0); normals face inward."""
    parts = []
    parts.append(_pack(_quad([0, 0, 0], [0, 0, L], [L, 0, L], [L, 0, 0]), WHITE))      # floor
    parts.append(_pack(_quad([0, L, 0], [L, L, 0], [L, L, L], [0, L, L]), WHITE))      # ceiling
    parts.append(_pack(_quad([0, 0, L], [0, L, L], [L, L, L], [L, 0, L]), WHITE))      # back
    parts.append(_pack(_quad([L, 0, 0], [L, 0, L], [L, L, L], [L, L, 0]), RED))        # left
    parts.append(_pack(_quad([0, 0, 0], [0, L, 0], [0, L, L], [0, 0, L]), GREEN))      # right
    return parts


def cornell_box():
    """The benchmark Cornell box: 555-unit box, short (165) and tall (330)
    boxes rotated -17 / +17 degrees, a 2-triangle light quad (diffuse, as in
    cornell_box.blend), a point light below the ceiling centre and the
    camera at the open side with a 39.3 degree field of view: 36 triangles."""
    parts = cornell_walls()
    parts.append(box_mesh((165, 165, 165), (185.5, 0.0, 168.5), -17.0, WHITE))
    parts.append(box_mesh((165, 330, 165), (368.5, 0.0, 351.5), 17.0, WHITE))
    y = 554.9  # faces down, into the room
    parts.append(_pack(_quad([213, y, 227], [343, y, 227], [343, y, 332], [213, y, 332]),
                       (0.78, 0.78, 0.78)))
    verts, normals, albedo, emission = concat(*parts)
    return {"verts": verts, "normals": normals, "albedo": albedo, "emission": emission,
            "camera": camera((277.5, 273.0, -800.0), 39.3),
            "lights": [((277.5, 545.0, 277.5), (1.0, 1.0, 1.0, 1.0))]}


# ---------------------------------------------------------------- icosphere
def _hash3(ix, iy, iz, seed):
    h = (ix.astype(np.uint32) * np.uint32(0x8DA6B343) ^ iy.astype(np.uint32) * np.uint32(0xD8163841)
         ^ iz.astype(np.uint32) * np.uint32(0xCB1AB31F) ^ np.uint32(seed * 0x9E3779B1 & 0xFFFFFFFF))
    h ^= h >> np.uint32(16)
    h *= np.uint32(0x7FEB352D)
    h ^= h >> np.uint32(15)
    h *= np.uint32(0x846CA68B)
    h ^= h >> np.uint32(16)
    return h.astype(np.float64) / 4294967295.0 * 2.0 - 1.0


def value_noise(p, seed, octaves=3, base_freq=2.0):
    """Seeded lattice value noise in [-1, 1]: hashed lattice values, smoothstep
    trilinear interpolation, octaves at doubling frequency and halving weight."""
    total = np.zeros(len(p))
    norm = 0.0
    for o in range(octaves):
        f = base_freq * 2 ** o
        w = 0.5 ** o
        q = p * f
        i0 = np.floor(q).astype(np.int64)
        fr = q - i0
        s = fr * fr * (3 - 2 * fr)
        acc = 0.0
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    wgt = ((s[:, 0] if dx else 1 - s[:, 0]) * (s[:, 1] if dy else 1 - s[:, 1])
                           * (s[:, 2] if dz else 1 - s[:, 2]))
                    acc = acc + wgt * _hash3(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz,
                                             seed + 7919 * o)
        total += w * acc
        norm += w
    return total / norm


def icosphere(level, radius=1.0, amplitude=0.0, seed=1, smooth=True):
    """Level-k icosphere (20 * 4^k triangles, 10 * 4^k + 2 shared vertices),
    vertices displaced radially by radius * (1 + amplitude * noise).  Returns
    (verts (N,9), normals (N,9)); outward winding; smooth normals average the
    adjacent face normals (area weighted), flat normals repeat the face normal."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    V = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]],
                 np.float64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    F = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9],
                  [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2],
                  [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
                  [8, 6, 7], [9, 8, 1]], np.int64)
    for _ in range(level):
        nv = len(V)
        e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
        key = np.minimum(e[:, 0], e[:, 1]) * nv + np.maximum(e[:, 0], e[:, 1])
        uk, inv = np.unique(key, return_inverse=True)
        a, b = uk // nv, uk % nv
        mid = V[a] + V[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        V = np.concatenate([V, mid])
        m = inv.reshape(3, -1).T + nv  # midpoints of (0,1), (1,2), (2,0)
        F = np.concatenate([np.stack([F[:, 0], m[:, 0], m[:, 2]], 1),
                            np.stack([F[:, 1], m[:, 1], m[:, 0]], 1),
                            np.stack([F[:, 2], m[:, 2], m[:, 1]], 1),
                            m])
    if amplitude:
        V = V * (1.0 + amplitude * value_noise(V, seed))[:, None]
    V = V * radius
    tri = V[F]  # (N, 3, 3)
    fn = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    if smooth:
        vn = np.zeros_like(V)
        for k in range(3):
            np.add.at(vn, F[:, k], fn)
        vn /= np.linalg.norm(vn, axis=1, keepdims=True)
        normals = vn[F].reshape(-1, 9)
    else:
        fnn = fn / np.linalg.norm(fn, axis=1, keepdims=True)
        normals = np.tile(fnn, (1, 3))
    return tri.reshape(-1, 9).astype(np.float32), normals.astype(np.float32)


def cube_inward(half=2.0, center=(0.0, 0.0, 0.0)):
    """12-triangle cube [-half, half]^3 with inward-facing winding."""
    h = half
    c = np.array(center, np.float32)
    P = np.array([[-h, -h, -h], [h, -h, -h], [h, h, -h], [-h, h, -h],
                  [-h, -h, h], [h, -h, h], [h, h, h], [-h, h, h]], np.float32) + c
    faces = [(0, 1, 2, 3), (4, 7, 6, 5), (0, 4, 5, 1), (3, 2, 6, 7), (0, 3, 7, 4), (1, 5, 6, 2)]
    tris = []
    for f in faces:
        tris += _quad(P[f[0]], P[f[1]], P[f[2]], P[f[3]])
    return np.stack(tris).reshape(-1, 9)


def intersect_scene(level=8, seed=1):
    """Ray-batch microbenchmark mesh: level-k icosphere (unit radius, radial
    3-octave value noise of amplitude 0.05) inside a 4-unit cube of 12
    triangles.  Level 8 gives 1,310,720 + 12 triangles."""
    v, n = icosphere(level, 1.0, 0.05, seed)
    c = cube_inward(2.0)
    cn = np.tile(_face_normals(c), (1, 3))
    verts = np.concatenate([v, c]).astype(np.float32)
    normals = np.concatenate([n, cn]).astype(np.float32)
    return {"verts": verts, "normals": normals}


def random_rays(n, seed, lo=-2.0, hi=2.0):
    """n rays, origins uniform in [lo, hi]^3, directions uniform on the unit
    sphere (z ~ U(-1, 1), phi ~ U(0, 2 pi)); float32 (n, 3) each."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    z = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2 * math.pi, n)
    r = np.sqrt(np.maximum(0.0, 1 - z * z))
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1).astype(np.float32)
    return o, d


# ------------------------------------------------- extended (area-light) scenes
def unit_cornell_walls():
    parts = cornell_walls(L=1.0)
    return parts


def emitter_quad(cx, cz, size, y, radiance):
    """Horizontal emitter facing down (normal -y) at height y."""
    h = size / 2
    tris = _quad([cx - h, y, cz - h], [cx + h, y, cz - h], [cx + h, y, cz + h], [cx - h, y, cz + h])
    return _pack(tris, (0.0, 0.0, 0.0), (radiance, radiance, radiance))


def _mesh_part(verts, normals, albedo):
    n = len(verts)
    return (verts.astype(np.float32), normals.astype(np.float32),
            np.tile(np.array([*albedo, 1.0], np.float32), (n, 1)), np.zeros((n, 4), np.float32))


def sphere_in_cornell(level=8, seed=1, emitter_size=0.2, radiance=15.0):
    """The displaced icosphere (albedo 0.8) in the unit Cornell enclosure with a
    square area emitter on the ceiling (camera at the open side, 39.3 deg)."""
    v, n = icosphere(level, 0.25, 0.05, seed)
    v = (v.reshape(-1, 3) + np.array([0.5, 0.3, 0.5], np.float32)).reshape(-1, 9)
    parts = unit_cornell_walls()
    parts.append(_mesh_part(v, n, (0.8, 0.8, 0.8)))
    parts.append(emitter_quad(0.5, 0.5, emitter_size, 0.999, radiance))
    verts, normals, albedo, emission = concat(*parts)
    return {"verts": verts, "normals": normals, "albedo": albedo, "emission": emission,
            "camera": camera((0.5, 0.5, -800.0 / 555.0), 39.3), "lights": []}


def occluded_light_cornell(level=8, seed=1):
    """Unit Cornell enclosure whose only emitter (0.3 x 0.3, radiance 50, facing
    down just below the ceiling) hides behind a 0.5 x 0.5 baffle 0.1 below it,
    so most light reaches the room after bouncing off baffle and ceiling;
    contains the displaced icosphere."""
    v, n = icosphere(level, 0.25, 0.05, seed)
    v = (v.reshape(-1, 3) + np.array([0.5, 0.3, 0.5], np.float32)).reshape(-1, 9)
    parts = unit_cornell_walls()
    parts.append(_mesh_part(v, n, (0.8, 0.8, 0.8)))
    parts.append(emitter_quad(0.5, 0.5, 0.3, 0.999, 50.0))
    h, y = 0.25, 0.9
    # two-sided baffle: an up-facing quad 1e-4 above a down-facing one
    up = _quad([0.5 - h, y, 0.5 - h], [0.5 - h, y, 0.5 + h], [0.5 + h, y, 0.5 + h],
               [0.5 + h, y, 0.5 - h])
    up = [t + np.array([0, 1e-4, 0], np.float32) for t in up]
    down = _quad([0.5 - h, y, 0.5 - h], [0.5 + h, y, 0.5 - h], [0.5 + h, y, 0.5 + h],
                 [0.5 - h, y, 0.5 + h])
    parts.append(_pack(up, WHITE))
    parts.append(_pack(down, WHITE))
    verts, normals, albedo, emission = concat(*parts)
    return {"verts": verts, "normals": normals, "albedo": albedo, "emission": emission,
            "camera": camera((0.5, 0.5, -800.0 / 555.0), 39.3), "lights": []}


def furnace(level=3, rho=0.8, half=4.0, radiance=1.0):
    """White furnace: an (undisplaced, convex) icosphere of albedo rho, flat
    normals, inside a closed cube whose inward faces emit `radiance` and
    reflect nothing.  Closed form for max_depth >= 1: a camera ray that hits
    the sphere sees rho * radiance, one that hits the enclosure sees radiance."""
    v, n = icosphere(level, 1.0, 0.0, smooth=False)
    c = cube_inward(half)
    cn = np.tile(_face_normals(c), (1, 3))
    sphere = _mesh_part(v, n, (rho, rho, rho))
    enc = (c.astype(np.float32), cn.astype(np.float32), np.zeros((12, 4), np.float32),
           np.tile(np.array([radiance, radiance, radiance, 1.0], np.float32), (12, 1)))
    enc[2][:, 3] = 1.0
    verts, normals, albedo, emission = concat(sphere, enc)
    cam = camera((0.0, 0.0, -3.5), 40.0)
    return {"verts": verts, "normals": normals, "albedo": albedo, "emission": emission,
            "camera": cam, "lights": []}


def reference_furnace(level=3, rho=0.8):
    """The reference's furnace_test.blend: an icosphere, no lamp, lit only by
    the --background colour of escaping rays (pathtracer.cpp:26-28)."""
    v, n = icosphere(level, 1.0, 0.0, smooth=False)
    part = _mesh_part(v, n, (rho, rho, rho))
    return {"verts": part[0], "normals": part[1], "albedo": part[2], "emission": part[3],
            "camera": camera((0.0, 0.0, -4.0), 40.0), "lights": []}


# ------------------------------------------------------------ scene files
def load_obj(path, albedo=(0.8, 0.8, 0.8), transform=None):
    """Triangles of a Wavefront OBJ file (with its MTL library) as scene
    arrays: verts (N, 9), normals (N, 9), albedo (N, 4), emission (N, 4).

    Faces may be polygons (fan-triangulated) with v, v/vt, v//vn or v/vt/vn
    references, negative (relative) indices allowed.  Corners without a
    normal take the face normal.  usemtl picks the MTL material: Kd is the
    diffuse colour (default `albedo`), Ke the emission.  transform (4 x 4)
    maps vertices, its 3 x 3 part the normals, as the reference applies a
    node's transformation (main.cpp:33-61).  Camera and lights are the
    caller's (an OBJ file has none)."""
    import os
    base = os.path.dirname(os.path.abspath(path))
    V, N = [], []
    mats, cur = {}, None
    tv, tn, tm = [], [], []

    def load_mtl(fn):
        name = None
        with open(os.path.join(base, fn)) as f:
            for line in f:
                t = line.split()
                if not t or t[0].startswith("#"):
                    continue
                if t[0] == "newmtl":
                    name = " ".join(t[1:])
                    mats[name] = {"Kd": tuple(albedo), "Ke": (0.0, 0.0, 0.0)}
                elif t[0] in ("Kd", "Ke") and name is not None:
                    mats[name][t[0]] = tuple(float(x) for x in t[1:4])

    def index(tok, n):
        k = int(tok)
        return k - 1 if k > 0 else n + k

    with open(path) as f:
        for line in f:
            t = line.split()
            if not t or t[0].startswith("#"):
                continue
            if t[0] == "v":
                V.append([float(x) for x in t[1:4]])
            elif t[0] == "vn":
                N.append([float(x) for x in t[1:4]])
            elif t[0] == "mtllib":
                load_mtl(" ".join(t[1:]))
            elif t[0] == "usemtl":
                cur = " ".join(t[1:])
            elif t[0] == "f":
                corners = []
                for c in t[1:]:
                    parts = c.split("/")
                    vi = index(parts[0], len(V))
                    ni = index(parts[2], len(N)) if len(parts) > 2 and parts[2] else None
                    corners.append((vi, ni))
                for k in range(1, len(corners) - 1):
                    tri = (corners[0], corners[k], corners[k + 1])
                    tv.append([V[c[0]] for c in tri])
                    tn.append([N[c[1]] if c[1] is not None else None for c in tri])
                    tm.append(cur)
    verts = np.asarray(tv, np.float64).reshape(-1, 3, 3)
    if transform is not None:
        T = np.asarray(transform, np.float64)
        verts = verts @ T[:3, :3].T + T[:3, 3]
    fn = np.cross(verts[:, 1] - verts[:, 0], verts[:, 2] - verts[:, 0])
    fn /= np.maximum(np.linalg.norm(fn, axis=1, keepdims=True), 1e-300)
    normals = np.empty_like(verts)
    for i, corner in enumerate(tn):
        for k in range(3):
            n = corner[k]
            if n is None:
                normals[i, k] = fn[i]
            else:
                n = np.asarray(n, np.float64)
                normals[i, k] = n @ np.asarray(transform, np.float64)[:3, :3].T if transform is not None else n
    alb = np.array([[*(mats.get(m, {}).get("Kd", albedo)), 1.0] for m in tm], np.float32).reshape(-1, 4)
    ke = [mats.get(m, {}).get("Ke", (0.0, 0.0, 0.0)) for m in tm]
    emi = np.array([[*e, 1.0 if any(e) else 0.0] for e in ke], np.float32).reshape(-1, 4)
    return {"verts": verts.reshape(-1, 9).astype(np.float32), "normals": normals.reshape(-1, 9).astype(np.float32),
            "albedo": alb, "emission": emi}


def save_obj(scene, path):
    """Writes the triangles of a scene dict as OBJ (three v and vn per
    triangle, exact float32 text) with an MTL library of its distinct
    (albedo, emission) pairs next to it; load_obj reads it back."""
    import os
    verts = np.asarray(scene["verts"], np.float32).reshape(-1, 3, 3)
    normals = np.asarray(scene["normals"], np.float32).reshape(-1, 3, 3)
    alb = np.asarray(scene["albedo"], np.float32).reshape(-1, 4)
    emi = np.asarray(scene["emission"], np.float32).reshape(-1, 4)
    mtl_path = os.path.splitext(path)[0] + ".mtl"
    keys, names = {}, []
    for a, e in zip(alb, emi):
        k = (tuple(a[:3].tolist()), tuple(e[:3].tolist()))
        if k not in keys:
            keys[k] = f"m{len(keys)}"
        names.append(keys[k])
    g = lambda x: repr(float(x))
    with open(mtl_path, "w") as f:
        for (kd, ke), name in keys.items():
            f.write(f"newmtl {name}\nKd {' '.join(map(g, kd))}\nKe {' '.join(map(g, ke))}\n")
    with open(path, "w") as f:
        f.write(f"mtllib {os.path.basename(mtl_path)}\n")
        last = None
        for i in range(len(verts)):
            if names[i] != last:
                f.write(f"usemtl {names[i]}\n")
                last = names[i]
            for k in range(3):
                f.write("v " + " ".join(map(g, verts[i, k])) + "\n")
            for k in range(3):
                f.write("vn " + " ".join(map(g, normals[i, k])) + "\n")
            f.write("f -3//-3 -2//-2 -1//-1\n")


def save_tns(scene, path, aspect=1.0):
    """Writes a scene for the C++ front ends (turner_amd/pathtracer,
    raycaster, raytracer): "TNSC", version (1, or 2 when the scene has
    "reflectivity"), triangle and light counts, verts (9N), normals (9N),
    albedo (4N), emission (4N) [, reflective (4N), reflectivity (N)] as
    float32, camera (pos 3, trafo 9, delta_x, delta_y, aspect) and lights
    (pos 3, rgba 4)."""
    cam = scene["camera"]
    lights = scene.get("lights", [])
    keys = [("verts", 9), ("normals", 9), ("albedo", 4), ("emission", 4)]
    version = 1
    if "reflectivity" in scene:
        keys += [("reflective", 4), ("reflectivity", 1)]
        version = 2
    with open(path, "wb") as f:
        f.write(b"TNSC")
        f.write(np.array([version, len(scene["verts"]), len(lights)], np.uint32).tobytes())
        for k, cols in keys:
            f.write(np.ascontiguousarray(scene[k], np.float32).reshape(-1, cols).tobytes())
        f.write(np.concatenate([np.asarray(cam["pos"], np.float32).ravel(),
                                np.asarray(cam["trafo"], np.float32).ravel(),
                                np.array([cam["delta_x"], cam["delta_y"], aspect], np.float32)]).tobytes())
        for p, c in lights:
            f.write(np.array([*p, *c], np.float32).tobytes())

"""ctypes wrapper of liboracle.so -- the CPU restatement of turner's
pathtracer hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
INVALID_ID = 1 << 30

FLAG_COSINE = 1
FLAG_AREA_LIGHTS = 2
FLAG_REF_RNG = 4


class Camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("trafo", C.c_float * 9),
                ("delta_x", C.c_float), ("delta_y", C.c_float)]


class Light(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("color", C.c_float * 4)]


class RenderParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("max_depth", C.c_int32),
                ("mc_samples", C.c_int32), ("pixel_samples", C.c_int32),
                ("flags", C.c_uint32), ("seed", C.c_uint64), ("bg", C.c_float * 4),
                ("nthreads", C.c_int32), ("row_step", C.c_uint32), ("row_offset", C.c_uint32),
                ("renderer", C.c_uint32), ("max_visibility", C.c_float), ("shadow_intensity", C.c_float)]


class Scene(C.Structure):
    _fields_ = [("pre16", C.c_void_p), ("verts9", C.c_void_p), ("normals9", C.c_void_p),
                ("albedo4", C.c_void_p), ("emission4", C.c_void_p), ("ntris", C.c_uint32),
                ("nodes", C.c_void_p), ("box", C.c_float * 6), ("reflective4", C.c_void_p),
                ("reflectivity", C.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.or_precompute.argtypes = [P, C.c_uint32, P]
        L.or_clip_triangle.argtypes = [P, P, P]
        L.or_clip_triangle.restype = C.c_int
        L.or_ray_box.argtypes = [P, P, P, P, P]
        L.or_ray_box.restype = C.c_int
        L.or_kdtree_build.argtypes = [P, C.c_uint32, C.POINTER(C.POINTER(C.c_uint64)), P]
        L.or_kdtree_build.restype = C.c_size_t
        L.or_kdtree_height.argtypes = [P, C.c_size_t]
        L.or_kdtree_height.restype = C.c_uint32
        L.or_free.argtypes = [P]
        L.or_intersect.argtypes = [P, P, P, P, P, C.c_uint32, P, P, P, P, C.c_int]
        L.or_intersect_count.argtypes = [P, P, P, P, P, C.c_uint32, P, C.c_int]
        L.or_intersect_brute.argtypes = [P, C.c_uint32, P, P, C.c_uint32, P, P, P, P, C.c_int]
        L.or_mt64.argtypes = [P, P, P, P, C.c_uint32, P, P, P]
        L.or_xorshift64s_next.argtypes = [C.POINTER(C.c_uint64)]
        L.or_xorshift64s_next.restype = C.c_uint64
        L.or_xorshift64s_float.argtypes = [C.POINTER(C.c_uint64)]
        L.or_xorshift64s_float.restype = C.c_float
        L.or_raster2cam.argtypes = [C.POINTER(Camera), C.c_float, C.c_float, C.c_float,
                                    C.c_float, P]
        L.or_render.argtypes = [C.POINTER(Scene), C.POINTER(Camera), C.POINTER(Light),
                                C.c_uint32, C.POINTER(RenderParams), P, P,
                                C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.or_trace_rays.argtypes = [C.POINTER(Scene), C.POINTER(Light), C.c_uint32,
                                    C.POINTER(RenderParams), P, P, C.c_uint32, C.c_int32, P,
                                    C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.or_exposure.argtypes = [C.c_float, C.c_float]
        L.or_exposure.restype = C.c_float
        L.or_gamma.argtypes = [C.c_float, C.c_float]
        L.or_gamma.restype = C.c_float
        L.or_ppm.argtypes = [P, C.c_uint32, C.c_uint32, C.c_char_p, C.c_size_t]
        L.or_ppm.restype = C.c_size_t
        L.or_introsort_u64.argtypes = [P, C.c_size_t]
        L.or_flat_inner.argtypes = [C.c_int, C.c_float, C.c_uint32]
        L.or_flat_inner.restype = C.c_uint64
        L.or_flat_leaf.argtypes = [C.c_uint32, C.c_uint32]
        L.or_flat_leaf.restype = C.c_uint64
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _c(a, dtype=np.float32):
    return np.ascontiguousarray(a, dtype=dtype)


def precompute(verts):
    v = _c(verts).reshape(-1, 9)
    out = np.zeros((len(v), 16), np.float32)
    lib().or_precompute(_p(v), len(v), _p(out))
    return out


def clip_triangle(tri, box):
    t = _c(tri).reshape(9)
    b = _c(box).reshape(6)
    out = np.zeros(6, np.float32)
    n = lib().or_clip_triangle(_p(t), _p(b), _p(out))
    return out, n


def ray_box(o, d, box):
    o, d, b = _c(o), _c(d), _c(box)
    tmin = np.zeros(1, np.float32)
    tmax = np.zeros(1, np.float32)
    hit = lib().or_ray_box(_p(o), _p(d), _p(b), _p(tmin), _p(tmax))
    return bool(hit), float(tmin[0]), float(tmax[0])


class Tree:
    """Oracle kd-tree: reference FlatNode array + root box + triangle records."""

    def __init__(self, verts):
        self.verts = _c(verts).reshape(-1, 9)
        ptr = C.POINTER(C.c_uint64)()
        self.box = np.zeros(6, np.float32)
        n = lib().or_kdtree_build(_p(self.verts), len(self.verts), C.byref(ptr), _p(self.box))
        self.nodes = np.ctypeslib.as_array(ptr, shape=(n,)).copy()
        lib().or_free(ptr)
        self.pre = precompute(self.verts)

    def height(self):
        return lib().or_kdtree_height(_p(self.nodes), len(self.nodes))

    def num_nodes(self):
        return len(self.nodes)


def intersect(pre, nodes, box, o, d, nthreads=8):
    o, d = _c(o).reshape(-1, 3), _c(d).reshape(-1, 3)
    pre, nodes, box = _c(pre), _c(nodes, np.uint64), _c(box)
    n = len(o)
    ids = np.zeros(n, np.uint32)
    r, s, t = (np.zeros(n, np.float32) for _ in range(3))
    lib().or_intersect(_p(pre), _p(nodes), _p(box), _p(o), _p(d), n, _p(ids), _p(r), _p(s),
                       _p(t), nthreads)
    return ids, r, s, t


def intersect_count(pre, nodes, box, o, d, nthreads=8):
    o, d = _c(o).reshape(-1, 3), _c(d).reshape(-1, 3)
    pre, nodes, box = _c(pre), _c(nodes, np.uint64), _c(box)
    cnt = np.zeros(4, np.uint64)
    lib().or_intersect_count(_p(pre), _p(nodes), _p(box), _p(o), _p(d), len(o), _p(cnt), nthreads)
    return {"inner": int(cnt[0]), "leaves": int(cnt[1]), "refs": int(cnt[2]), "tests": int(cnt[3])}


def intersect_brute(pre, o, d, nthreads=8):
    o, d = _c(o).reshape(-1, 3), _c(d).reshape(-1, 3)
    pre = _c(pre)
    n = len(o)
    ids = np.zeros(n, np.uint32)
    r, s, t = (np.zeros(n, np.float32) for _ in range(3))
    lib().or_intersect_brute(_p(pre), len(pre), _p(o), _p(d), n, _p(ids), _p(r), _p(s), _p(t),
                             nthreads)
    return ids, r, s, t


def mt64(verts, ids, o, d):
    v = _c(verts).reshape(-1, 9)
    ids = _c(ids, np.uint32)
    o, d = _c(o).reshape(-1, 3), _c(d).reshape(-1, 3)
    n = len(o)
    r, b1, b2 = (np.zeros(n, np.float64) for _ in range(3))
    lib().or_mt64(_p(v), _p(ids), _p(o), _p(d), n, _p(r), _p(b1), _p(b2))
    return r, b1, b2


def xorshift64s(seed, count):
    s = C.c_uint64(seed)
    return [lib().or_xorshift64s_next(C.byref(s)) for _ in range(count)]


def xorshift64s_floats(seed, count):
    s = C.c_uint64(seed)
    return np.array([lib().or_xorshift64s_float(C.byref(s)) for _ in range(count)], np.float32)


def make_camera(cam):
    c = Camera()
    c.pos[:] = [float(x) for x in np.asarray(cam["pos"]).ravel()]
    c.trafo[:] = [float(x) for x in np.asarray(cam["trafo"]).ravel()]
    c.delta_x = float(cam["delta_x"])
    c.delta_y = float(cam["delta_y"])
    return c


def raster2cam(cam, px, py, w, h):
    out = np.zeros(3, np.float32)
    lib().or_raster2cam(C.byref(make_camera(cam)), px, py, w, h, _p(out))
    return out


RENDERER_PATHTRACER = 0
RENDERER_RAYCASTER = 1
RENDERER_RAYTRACER = 2


def _scene_struct(scene, tree, keep):
    def arr(a, cols):
        a = _c(a).reshape(-1, cols)
        keep.append(a)
        return a

    sc = Scene()
    sc.pre16 = tree.pre.ctypes.data
    sc.verts9 = arr(scene["verts"], 9).ctypes.data
    sc.normals9 = arr(scene["normals"], 9).ctypes.data
    sc.albedo4 = arr(scene["albedo"], 4).ctypes.data
    sc.emission4 = arr(scene["emission"], 4).ctypes.data
    sc.ntris = len(tree.pre)
    sc.nodes = tree.nodes.ctypes.data
    sc.box[:] = [float(x) for x in tree.box]
    if "reflectivity" in scene:
        sc.reflective4 = arr(scene["reflective"], 4).ctypes.data
        sc.reflectivity = arr(scene["reflectivity"], 1).ctypes.data
    lights = scene.get("lights", [])
    la = (Light * max(1, len(lights)))()
    for k, (p, col) in enumerate(lights):
        la[k].pos[:] = [float(x) for x in p]
        la[k].color[:] = [float(x) for x in col]
    return sc, la, len(lights)


def _params(width, height, max_depth, mc, spp, flags, seed, bg, nthreads, row_step, row_offset,
            renderer, max_visibility, shadow_intensity):
    prm = RenderParams()
    prm.width, prm.height = width, height
    prm.max_depth, prm.mc_samples, prm.pixel_samples = max_depth, mc, spp
    prm.flags, prm.seed = flags, seed
    prm.bg[:] = [float(x) for x in bg]
    prm.nthreads, prm.row_step, prm.row_offset = nthreads, row_step, row_offset
    prm.renderer, prm.max_visibility, prm.shadow_intensity = renderer, max_visibility, shadow_intensity
    return prm


def render(scene, tree, width, height, max_depth, mc, spp, flags=0, seed=1, bg=(0, 0, 0, 1),
           nthreads=8, row_step=1, row_offset=0, want_sq=False, renderer=RENDERER_PATHTRACER,
           max_visibility=2.0, shadow_intensity=0.5):
    """Oracle render: returns (mean RGBA (h, w, 4), sum of squares or None, rays, shadow rays).
    scene may hold "reflective" (N, 4) and "reflectivity" (N,) for the raytracer."""
    keep = []
    sc, la, nl = _scene_struct(scene, tree, keep)
    prm = _params(width, height, max_depth, mc, spp, flags, seed, bg, nthreads, row_step, row_offset,
                  renderer, max_visibility, shadow_intensity)
    out = np.zeros((height, width, 4), np.float32)
    sq = np.zeros((height, width, 4), np.float32) if want_sq else None
    nr, ns = C.c_uint64(0), C.c_uint64(0)
    lib().or_render(C.byref(sc), C.byref(make_camera(scene["camera"])), la, nl,
                    C.byref(prm), _p(out), _p(sq), C.byref(nr), C.byref(ns))
    return out, sq, nr.value, ns.value


def trace_rays(scene, tree, o, d, depth=0, max_depth=3, mc=1, flags=0, seed=1, bg=(0, 0, 0, 1),
               nthreads=8, renderer=RENDERER_PATHTRACER, max_visibility=2.0, shadow_intensity=0.5):
    """trace() of trace.h:23-25 for caller rays entering at level `depth`:
    returns (RGBA per ray (n, 4), closest-hit segments, shadow rays)."""
    keep = []
    sc, la, nl = _scene_struct(scene, tree, keep)
    prm = _params(0, 0, max_depth, mc, 1, flags, seed, bg, nthreads, 1, 0, renderer, max_visibility,
                  shadow_intensity)
    o, d = _c(o).reshape(-1, 3), _c(d).reshape(-1, 3)
    out = np.zeros((len(o), 4), np.float32)
    nr, ns = C.c_uint64(0), C.c_uint64(0)
    lib().or_trace_rays(C.byref(sc), la, nl, C.byref(prm), _p(o), _p(d), len(o), depth, _p(out),
                        C.byref(nr), C.byref(ns))
    return out, nr.value, ns.value


def ppm(rgba):
    rgba = _c(rgba)
    h, w = rgba.shape[:2]
    cap = w * h * 13 + 64
    buf = C.create_string_buffer(cap)
    n = lib().or_ppm(_p(rgba), w, h, buf, cap)
    return buf.raw[:n].decode()


def exposure(x, v):
    return lib().or_exposure(x, v)


def gamma(x, g):
    return lib().or_gamma(x, g)


def introsort_u64(keys):
    k = _c(keys, np.uint64).copy()
    lib().or_introsort_u64(_p(k), len(k))
    return k


def flat_inner(axis, split, right):
    return int(lib().or_flat_inner(axis, split, right))


def flat_leaf(a, b=0x3FFFFFFF):
    return int(lib().or_flat_leaf(a, b))


def first_id(node):
    return int(node) >> 32


def has_second(node):
    return (int(node) & 0xFFFFFFFF) != 0xFFFFFFFF


def second_id(node):
    return (int(node) & 0xFFFFFFFF) >> 2


# FlatNode field access (lib/kdtree.h:62-154): low 2 bits 3 = leaf, else axis.
def is_leaf(node):
    return (int(node) & 3) == 3


def split_axis(node):
    return int(node) & 3


def split_pos(node):
    return float(np.array([int(node) >> 32], np.uint32).view(np.float32)[0])


def right(node):
    return (int(node) & 0xFFFFFFFF) >> 2

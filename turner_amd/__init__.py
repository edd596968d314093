"""turner_amd: MI355X (gfx950) wavefront path tracer for turner's pathtracer hot path.

Python mirror of the reference's host interface over the C ABI of
libturner_amd.so (include/turner_amd.h):

* KDTree(verts, ...)                -- KDTree::KDTree(Triangles)   lib/kdtree.h:195
    .height() .num_nodes() .num_triangles() .box()                  lib/kdtree.h:197-221
* KDTreeIntersection(tree)          -- lib/kdtree.h:270-309
    .intersect(o, d) -> (id | None, r, s, t)                        lib/kdtree.cpp:515-578
    .intersect_batch(o, d)          -- batched twin (device tensors or host arrays)
* TracerConfig                      -- config.h:103-153 (+ pathtracer.h:4-25 flags)
* render(tree, camera, lights, conf) -- the main.cpp:187-226 pixel loop over trace()

Every call runs the HIP kernels; there is no CPU fallback.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import (TN_INVALID_ID, TN_RENDERER_PATHTRACER, TN_RENDERER_RAYCASTER, TN_RENDERER_RAYTRACER,
                   TurnerError, check, lib, ptr)

__all__ = ["KDTree", "KDTreeIntersection", "TracerConfig", "render", "TurnerError",
           "TN_INVALID_ID", "device_info", "build_tree", "tree_cache_save", "tree_cache_load",
           "write_ppm", "write_pfm", "read_pfm", "render_accumulate", "TN_RENDERER_PATHTRACER", "TN_RENDERER_RAYCASTER",
           "TN_RENDERER_RAYTRACER"]


def build_tree(verts, serial=False):
    """Host-only SAH build (no GPU): (reference FlatNode array, box, height)."""
    v = _f32(verts, 9)
    nodes = C.POINTER(C.c_uint64)()
    nn = C.c_size_t(0)
    box = np.zeros(6, np.float32)
    h = C.c_uint32(0)
    check(lib().tn_kdtree_build(ptr(v), len(v), _lib.TN_SCENE_SERIAL_BUILD if serial else 0,
                                C.byref(nodes), C.byref(nn), ptr(box), C.byref(h)))
    out = np.ctypeslib.as_array(nodes, shape=(nn.value,)).copy()
    lib().tn_buffer_free(nodes)
    return out, box, h.value


def triangle_records(verts):
    """Host precompute of the triangle records tn_scene_create uploads
    (lib/triangle.h:22-39): (N, 16) float32, v0 n u v uv vv uu denom."""
    v = _f32(verts, 9)
    out = np.zeros((len(v), 16), np.float32)
    check(lib().tn_triangle_records_host(ptr(v), len(v), ptr(out)))
    return out


def build_device_tree(verts, serial=False, empty_leaves=False, refined=False):
    """Host-only: the device layout tn_scene_create would upload -- nodes
    (M, 2) uint32 (inner: split bits, right << 2 | axis; leaf: offset,
    count << 2 | 3), refs (K,) uint32, box (6,), height."""
    v = _f32(verts, 9)
    pn = C.POINTER(C.c_uint32)()
    pr = C.POINTER(C.c_uint32)()
    nn, nr, h = C.c_size_t(), C.c_size_t(), C.c_uint32()
    box = np.zeros(6, np.float32)
    flags = (_lib.TN_SCENE_SERIAL_BUILD if serial else 0) | (_lib.TN_SCENE_EMPTY_LEAVES if empty_leaves else 0) | \
        (_lib.TN_SCENE_REFINED if refined else 0)
    check(lib().tn_kdtree_build_device(ptr(v), len(v), flags, C.byref(pn), C.byref(nn), C.byref(pr),
                                       C.byref(nr), ptr(box), C.byref(h)))
    try:
        nodes = np.ctypeslib.as_array(pn, shape=(nn.value * 2,)).copy().reshape(-1, 2)
        refs = np.ctypeslib.as_array(pr, shape=(max(nr.value, 1),)).copy()[:nr.value]
    finally:
        lib().tn_buffer_free(pn)
        lib().tn_buffer_free(pr)
    return nodes, refs, box, h.value


def tree_cache_save(path, verts, nodes, box):
    v = _f32(verts, 9)
    nodes = np.ascontiguousarray(nodes, np.uint64)
    box = np.ascontiguousarray(box, np.float32)
    check(lib().tn_tree_cache_save(path.encode(), ptr(v), len(v), ptr(nodes), len(nodes),
                                   ptr(box)))


def tree_cache_load(path, verts):
    """Returns (nodes, box) or raises TurnerError for a missing/stale/corrupt cache."""
    v = _f32(verts, 9)
    nodes = C.POINTER(C.c_uint64)()
    nn = C.c_size_t(0)
    box = np.zeros(6, np.float32)
    check(lib().tn_tree_cache_load(path.encode(), ptr(v), len(v), C.byref(nodes), C.byref(nn),
                                   ptr(box)))
    out = np.ctypeslib.as_array(nodes, shape=(nn.value,)).copy()
    lib().tn_buffer_free(nodes)
    return out, box


def device_info():
    name = C.create_string_buffer(256)
    cus = C.c_int(0)
    check(lib().tn_device_info(name, 256, C.byref(cus)))
    return name.value.decode(), cus.value


def _f32(a, cols):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a.reshape(-1, cols)


class KDTree:
    """Scene handle: triangles + SAH kd-tree resident in HBM.

    verts (N, 9) float32; normals (N, 9) or None (face normals); albedo (N, 4)
    or None (0.8 grey); emission (N, 4) or None.  `flat_nodes`/`box` upload a
    tree built earlier (e.g. from the tree cache) instead of building one.
    empty_leaves=True builds the rendering variant of the tree that keeps
    empty-space splits (TN_SCENE_EMPTY_LEAVES; no reference FlatNode array).
    refined=True adds the refined reference tree that renders traverse with
    the reference's hit ids (TN_SCENE_REFINED).
    """

    def __init__(self, verts, normals=None, albedo=None, emission=None, serial=False,
                 flat_nodes=None, box=None, empty_leaves=False, refined=False):
        self._v = _f32(verts, 9)
        self._n = _f32(normals, 9)
        self._a = _f32(albedo, 4)
        self._e = _f32(emission, 4)
        n = len(self._v)
        h = C.c_void_p()
        if flat_nodes is not None:
            fl = np.ascontiguousarray(flat_nodes, dtype=np.uint64)
            bx = np.ascontiguousarray(box, dtype=np.float32)
            check(lib().tn_scene_create_with_tree(ptr(self._v), ptr(self._n), ptr(self._a),
                                                  ptr(self._e), n, ptr(fl), len(fl), ptr(bx),
                                                  C.byref(h)))
        else:
            flags = (_lib.TN_SCENE_SERIAL_BUILD if serial else 0) | \
                (_lib.TN_SCENE_EMPTY_LEAVES if empty_leaves else 0) | (_lib.TN_SCENE_REFINED if refined else 0)
            check(lib().tn_scene_create(ptr(self._v), ptr(self._n), ptr(self._a), ptr(self._e),
                                        n, flags, C.byref(h)))
        self.handle = h
        self._info = _lib.SceneInfo()
        check(lib().tn_scene_get_info(self.handle, C.byref(self._info)))

    def set_reflection(self, reflective, reflectivity):
        """Reflective colour (N, 4) and reflectivity (N,) per triangle, read by
        the raytracer (raytracer.cpp:37-46; lib/triangle.h:23-33)."""
        self._rc = _f32(reflective, 4)
        self._rf = np.ascontiguousarray(reflectivity, np.float32).reshape(-1)
        if self._rc is None or len(self._rc) != len(self._v) or len(self._rf) != len(self._v):
            raise ValueError("one reflective colour and one reflectivity per triangle")
        check(lib().tn_scene_set_reflection(self.handle, ptr(self._rc), ptr(self._rf)))

    def close(self):
        if getattr(self, "handle", None):
            lib().tn_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # lib/kdtree.h:197-224
    def height(self):
        return self._info.height

    def num_nodes(self):
        return self._info.num_nodes

    def num_triangles(self):
        return self._info.num_triangles

    def box(self):
        return np.array(self._info.box, np.float32)

    @property
    def info(self):
        return self._info

    def flat_nodes(self):
        out = np.zeros(self._info.num_nodes, np.uint64)
        check(lib().tn_scene_get_flat_nodes(self.handle, ptr(out), len(out)))
        return out

    def triangle_records(self):
        out = np.zeros((self._info.num_triangles, 16), np.float32)
        check(lib().tn_scene_get_triangle_records(self.handle, ptr(out), len(out)))
        return out

    @property
    def verts(self):
        return self._v


class KDTreeIntersection:
    """lib/kdtree.h:270-309.  intersect() for one ray mirrors the reference's
    out-parameter call; intersect_batch() is the batched GPU twin."""

    def __init__(self, tree, flags=None):
        """flags: TN_ISECT_*; by default the refined traversal on a tree built
        with refined=True (the reference tree's hits, faster), else 0."""
        self.tree = tree
        if flags is None:
            flags = _lib.TN_ISECT_REFINED if tree.info.refined_nodes else 0
        self.flags = flags

    def intersect(self, o, d):
        ids, r, s, t = self.intersect_batch(np.asarray([o], np.float32),
                                            np.asarray([d], np.float32))
        if ids[0] == TN_INVALID_ID:
            return None, float(r[0]), 0.0, 0.0
        return int(ids[0]), float(r[0]), float(s[0]), float(t[0])

    def intersect_batch(self, o, d, flags=None):
        """o, d: (n, 3) float32 host arrays or device tensors.  Returns
        (id, r, s, t) of the same kind."""
        flags = self.flags if flags is None else flags
        if hasattr(o, "is_cuda") and o.is_cuda:
            import torch
            n = o.shape[0]
            ids = torch.empty(n, dtype=torch.int32, device=o.device)
            r = torch.empty(n, dtype=torch.float32, device=o.device)
            s = torch.empty_like(r)
            t = torch.empty_like(r)
            stream = torch.cuda.current_stream(o.device).cuda_stream
            check(lib().tn_intersect(self.tree.handle, ptr(o), ptr(d), n, ptr(ids), ptr(r),
                                     ptr(s), ptr(t), flags, stream))
            return ids, r, s, t
        o = np.ascontiguousarray(o, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(d, np.float32).reshape(-1, 3)
        n = len(o)
        ids = np.zeros(n, np.uint32)
        r = np.zeros(n, np.float32)
        s = np.zeros(n, np.float32)
        t = np.zeros(n, np.float32)
        check(lib().tn_intersect_host(self.tree.handle, ptr(o), ptr(d), n, ptr(ids), ptr(r),
                                      ptr(s), ptr(t), flags))
        return ids, r, s, t


@dataclass
class TracerConfig:
    """config.h:45-60,103-153 with the pathtracer USAGE defaults (pathtracer.h:4-25)."""
    width: int = 640
    aspect: float = 1.0
    bg_color: tuple = (0.0, 0.0, 0.0, 1.0)
    num_threads: int = 1
    inverse_gamma: float = 0.454545
    exposure: float = 1.0
    gamma_correction_enabled: bool = True
    max_recursion_depth: int = 3
    num_pixel_samples: int = 1
    num_monte_carlo_samples: int = 8
    # extensions (no reference counterpart)
    flags: int = 0
    seed: int = 1
    chunk_paths: int = 0
    # the other trace() plugins (TN_RENDERER_*; raycaster.h / raytracer.h defaults)
    renderer: int = 0
    max_visibility: float = 2.0
    shadow_intensity: float = 0.5

    def height(self):
        return int(self.width / self.aspect)

    def check(self):  # config.h:29-33,119-126
        assert 0 < self.aspect and 1 <= self.num_threads and 0 <= self.exposure
        assert 0 < self.max_recursion_depth and 1 <= self.num_pixel_samples
        assert 0 <= self.num_monte_carlo_samples
        assert 0 <= self.max_visibility and 0 <= self.shadow_intensity <= 1


def _camera(cam):
    c = _lib.Camera()
    c.pos[:] = [float(x) for x in np.asarray(cam["pos"]).ravel()]
    c.trafo[:] = [float(x) for x in np.asarray(cam["trafo"]).ravel()]
    c.delta_x = float(cam["delta_x"])
    c.delta_y = float(cam["delta_y"])
    return c


def _lights(lights):
    arr = (_lib.Light * max(1, len(lights)))()
    for k, (p, col) in enumerate(lights):
        arr[k].pos[:] = [float(x) for x in p]
        arr[k].color[:] = [float(x) for x in col]
    return arr


def _params(conf, width, height):
    p = _lib.RenderParams()
    p.width, p.height = width, height
    p.max_depth = conf.max_recursion_depth
    p.mc_samples = conf.num_monte_carlo_samples
    p.pixel_samples = conf.num_pixel_samples
    p.flags = conf.flags
    p.seed = conf.seed
    p.bg[:] = [float(x) for x in conf.bg_color]
    p.exposure = conf.exposure
    p.inverse_gamma = conf.inverse_gamma
    p.gamma_enabled = 1 if conf.gamma_correction_enabled else 0
    p.chunk_paths = conf.chunk_paths
    p.renderer = conf.renderer
    p.max_visibility = conf.max_visibility
    p.shadow_intensity = conf.shadow_intensity
    return p


def render(tree, cam, lights, conf, display=False):
    """Renders on the GPU.  Returns (linear RGBA (h, w, 4) mean over pixel
    samples before exposure, stats dict[, display RGBA after exposure/gamma])."""
    conf.check()
    w, h = conf.width, conf.height()
    lin = np.zeros((h, w, 4), np.float32)
    disp = np.zeros((h, w, 4), np.float32) if display else None
    st = _lib.RenderStats()
    ls = _lights(lights)
    check(lib().tn_render(tree.handle, C.byref(_camera(cam)), ls, len(lights),
                          C.byref(_params(conf, w, h)), ptr(lin), ptr(disp), C.byref(st)))
    stats = {"rays": st.rays, "primary_rays": st.primary_rays, "shadow_rays": st.shadow_rays,
             "ms": st.ms}
    return (lin, stats, disp) if display else (lin, stats)


def render_accumulate(tree, cam, lights, conf, accum, sample_offset=0):
    """tn_render_accumulate: adds the per-pixel radiance SUMS of
    conf.num_pixel_samples samples (pixel-sample indices from sample_offset,
    fresh random streams per offset) into accum, a (h, w, 4) float32 device
    tensor.  Enqueued on torch's current stream; the call returns once the
    render has finished (its stats and the queue-overflow check need the
    device counters), so the sums are complete on return."""
    import torch
    w, h = conf.width, conf.height()
    if not (isinstance(accum, torch.Tensor) and accum.is_cuda and accum.dtype == torch.float32
            and accum.is_contiguous() and accum.numel() == w * h * 4):
        raise ValueError("accum must be a contiguous float32 device tensor of h * w * 4")
    p = _params(conf, w, h)
    p.sample_offset = sample_offset
    st = _lib.RenderStats()
    stream = torch.cuda.current_stream(accum.device).cuda_stream
    check(lib().tn_render_accumulate(tree.handle, C.byref(_camera(cam)), _lights(lights), len(lights),
                                     C.byref(p), ptr(accum), C.byref(st), stream))
    return {"rays": st.rays, "primary_rays": st.primary_rays, "shadow_rays": st.shadow_rays,
            "ms": st.ms}


def trace(tree, o, d, lights, depth, conf, sample_offset=0):
    """trace(ray, tree_intersection, lights, depth, conf) of trace.h:23-25 for
    a batch of rays (tn_trace): o, d (n, 3) float32 host arrays; returns the
    radiance of each ray entering at recursion level `depth` as (n, 4) RGBA
    and the stats.  Ray j draws its streams from (conf.seed, sample_offset, j)."""
    import torch
    o, d = _f32(o, 3), _f32(d, 3)
    n = len(o)
    dev = torch.device("cuda:0")
    to, td = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
    rgba = torch.zeros(n * 4, device=dev)
    p = _params(conf, 1, 1)
    p.sample_offset = sample_offset
    st = _lib.RenderStats()
    check(lib().tn_trace(tree.handle, ptr(to), ptr(td), n, depth, _lights(lights), len(lights), C.byref(p),
                         ptr(rgba), C.byref(st), torch.cuda.current_stream(dev).cuda_stream))
    torch.cuda.current_stream(dev).synchronize()
    return rgba.cpu().numpy().reshape(n, 4), {"rays": st.rays, "shadow_rays": st.shadow_rays}


def write_pfm(path, rgb):
    """Portable float map of an (h, w, 3|4) image: "PF", width height, -1.0
    (little-endian), FP32 RGB rows bottom first -- the linear images the
    relMSE measurements compare."""
    a = np.asarray(rgb, np.float32)[..., :3]
    h, w = a.shape[:2]
    with open(path, "wb") as f:
        f.write(f"PF\n{w} {h}\n-1.0\n".encode())
        f.write(np.ascontiguousarray(a[::-1], "<f4").tobytes())


def read_pfm(path):
    """(h, w, 3) float32 image of a colour PF file (either byte order)."""
    with open(path, "rb") as f:
        data = f.read()
    toks, pos = [], 0
    while len(toks) < 4:  # PF, width, height, scale: whitespace-separated
        while data[pos:pos + 1].isspace():
            pos += 1
        end = pos
        while end < len(data) and not data[end:end + 1].isspace():
            end += 1
        toks.append(data[pos:end].decode())
        pos = end
    pos += 1  # one whitespace byte ends the header
    if toks[0] != "PF":
        raise ValueError(f"{path} is not a colour PFM")
    w, h, scale = int(toks[1]), int(toks[2]), float(toks[3])
    img = np.frombuffer(data, "<f4" if scale < 0 else ">f4", count=w * h * 3, offset=pos).reshape(h, w, 3)
    return img[::-1].astype(np.float32)


def write_ppm(rgba):
    """PPM P3 text of an (h, w, 4) image (lib/raster.h:79-100)."""
    rgba = np.ascontiguousarray(rgba, np.float32)
    h, w = rgba.shape[:2]
    n = check(lib().tn_write_ppm(ptr(rgba), w, h, None, 0))
    buf = C.create_string_buffer(n + 1)
    check(lib().tn_write_ppm(ptr(rgba), w, h, buf, n + 1))
    return buf.value.decode()

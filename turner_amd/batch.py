"""Batched heterogeneous rendering (tn_batch_* of include/turner_amd.h): many
scenes packed into one kd-tree forest and rendered through one wavefront
launch sequence per chunk of samples, with an optional hipGraph replay of
the chunk sequence and an autotuner for chunk size and grid.  No reference
counterpart (the reference renders one scene per process, main.cpp:88-244).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import TN_RENDER_AUTOTUNE, TN_RENDER_GRAPH, TN_RENDER_PERSISTENT, check, lib, ptr

__all__ = ["Batch", "heterogeneous_workload", "TN_RENDER_AUTOTUNE", "TN_RENDER_GRAPH", "TN_RENDER_PERSISTENT"]


def heterogeneous_workload(seed=2024, count=32):
    """The batched-rendering workload of 32 independent scenes: scene 0 the
    white furnace (unit-radiance enclosing emitter, level-4 sphere of albedo
    0.8, whose pixels see exactly 0.8 where a camera ray hits the sphere);
    scenes 1..count-1 seeded displaced icospheres of subdivision level
    uniform in 3..8 (1,280 to 1.31 M triangles, log-uniform in the count) in
    the Cornell enclosure with a 0.2 x 0.2 ceiling emitter of radiance 15;
    square images of width uniform in 128..1024 and 16..64 pixel samples.
    Returns (scenes, levels, jobs) with one job dict per scene."""
    from . import scenes
    rng = np.random.default_rng(seed)
    scs = [scenes.furnace(level=4, rho=0.8, radiance=1.0)]
    levels = [4]
    for k in range(1, count):
        lv = int(rng.integers(3, 9))
        levels.append(lv)
        scs.append(scenes.sphere_in_cornell(level=lv, seed=k))
    jobs = []
    for sc in scs:
        w = int(rng.integers(128, 1025))
        jobs.append({"camera": sc["camera"], "lights": sc["lights"], "width": w, "height": w,
                     "spp": int(rng.integers(16, 65))})
    return scs, levels, jobs


class Batch:
    """Packs the given KDTree objects (their scenes) into one forest."""

    def __init__(self, trees):
        arr = (C.c_void_p * len(trees))(*[t.handle.value if hasattr(t.handle, "value") else t.handle
                                         for t in trees])
        h = C.c_void_p()
        check(lib().tn_batch_create(arr, len(trees), C.byref(h)))
        self.handle = h
        self.n = len(trees)

    def close(self):
        if getattr(self, "handle", None):
            lib().tn_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, jobs, max_depth=3, mc_samples=1, flags=0, seed=1, chunk_paths=0, sample_offset=0,
               accum=None, renderer=0):
        """jobs: one dict per scene with camera, lights (<= 1), width, height,
        spp and optional bg.  Returns (list of (h, w, 4) device tensors of
        per-pixel MEANS, stats dict, accum tensor of the sums)."""
        import torch
        from . import _camera
        if len(jobs) != self.n:
            raise ValueError("one job per scene")
        J = (_lib.BatchJob * self.n)()
        npix = 0
        for k, j in enumerate(jobs):
            J[k].cam = _camera(j["camera"])
            ls = j.get("lights", [])
            if len(ls) > 1:
                raise ValueError("at most one point light per scene")
            J[k].num_lights = len(ls)
            if ls:
                J[k].light.pos[:] = [float(x) for x in ls[0][0]]
                J[k].light.color[:] = [float(x) for x in ls[0][1]]
            J[k].width, J[k].height = j["width"], j["height"]
            J[k].pixel_samples = j["spp"]
            J[k].bg[:] = [float(x) for x in j.get("bg", (0.0, 0.0, 0.0, 1.0))]
            npix += j["width"] * j["height"]
        if accum is None:
            accum = torch.zeros(npix * 4, dtype=torch.float32, device="cuda")
        p = _lib.RenderParams()
        p.max_depth, p.mc_samples, p.flags, p.seed = max_depth, mc_samples, flags, seed
        p.chunk_paths, p.sample_offset = chunk_paths, sample_offset
        p.renderer = renderer  # batches render with the path tracer only
        st = _lib.RenderStats()
        stream = torch.cuda.current_stream(accum.device).cuda_stream
        check(lib().tn_batch_render(self.handle, J, C.byref(p), ptr(accum), C.byref(st), stream))
        imgs, off = [], 0
        for j in jobs:
            n = j["width"] * j["height"] * 4
            imgs.append(accum[off:off + n].view(j["height"], j["width"], 4) / j["spp"])
            off += n
        return imgs, {"rays": st.rays, "primary_rays": st.primary_rays, "shadow_rays": st.shadow_rays,
                      "ms": st.ms}, accum

    def tuning(self):
        t = _lib.BatchTuning()
        check(lib().tn_batch_get_tuning(self.handle, C.byref(t)))
        return {"chunk_paths": t.chunk_paths, "blocks_per_cu": t.blocks_per_cu, "graph": bool(t.graph),
                "candidates": t.candidates, "probe_ms": t.probe_ms, "best_rate": t.best_rate}

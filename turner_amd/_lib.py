"""ctypes binding of libturner_amd.so (C ABI in include/turner_amd.h).

The library is built in-tree by turner_amd/build.py.  Loading fails loudly
when it is missing: there is no CPU fallback for any entry point.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libturner_amd.so")

TN_OK = 0
TN_ERR_INVALID = -1
TN_ERR_OOM = -2
TN_ERR_HIP = -3
TN_ERR_QUEUE = -4
TN_ERR_NODEVICE = -5
TN_ERR_LIMIT = -6
TN_ERR_INCOMPLETE = -7
TN_INVALID_ID = 1 << 30
TN_GUIDE_PARAMS = 15976
TN_GUIDE_OUTPUTS = 40

TN_SCENE_SERIAL_BUILD = 1
TN_SCENE_EMPTY_LEAVES = 2
TN_SCENE_REFINED = 4
TN_ISECT_EXHAUSTIVE = 1
TN_ISECT_MT = 2
TN_ISECT_COUNT = 4
TN_ISECT_TICK = 8
TN_ISECT_LEAF = 16
TN_ISECT_REFINED = 32
TN_RENDER_COSINE = 1
TN_RENDER_AREA_LIGHTS = 2
TN_RENDER_REF_JITTER = 4
TN_RENDER_GUIDED = 8
TN_RENDER_EXHAUSTIVE = 16
TN_RENDER_GRAPH = 32
TN_RENDER_AUTOTUNE = 64
TN_RENDER_PERSISTENT = 128
TN_RENDER_DEFER_TRAINING = 256
TN_RENDERER_PATHTRACER = 0
TN_RENDERER_RAYCASTER = 1
TN_RENDERER_RAYTRACER = 2


class TurnerError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"turner_amd error {code}: {msg}")
        self.code = code


class SceneInfo(C.Structure):
    _fields_ = [("num_triangles", C.c_uint32), ("num_nodes", C.c_uint32),
                ("height", C.c_uint32), ("device_nodes", C.c_uint32),
                ("device_refs", C.c_uint32), ("num_emitters", C.c_uint32),
                ("box", C.c_float * 6), ("build_ms", C.c_double),
                ("refined_nodes", C.c_uint32), ("refined_refs", C.c_uint32)]


class Camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("trafo", C.c_float * 9),
                ("delta_x", C.c_float), ("delta_y", C.c_float)]


class Light(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("color", C.c_float * 4)]


class RenderParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("max_depth", C.c_int32),
                ("mc_samples", C.c_int32), ("pixel_samples", C.c_int32), ("flags", C.c_uint32),
                ("seed", C.c_uint64), ("bg", C.c_float * 4), ("exposure", C.c_float),
                ("inverse_gamma", C.c_float), ("gamma_enabled", C.c_int32),
                ("chunk_paths", C.c_uint32), ("sample_offset", C.c_uint32),
                ("renderer", C.c_uint32), ("max_visibility", C.c_float), ("shadow_intensity", C.c_float),
                ("reserved", C.c_uint32)]


class BatchJob(C.Structure):
    _fields_ = [("cam", Camera), ("light", Light), ("num_lights", C.c_uint32), ("width", C.c_uint32),
                ("height", C.c_uint32), ("pixel_samples", C.c_int32), ("bg", C.c_float * 4)]


class BatchTuning(C.Structure):
    _fields_ = [("chunk_paths", C.c_uint32), ("blocks_per_cu", C.c_uint32), ("graph", C.c_uint32),
                ("candidates", C.c_uint32), ("probe_ms", C.c_double), ("best_rate", C.c_double)]


class RenderStats(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("primary_rays", C.c_uint64),
                ("shadow_rays", C.c_uint64), ("ms", C.c_double)]


_P = C.c_void_p
_FP = C.POINTER(C.c_float)
_U32 = C.c_uint32

# name -> (restype, argtypes); every symbol declared in include/turner_amd.h
SIGNATURES = {
    "tn_scene_create": (C.c_int, [_P, _P, _P, _P, _U32, _U32, C.POINTER(_P)]),
    "tn_scene_create_with_tree": (C.c_int, [_P, _P, _P, _P, _U32, _P, C.c_size_t, _P,
                                            C.POINTER(_P)]),
    "tn_scene_set_reflection": (C.c_int, [_P, _P, _P]),
    "tn_scene_destroy": (C.c_int, [_P]),
    "tn_scene_get_info": (C.c_int, [_P, C.POINTER(SceneInfo)]),
    "tn_scene_get_flat_nodes": (C.c_int, [_P, _P, C.c_size_t]),
    "tn_scene_get_triangle_records": (C.c_int, [_P, _P, C.c_size_t]),
    "tn_intersect": (C.c_int, [_P, _P, _P, _U32, _P, _P, _P, _P, _U32, _P]),
    "tn_intersect_noalloc": (C.c_int, [_P, _P, _P, _U32, _P, _P, _P, _P, _U32, _P]),
    "tn_intersect_host": (C.c_int, [_P, _P, _P, _U32, _P, _P, _P, _P, _U32]),
    "tn_occluded": (C.c_int, [_P, _P, _P, _P, _U32, _P, _U32, _P]),
    "tn_get_counters": (C.c_int, [_P, _P, C.c_int]),
    "tn_render": (C.c_int, [_P, C.POINTER(Camera), C.POINTER(Light), _U32,
                            C.POINTER(RenderParams), _P, _P, C.POINTER(RenderStats)]),
    "tn_render_accumulate": (C.c_int, [_P, C.POINTER(Camera), C.POINTER(Light), _U32,
                                       C.POINTER(RenderParams), _P, C.POINTER(RenderStats), _P]),
    "tn_trace": (C.c_int, [_P, _P, _P, _U32, C.c_int32, C.POINTER(Light), _U32,
                           C.POINTER(RenderParams), _P, C.POINTER(RenderStats), _P]),
    "tn_tonemap": (C.c_int, [_P, _P, _U32, C.c_float, C.c_float, C.c_float, C.c_int, _P]),
    "tn_write_ppm": (C.c_long, [_P, _U32, _U32, _P, C.c_size_t]),
    "tn_kdtree_build": (C.c_int, [_P, _U32, _U32, C.POINTER(C.POINTER(C.c_uint64)),
                                  C.POINTER(C.c_size_t), _P, C.POINTER(C.c_uint32)]),
    "tn_buffer_free": (None, [_P]),
    "tn_triangle_records_host": (C.c_int, [_P, _U32, _P]),
    "tn_kdtree_build_device": (C.c_int, [_P, _U32, _U32, C.POINTER(C.POINTER(C.c_uint32)),
                                         C.POINTER(C.c_size_t), C.POINTER(C.POINTER(C.c_uint32)),
                                         C.POINTER(C.c_size_t), _P, C.POINTER(C.c_uint32)]),
    "tn_tree_cache_save": (C.c_int, [C.c_char_p, _P, _U32, _P, C.c_size_t, _P]),
    "tn_tree_cache_load": (C.c_int, [C.c_char_p, _P, _U32, C.POINTER(C.POINTER(C.c_uint64)),
                                     C.POINTER(C.c_size_t), _P]),
    "tn_guide_create": (C.c_int, [_P, _P, C.c_uint64, C.POINTER(_P)]),
    "tn_guide_destroy": (C.c_int, [_P]),
    "tn_guide_get_params": (C.c_int, [_P, _P]),
    "tn_guide_set_params": (C.c_int, [_P, _P]),
    "tn_guide_set_options": (C.c_int, [_P, C.c_float, _U32, C.c_float, _U32]),
    "tn_guide_forward": (C.c_int, [_P, _P, _P, _U32, _P, _P]),
    "tn_guide_pdf": (C.c_int, [_P, _P, _P, _P, _U32, _P, _P]),
    "tn_guide_sample": (C.c_int, [_P, _P, _P, _P, _U32, _P, _P, _P]),
    "tn_guide_loss_grad": (C.c_int, [_P, _P, _P, _P, _P, _U32, _P, _FP, _P]),
    "tn_guide_train_step": (C.c_int, [_P, _P, _P, _P, _P, _U32, _FP, _P]),
    "tn_guide_info": (C.c_int, [_P, C.POINTER(_U32), C.POINTER(_U32), _FP]),
    "tn_guide_get_train_set": (C.c_long, [_P, _P, C.c_size_t]),
    "tn_guide_train_recorded": (C.c_int, [_P, _U32, _P]),
    "tn_guide_flush": (C.c_int, [_P, _P]),
    "tn_render_guided": (C.c_int, [_P, C.POINTER(Camera), C.POINTER(Light), _U32,
                                   C.POINTER(RenderParams), _P, _U32, _P,
                                   C.POINTER(RenderStats), _P]),
    "tn_batch_create": (C.c_int, [_P, _U32, C.POINTER(_P)]),
    "tn_batch_destroy": (C.c_int, [_P]),
    "tn_batch_render": (C.c_int, [_P, _P, C.POINTER(RenderParams), _P, C.POINTER(RenderStats), _P]),
    "tn_batch_get_tuning": (C.c_int, [_P, C.POINTER(BatchTuning)]),
    "tn_sample_triangle_host": (C.c_int, [_P, _P, _P, _U32, _P, _FP, _FP]),
    "tn_last_error": (C.c_char_p, []),
    "tn_device_info": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(C.c_int)]),
}

_lib = None


def lib():
    """Loads libturner_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                              "(python turner_amd/build.py) first")
        # Share one HIP runtime with torch: torch loads its own libamdhip64
        # (SONAME libamdhip64.so.7) through a NEEDED entry that names
        # libamdhip64.so, so if this library were loaded first the process
        # would end up with two HIP runtimes.  Importing torch first makes
        # this library's NEEDED libamdhip64.so.7 resolve to torch's copy.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc < 0:
        msg = lib().tn_last_error()
        raise TurnerError(rc, msg.decode() if msg else "")
    return rc


def ptr(a):
    """Address of a numpy array or torch tensor (None -> NULL)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data

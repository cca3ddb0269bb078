"""ctypes declarations of the C ABI in include/spt.h (libspt.so).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C swpathtracer_amd/csrc``). Loading fails loudly when it is missing:
there is no CPU fallback for the product path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspt.so")

SPT_OK = 0
SPT_ERR_INVALID = 1
SPT_ERR_DEVICE = 2
SPT_ERR_NOMEM = 3
SPT_ERR_STATE = 4
SPT_ERR_UNSUPPORTED = 5
SPT_ERR_OVERFLOW = 6
SPT_MISS = 0xFFFFFFFF

SPT_BSDF_LAMBERTIAN = 1
SPT_BSDF_COOK_TORRANCE = 2
SPT_BSDF_SMOOTH_CONDUCTOR = 3
SPT_BSDF_SMOOTH_DIELECTRIC = 4
SPT_BSDF_GGX_CONDUCTOR = 5
SPT_PRIM_TRIANGLE = 0
SPT_PRIM_SPHERE = 1
SPT_SAMPLER_IMPORTANCE = 0
SPT_SAMPLER_UNIFORM = 1
SPT_MAX_LAYERS = 4
SPT_CAMERA_PINHOLE = 0
SPT_CAMERA_ORTHOGRAPHIC = 1
SPT_PIXEL_REGULAR = 0
SPT_PIXEL_JITTERED = 1
SPT_PIXEL_UNIFORM = 2
SPT_RNG_COUNTER = 0
SPT_RNG_STREAM = 1
SPT_RENDER_FILM32 = 1 << 0
SPT_RENDER_FILM64 = 1 << 1
SPT_RENDER_PER_PATH = 1 << 2
SPT_RENDER_NO_SORT = 1 << 3
SPT_RENDER_GUIDED = 1 << 4
SPT_RENDER_FETCH_STATS = 1 << 5
SPT_RENDER_TIME_KERNELS = 1 << 6
SPT_RENDER_GUIDE_TRAIN = 1 << 7
SPT_GUIDE_COATED = 1 << 0
SPT_GUIDE_PRODUCT = 1 << 1
SPT_GUIDE_PARAMS = 13552
SPT_GUIDE_RECORD_FLOATS = 24
SPT_TRACE_CLOSEST = 0
SPT_TRACE_COUNT = 1
SPT_TRACE_UNPRUNED = 2
SPT_TRACE_FETCH = 3
SPT_RAYCAST_NORMAL = 0
SPT_RAYCAST_DEPTH = 1
SPT_RAYCAST_COUNT = 2
SPT_NOALLOC_SCRATCH_WORDS = 8 * 1024 * 1024
SPT_NOALLOC_MAX_RAYS = 0xFFFFFF00
SPT_LUA_NONE = 0
SPT_LUA_PATHTRACER = 1
SPT_LUA_RAYCAST_NORMAL = 2
SPT_LUA_RAYCAST_DEPTH = 3
SPT_LUA_RAYCAST_COUNT = 4
SPT_ACCEL_NONE = 0
SPT_ACCEL_BVH_SAH = 1
SPT_ACCEL_SBVH_SAH = 2
SPT_LUA_BUILTIN_CORE = 1

F3 = C.c_float * 3


class Layer(C.Structure):
    _fields_ = [("type", C.c_uint32), ("sampler", C.c_uint32), ("kd", F3), ("r0", F3),
                ("roughness", C.c_float), ("ior", C.c_float), ("eta", F3), ("k", F3)]


class Material(C.Structure):
    _fields_ = [("n_layers", C.c_uint32), ("emission", F3), ("layers", Layer * SPT_MAX_LAYERS)]


class Camera(C.Structure):
    _fields_ = [("type", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32), ("pad", C.c_uint32),
                ("position", F3), ("look_at", F3), ("up", F3), ("fov_deg", C.c_float),
                ("min_x", C.c_float), ("max_x", C.c_float), ("min_y", C.c_float), ("max_y", C.c_float)]


class RenderParams(C.Structure):
    _fields_ = [("pixel_sampler", C.c_uint32), ("spp", C.c_uint32), ("max_depth", C.c_int32),
                ("rr_depth", C.c_int32), ("seed", C.c_uint64), ("spp_begin", C.c_uint32),
                ("spp_count", C.c_uint32), ("viewport", C.c_uint32 * 4), ("background", F3),
                ("flags", C.c_uint32), ("rng", C.c_uint32), ("stream_stride", C.c_uint32),
                ("d_stream", C.c_void_p)]


class BvhInfo(C.Structure):
    _fields_ = [("n_prims", C.c_uint64), ("n_nodes", C.c_uint64), ("n_inner", C.c_uint64),
                ("n_leaves", C.c_uint64), ("tree_height", C.c_uint64), ("n_references", C.c_uint64),
                ("min_leaf", C.c_uint64), ("max_leaf", C.c_uint64), ("n_pair_nodes", C.c_uint64),
                ("max_depth", C.c_uint64), ("root_min", F3), ("root_max", F3), ("build_ms", C.c_double)]


class RenderStats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("rays", C.c_uint64), ("bounces", C.c_uint64),
                ("stream_overflows", C.c_uint64), ("ms", C.c_double), ("extend_ms", C.c_double),
                ("extend_launches", C.c_uint64), ("node_fetches", C.c_uint64), ("tri_fetches", C.c_uint64),
                ("guide_ms", C.c_double), ("guide_queries", C.c_uint64)]


class GuideConfig(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("learning_rate", C.c_float), ("loss_scale", C.c_float),
                ("bsdf_fraction", C.c_float), ("batch", C.c_uint32), ("steps_per_pass", C.c_uint32),
                ("flags", C.c_uint32), ("train_every", C.c_uint32)]


class GuideStats(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("skipped_steps", C.c_uint64), ("records", C.c_uint64),
                ("last_loss", C.c_float), ("loss_scale", C.c_float), ("train_ms", C.c_double),
                ("train_kernel_ms", C.c_double), ("trained_records", C.c_uint64), ("candidates", C.c_uint64),
                ("record_stride", C.c_uint32), ("pad", C.c_uint32)]


class Tuning(C.Structure):
    _fields_ = [("path_capacity", C.c_uint32), ("extend_blocks_per_cu", C.c_uint32),
                ("shade_blocks_per_cu", C.c_uint32), ("use_graph", C.c_uint32), ("tile_pixels", C.c_uint32),
                ("tile_samples", C.c_uint32), ("check_every", C.c_uint32), ("refill_lanes", C.c_uint32),
                ("wavefronts", C.c_uint32)]


class BatchScene(C.Structure):
    _fields_ = [("bvh", C.c_void_p), ("verts", C.c_void_p), ("normals", C.c_void_p), ("mat_ids", C.c_void_p),
                ("mats", C.POINTER(Material)), ("n_mats", C.c_uint32), ("pad", C.c_uint32),
                ("camera", Camera), ("params", RenderParams), ("d_rgb32", C.c_void_p), ("d_rgb64", C.c_void_p),
                ("d_path_rgb", C.c_void_p)]


class BatchStats(C.Structure):
    _fields_ = [("scenes", C.c_uint64), ("paths", C.c_uint64), ("rays", C.c_uint64), ("bounces", C.c_uint64),
                ("ms", C.c_double), ("tiles", C.c_uint64), ("triangles", C.c_uint64),
                ("device_bytes", C.c_uint64)]


class LuaInfo(C.Structure):
    _fields_ = [("camera", Camera), ("params", RenderParams), ("integrator", C.c_uint32), ("accel", C.c_uint32),
                ("overlap_threshold", C.c_float), ("min_value", C.c_float), ("max_value", C.c_float),
                ("path_length", C.c_int32), ("n_prims", C.c_uint64), ("n_materials", C.c_uint64),
                ("n_objects", C.c_uint64), ("n_bsdfs", C.c_uint64), ("file_name", C.c_char * 256)]


assert C.sizeof(Layer) == 64 and C.sizeof(Material) == 272

# name -> (restype, argtypes); every entry point declared in include/spt.h
PROTOTYPES = {
    "spt_bvh_build": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)]),
    "spt_bvh_build_prims": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)]),
    "spt_bvh_get_info": (C.c_int32, [C.c_void_p, C.POINTER(BvhInfo)]),
    "spt_bvh_export": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "spt_bvh_dump": (C.c_int32, [C.c_void_p, C.c_char_p]),
    "spt_bvh_wide": (C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64), C.c_void_p, C.POINTER(C.c_uint64)]),
    "spt_bvh_destroy": (None, [C.c_void_p]),
    "spt_create": (C.c_int32, [C.c_int, C.POINTER(C.c_void_p)]),
    "spt_destroy": (None, [C.c_void_p]),
    "spt_last_error": (C.c_char_p, [C.c_void_p]),
    "spt_sync": (C.c_int32, [C.c_void_p]),
    "spt_stream": (C.c_void_p, [C.c_void_p]),
    "spt_malloc": (C.c_int32, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    "spt_free": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "spt_memcpy_h2d": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "spt_memcpy_d2h": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "spt_memcpy_d2d": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "spt_set_scene": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "spt_upload_scene": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                                     C.c_void_p, C.c_void_p]),
    "spt_set_materials": (C.c_int32, [C.c_void_p, C.POINTER(Material), C.c_uint32]),
    "spt_trace": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint32, C.c_void_p,
                              C.c_void_p]),
    "spt_camera_rays": (C.c_int32, [C.c_void_p, C.POINTER(Camera), C.c_void_p]),
    "spt_raycast": (C.c_int32, [C.c_void_p, C.POINTER(Camera), C.c_uint32, C.c_void_p]),
    "spt_render": (C.c_int32, [C.c_void_p, C.POINTER(Camera), C.POINTER(RenderParams), C.c_void_p,
                               C.c_void_p, C.c_void_p]),
    "spt_render_get_stats": (C.c_int32, [C.c_void_p, C.POINTER(RenderStats)]),
    "spt_event_record": (C.c_int32, [C.c_void_p, C.c_int]),
    "spt_event_elapsed": (C.c_int32, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_float)]),
    "spt_guide_init": (C.c_int32, [C.c_void_p, C.POINTER(GuideConfig)]),
    "spt_guide_set_params": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "spt_guide_get_params": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "spt_guide_eval": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "spt_guide_train": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_float)]),
    "spt_guide_gradient": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.POINTER(C.c_float)]),
    "spt_guide_train_collected": (C.c_int32, [C.c_void_p, C.POINTER(C.c_float)]),
    "spt_guide_get_stats": (C.c_int32, [C.c_void_p, C.POINTER(GuideStats)]),
    "spt_guide_records": (C.c_int32, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_uint32)]),
    "spt_set_tuning": (C.c_int32, [C.c_void_p, C.POINTER(Tuning)]),
    "spt_get_tuning": (C.c_int32, [C.c_void_p, C.POINTER(Tuning)]),
    "spt_autotune": (C.c_int32, [C.c_void_p, C.POINTER(Camera), C.POINTER(RenderParams), C.c_uint64,
                                 C.POINTER(Tuning)]),
    "spt_batch_create": (C.c_int32, [C.c_void_p, C.POINTER(BatchScene), C.c_uint32, C.POINTER(C.c_void_p)]),
    "spt_batch_render": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "spt_batch_get_stats": (C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(BatchStats), C.c_void_p]),
    "spt_batch_autotune": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(Tuning)]),
    "spt_batch_destroy": (None, [C.c_void_p]),
    "spt_trace_noalloc": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "spt_lua_load": (C.c_int32, [C.c_char_p, C.c_char_p, C.c_uint32, C.POINTER(C.c_void_p), C.c_char_p,
                                 C.c_size_t]),
    "spt_lua_get_info": (C.c_int32, [C.c_void_p, C.POINTER(LuaInfo)]),
    "spt_lua_get_prims": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "spt_lua_get_materials": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "spt_lua_registry_json": (C.c_int32, [C.c_void_p, C.c_char_p, C.POINTER(C.c_uint64)]),
    "spt_lua_printed": (C.c_int32, [C.c_void_p, C.c_char_p, C.POINTER(C.c_uint64)]),
    "spt_lua_destroy": (None, [C.c_void_p]),
}

_lib = None


def load() -> C.CDLL:
    """Loads libspt.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "or `make -C swpathtracer_amd/csrc`")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib

"""ctypes view of the C ABI in include/tinytrace.h.

Loads the in-tree ``libtinytrace_amd.so`` (built by ``__graft_entry__.build``
or ``make -C tinytrace_amd/csrc``) and raises if it is missing: there is no
fallback path.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_HERE, "libtinytrace_amd.so")

TT_OK = 0
STATUS_NAMES = {
    0: "TT_OK", 1: "TT_NO_DEVICE", 2: "TT_HIP_ERROR", 3: "TT_KERNEL_LOAD", 4: "TT_INVALID_ARGUMENT",
    5: "TT_STALE_HANDLE", 6: "TT_OUT_OF_MEMORY", 7: "TT_OVER_BUDGET", 8: "TT_UNSUPPORTED",
    9: "TT_INTERNAL_ERROR",
}
TT_INVALID_INDEX = 0xFFFFFFFF
TT_TRACE_CULL_BACK_FACES = 1
TT_TRACE_STATIC_MAX_RAYS = 65536  # tinytrace.h
TT_RENDER_COLLECT_STATS = 1
TT_RNG_REFERENCE = 0
TT_RNG_PCG = 1

F16 = C.c_float * 16


class Constants(C.Structure):
    _fields_ = [("view", F16), ("proj", F16), ("proj_view", F16), ("inverse_view", F16), ("inverse_proj", F16),
                ("camera_position", C.c_float * 4), ("frame_index", C.c_uint32),
                ("accumulated_frame_count", C.c_uint32), ("sample_count", C.c_uint32),
                ("bounce_count", C.c_uint32), ("screen_size", C.c_uint32 * 2), ("tonemap", C.c_uint32),
                ("light_sampling", C.c_uint32)]


class MaterialConstants(C.Structure):
    _fields_ = [(n, C.c_float * 4) for n in ("base_color", "specular_color", "transmission_color", "coat_color",
                                            "fuzz_color", "emission_color", "geometry_normal")] + \
               [(n, C.c_float) for n in ("base_weight", "base_metalness", "base_diffuse_roughness",
                                         "specular_weight", "specular_roughness", "specular_roughness_anisotropy",
                                         "specular_ior", "specular_rotation", "transmission_weight",
                                         "transmission_depth", "coat_weight", "coat_roughness",
                                         "coat_roughness_anisotropy", "coat_ior", "coat_darkening", "coat_rotation",
                                         "fuzz_weight", "fuzz_roughness", "emission_luminance", "geometry_opacity")]


class SceneDesc(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("position_count", C.c_uint32),
                ("vertices", C.c_void_p), ("vertex_count", C.c_uint32),
                ("indices", C.c_void_p), ("index_count", C.c_uint32),
                ("meshes", C.c_void_p), ("mesh_count", C.c_uint32),
                ("materials", C.c_void_p), ("material_count", C.c_uint32),
                ("instances", C.c_void_p), ("instance_count", C.c_uint32),
                ("emissive_triangles", C.c_void_p), ("emissive_triangle_count", C.c_uint32),
                ("textures", C.c_void_p), ("texture_count", C.c_uint32),
                ("texture_data", C.c_void_p), ("texture_data_bytes", C.c_uint64)]


class TextureDesc(C.Structure):
    _fields_ = [("channel", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32), ("texels", C.c_void_p)]


# MaterialTextures channels (tinytrace.h TT_TEXTURE_*), field order of asset::Material.
TEXTURE_CHANNELS = ("base_color", "specular_color", "transmission_color", "coat_color", "fuzz_color", "emission_color",
                    "geometry_normal", "base_weight", "base_metalness", "base_diffuse_roughness", "specular_weight",
                    "specular_roughness", "specular_roughness_anisotropy", "specular_ior", "specular_rotation",
                    "transmission_weight", "transmission_depth", "coat_weight", "coat_roughness",
                    "coat_roughness_anisotropy", "coat_ior", "coat_darkening", "coat_rotation", "fuzz_weight",
                    "fuzz_roughness", "emission_luminance", "geometry_opacity")
TT_TEXTURE_FORMAT_RGBA8_SRGB, TT_TEXTURE_FORMAT_RG8_UNORM, TT_TEXTURE_FORMAT_R8_UNORM, TT_TEXTURE_FORMAT_R16_FLOAT = 0, 1, 2, 3
TT_TEXTURE_FORMAT_BC1_SRGB, TT_TEXTURE_FORMAT_BC5_UNORM = 4, 5
TEXTURE_RECORD_DTYPE = np.dtype([("format", "<u4"), ("width", "<u4"), ("height", "<u4"), ("levels", "<u4"),
                                 ("offsets", "<u4", 16)])


class ContextDesc(C.Structure):
    _fields_ = [("device", C.c_int32), ("flags", C.c_uint32)]


class SceneInfo(C.Structure):
    _fields_ = [("triangle_count", C.c_uint32), ("unique_triangle_count", C.c_uint32),
                ("node_count", C.c_uint32), ("max_depth", C.c_uint32), ("sah_cost", C.c_double),
                ("expected_bytes_per_ray", C.c_double), ("build_ms", C.c_double), ("device_bytes", C.c_uint64)]


class RenderSettings(C.Structure):
    _fields_ = [("rng", C.c_uint32), ("seed", C.c_uint32), ("restir_candidates", C.c_uint32),
                ("restir_temporal", C.c_uint32), ("restir_spatial", C.c_uint32), ("flags", C.c_uint32)]


class RenderStats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("extend_launches", C.c_uint32), ("shadow_launches", C.c_uint32), ("extend_ms", C.c_double),
                ("shadow_ms", C.c_double), ("integrate_ms", C.c_double), ("restir_rays", C.c_uint64),
                ("restir_ms", C.c_double)]


class CameraDesc(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("yaw", C.c_float), ("pitch", C.c_float), ("z_near", C.c_float),
                ("z_far", C.c_float), ("fov", C.c_float)]


class ConfigDesc(C.Structure):
    _fields_ = [("bounce_count", C.c_uint32), ("sample_count", C.c_uint32), ("tonemap", C.c_uint32),
                ("light_sampling", C.c_uint32)]


class MeshDesc(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("texture_coordinates", C.c_void_p), ("indices", C.c_void_p),
                ("normals", C.c_void_p), ("tangents", C.c_void_p), ("vertex_count", C.c_uint32),
                ("index_count", C.c_uint32)]


MATERIAL_DESC_FIELDS = [
    ("base_weight", 1), ("base_color", 3), ("base_metalness", 1), ("base_diffuse_roughness", 1),
    ("specular_weight", 1), ("specular_color", 3), ("specular_roughness", 1), ("specular_roughness_anisotropy", 1),
    ("specular_ior", 1), ("specular_rotation", 1), ("transmission_weight", 1), ("transmission_color", 3),
    ("transmission_depth", 1), ("coat_weight", 1), ("coat_color", 3), ("coat_roughness", 1),
    ("coat_roughness_anisotropy", 1), ("coat_ior", 1), ("coat_darkening", 1), ("coat_rotation", 1),
    ("fuzz_weight", 1), ("fuzz_color", 3), ("fuzz_roughness", 1), ("emission_luminance", 1),
    ("emission_color", 3), ("geometry_opacity", 1), ("geometry_normal", 3),
]


class MaterialDesc(C.Structure):
    _fields_ = [(n, C.c_float if k == 1 else C.c_float * 3) for n, k in MATERIAL_DESC_FIELDS]


class SceneOptions(C.Structure):
    _fields_ = [("bvh_builder", C.c_uint32), ("reserved", C.c_uint32)]


TT_BVH_BUILDER_HOST_SAH, TT_BVH_BUILDER_DEVICE_LBVH = 0, 1


class ServerDesc(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("pool_paths", C.c_uint32), ("batches_in_flight", C.c_uint32)]


class ServerRequest(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("rng", C.c_uint32), ("seed", C.c_uint32), ("passes", C.c_uint32),
                ("constants", Constants), ("restir_candidates", C.c_uint32), ("restir_temporal", C.c_uint32),
                ("restir_spatial", C.c_uint32), ("reserved", C.c_uint32)]


class RequestStatus(C.Structure):
    _fields_ = [("state", C.c_uint32), ("passes_done", C.c_uint32), ("submitted_ns", C.c_uint64),
                ("admitted_ns", C.c_uint64), ("done_ns", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [("batches", C.c_uint64), ("paths", C.c_uint64), ("queued", C.c_uint64), ("active", C.c_uint64),
                ("done", C.c_uint64), ("budget_used_bytes", C.c_uint64), ("busy_ms", C.c_double),
                ("cache_queries", C.c_uint64), ("cache_samples", C.c_uint64), ("judge_samples", C.c_uint64),
                ("judge_traced", C.c_double), ("judge_cached", C.c_double)]


TT_REQUEST_QUEUED, TT_REQUEST_ADMITTED, TT_REQUEST_DONE = 0, 1, 2


class MlpDesc(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("learning_rate", C.c_float), ("reserved", C.c_uint32)]


class CacheSettings(C.Structure):
    _fields_ = [("train_samples", C.c_uint32), ("train_steps", C.c_uint32), ("judge_fraction", C.c_float),
                ("records_per_path", C.c_uint32), ("ordered_training", C.c_uint32)]


class CacheStats(C.Structure):
    _fields_ = [("queries", C.c_uint32), ("samples", C.c_uint32), ("loss", C.c_double),
                ("judge_samples", C.c_uint32), ("reserved", C.c_uint32), ("judge_traced", C.c_double),
                ("judge_cached", C.c_double), ("judge_loss", C.c_double), ("infer_ms", C.c_double),
                ("train_ms", C.c_double)]


MLP_PARAMS = 22720
MLP_IN = 96


# numpy dtypes of the reference byte layouts (little-endian, packed as C).
RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("tmin", "<f4"), ("direction", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("instance", "<u4"), ("primitive", "<u4"), ("barycentrics", "<f4", 2)])
BSDF_QUERY_DTYPE = np.dtype([("wi", "<f4", 3), ("wo", "<f4", 3)])
BSDF_RESULT_DTYPE = np.dtype([("bsdf", "<f4", 3), ("density", "<f4")])
VERTEX_DTYPE = np.dtype([("texture_coordinates", "<u2", 2), ("tangent_frame", "<u4")])
MESH_DTYPE = np.dtype([("center", "<f4", 3), ("radius", "<f4"), ("vertex_offset", "<u4"), ("vertex_count", "<u4"),
                       ("index_offset", "<u4"), ("index_count", "<u4")])
INSTANCE_DTYPE = np.dtype([("transform", "<f4", 16), ("inverse_transform", "<f4", 16),
                           ("normal_transform", "<f4", 16), ("vertex_offset", "<u4"), ("index_offset", "<u4"),
                           ("material_index", "<u4"), ("mesh_index", "<u4")])
EMISSIVE_DTYPE = np.dtype([("positions", "<i2", (3, 3)), ("hash", "<u2"), ("texture_coordinates", "<u2", (3, 2)),
                           ("instance_index", "<u4")])
MATERIAL_DTYPE = np.dtype([("constants", "<f4", 48), ("textures", "<u4", 27), ("padding", "<u4")])
assert RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 20 and VERTEX_DTYPE.itemsize == 8
assert MESH_DTYPE.itemsize == 32 and INSTANCE_DTYPE.itemsize == 208 and EMISSIVE_DTYPE.itemsize == 36
assert MATERIAL_DTYPE.itemsize == 304 and C.sizeof(Constants) == 368 and C.sizeof(MaterialConstants) == 192


class TinytraceError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


_lib = None

# (name, restype, argtypes)
_PROTOTYPES = [
    ("tt_context_create", C.c_int, [C.POINTER(ContextDesc), C.POINTER(C.c_void_p)]),
    ("tt_context_destroy", None, [C.c_void_p]),
    ("tt_last_error", C.c_char_p, [C.c_void_p]),
    ("tt_scene_create", C.c_int, [C.c_void_p, C.POINTER(SceneDesc), C.POINTER(C.c_void_p)]),
    ("tt_scene_create_ex", C.c_int, [C.c_void_p, C.POINTER(SceneDesc), C.POINTER(SceneOptions),
                                     C.POINTER(C.c_void_p)]),
    ("tt_scene_destroy", None, [C.c_void_p]),
    ("tt_scene_info_get", C.c_int, [C.c_void_p, C.POINTER(SceneInfo)]),
    ("tt_target_create", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)]),
    ("tt_target_destroy", None, [C.c_void_p]),
    ("tt_render", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Constants), C.c_void_p]),
    ("tt_render_ex", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Constants), C.POINTER(RenderSettings), C.c_void_p]),
    ("tt_target_download", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("tt_timing_ms", C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_double)]),
    ("tt_render_stats_get", C.c_int, [C.c_void_p, C.POINTER(RenderStats)]),
    ("tt_server_create", C.c_int, [C.c_void_p, C.POINTER(ServerDesc), C.POINTER(C.c_void_p)]),
    ("tt_server_destroy", None, [C.c_void_p]),
    ("tt_server_add_scene", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32)]),
    ("tt_server_submit", C.c_int, [C.c_void_p, C.POINTER(ServerRequest), C.POINTER(C.c_uint64)]),
    ("tt_server_step", C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    ("tt_server_status", C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(RequestStatus)]),
    ("tt_server_take_result", C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_size_t]),
    ("tt_server_stats_get", C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    ("tt_mlp_create", C.c_int, [C.c_void_p, C.POINTER(MlpDesc), C.POINTER(C.c_void_p)]),
    ("tt_mlp_destroy", None, [C.c_void_p]),
    ("tt_mlp_set_bounds", C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    ("tt_mlp_params", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]),
    ("tt_mlp_run", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                             C.POINTER(C.c_double)]),
    ("tt_render_cached", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Constants), C.POINTER(RenderSettings),
                                   C.c_void_p, C.POINTER(CacheSettings), C.c_void_p, C.POINTER(CacheStats)]),
    ("tt_server_set_cache", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.POINTER(CacheSettings)]),
    ("tt_trace_closest", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32]),
    ("tt_trace_closest_device", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32]),
    ("tt_trace_closest_static", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32]),
    ("tt_bsdf_evaluate", C.c_int, [C.c_void_p, C.POINTER(MaterialConstants), C.c_void_p, C.c_void_p, C.c_uint32,
                                   C.c_uint32]),
    ("tt_debug_math", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]),
    ("tt_camera_default", None, [C.POINTER(CameraDesc)]),
    ("tt_config_default", None, [C.POINTER(ConfigDesc)]),
    ("tt_constants_build", C.c_int, [C.POINTER(CameraDesc), C.POINTER(ConfigDesc), C.c_uint32, C.c_uint32,
                                      C.c_uint32, C.c_uint32, C.POINTER(Constants)]),
    ("tt_material_desc_default", None, [C.POINTER(MaterialDesc)]),
    ("tt_asset_scene_create", C.c_int, [C.POINTER(C.c_void_p)]),
    ("tt_asset_scene_destroy", None, [C.c_void_p]),
    ("tt_asset_set_block_compression", C.c_int, [C.c_void_p, C.c_uint32]),
    ("tt_asset_insert_mesh", C.c_int, [C.c_void_p, C.POINTER(MeshDesc), C.POINTER(C.c_uint32)]),
    ("tt_asset_insert_material", C.c_int, [C.c_void_p, C.POINTER(MaterialDesc), C.POINTER(C.c_uint32)]),
    ("tt_asset_add_node", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.POINTER(C.c_uint32)]),
    ("tt_asset_add_model", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32]),
    ("tt_asset_flatten", C.c_int, [C.c_void_p, C.POINTER(SceneDesc)]),
    ("tt_asset_import_gltf", C.c_int, [C.c_void_p, C.c_char_p]),
    ("tt_asset_insert_material_textured", C.c_int,
     [C.c_void_p, C.c_void_p, C.POINTER(TextureDesc), C.c_uint32, C.POINTER(C.c_uint32)]),
    ("tt_target_display", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Constants), C.c_void_p, C.c_size_t]),
]

EXPORTED_SYMBOLS = [p[0] for p in _PROTOTYPES]


# Live native handles by teardown order (servers, then scenes / targets /
# asset scenes, then contexts).  They are closed at interpreter exit, before
# the HIP runtime's own static teardown, whatever order the garbage collector
# would pick.
_LIVE = (weakref.WeakSet(), weakref.WeakSet(), weakref.WeakSet())


def track(obj, level: int) -> None:
    _LIVE[level].add(obj)


@atexit.register
def _close_live_handles() -> None:
    for live in _LIVE:
        for obj in list(live):
            try:
                obj.close()
            except Exception:
                pass


def lib():
    """The loaded native library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY_PATH):
            raise RuntimeError(f"{LIBRARY_PATH} is missing: build it with __graft_entry__.build() "
                               "or `make -C tinytrace_amd/csrc`")
        handle = C.CDLL(LIBRARY_PATH)
        for name, restype, argtypes in _PROTOTYPES:
            fn = getattr(handle, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _lib = handle
    return _lib


def check(status: int, context=None, what: str = "") -> None:
    if status != TT_OK:
        message = what
        if context:
            message = f"{what}: {lib().tt_last_error(context).decode()}"
        raise TinytraceError(status, message)


def ptr(array: np.ndarray) -> int:
    return array.ctypes.data

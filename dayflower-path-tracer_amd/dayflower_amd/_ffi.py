"""ctypes declarations for libdayflower_hip.so and libdayflower_scene.so.

The structures mirror include/dayflower_types.h field for field; the
function prototypes mirror include/dayflower_hip.h and
include/dayflower_scene.h.  The libraries are loaded from this package's
``lib/`` directory (built in-tree by ``build.py``); a missing library raises
instead of falling back to anything else.
"""

import ctypes
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
HIP_LIB_PATH = os.path.join(LIB_DIR, "libdayflower_hip.so")
SCENE_LIB_PATH = os.path.join(LIB_DIR, "libdayflower_scene.so")

c_float_p = ctypes.POINTER(ctypes.c_float)
c_int32_p = ctypes.POINTER(ctypes.c_int32)
c_int64_p = ctypes.POINTER(ctypes.c_int64)
c_uint8_p = ctypes.POINTER(ctypes.c_uint8)

DF_OK = 0
DF_EINVAL = -1
DF_ENODEVICE = -2
DF_EHIP = -3
DF_ESTATE = -4
DF_ENOMEM = -5
DF_EIO = -6

ERROR_NAMES = {
    DF_EINVAL: "DF_EINVAL",
    DF_ENODEVICE: "DF_ENODEVICE",
    DF_EHIP: "DF_EHIP",
    DF_ESTATE: "DF_ESTATE",
    DF_ENOMEM: "DF_ENOMEM",
    DF_EIO: "DF_EIO",
}

ARRAY_NAMES = (
    "point2fs", "point3fs", "vector3fs", "planes", "spheres", "terrains", "triangles", "bvh",
    "textures", "surfaces", "primitives", "object_to_world", "world_to_object", "camera",
)
INT_ARRAYS = {"planes", "triangles", "bvh", "primitives"}


class CompiledScene(ctypes.Structure):
    _fields_ = []
    for _n in ARRAY_NAMES:
        _fields_.append((_n, c_int32_p if _n in INT_ARRAYS else c_float_p))
        _fields_.append((_n + "_len", ctypes.c_size_t))
    del _n


class Sky(ctypes.Structure):
    _fields_ = [
        ("onb_u", ctypes.c_float * 3),
        ("onb_v", ctypes.c_float * 3),
        ("onb_w", ctypes.c_float * 3),
        ("sun_direction_world", ctypes.c_float * 3),
        ("sun_direction", ctypes.c_float * 3),
        ("sun_origin", ctypes.c_float * 3),
        ("theta", ctypes.c_float),
        ("turbidity", ctypes.c_float),
        ("zenith_relative_luminance", ctypes.c_double),
        ("zenith_x", ctypes.c_double),
        ("zenith_y", ctypes.c_double),
        ("perez_relative_luminance", ctypes.c_double * 5),
        ("perez_x", ctypes.c_double * 5),
        ("perez_y", ctypes.c_double * 5),
        ("is_sky_active", ctypes.c_int32),
        ("is_sun_active", ctypes.c_int32),
    ]


class Diff(ctypes.Structure):
    _fields_ = [
        ("n_albedo", ctypes.c_int32),
        ("n_emission", ctypes.c_int32),
        ("surface_albedo", ctypes.POINTER(ctypes.c_int32)),
        ("surface_emission", ctypes.POINTER(ctypes.c_int32)),
        ("n_surfaces", ctypes.c_size_t),
        ("albedo", ctypes.POINTER(ctypes.c_float)),
        ("emission", ctypes.POINTER(ctypes.c_float)),
        ("n_checker", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("surface_checker", ctypes.POINTER(ctypes.c_int32)),
        ("checker", ctypes.POINTER(ctypes.c_float)),
        ("turbidity", ctypes.c_double),
        ("elevation_deg", ctypes.c_double),
        ("azimuth_deg", ctypes.c_double),
    ]


class Schedule(ctypes.Structure):
    _fields_ = [("engine", ctypes.c_int32), ("wave_tile", ctypes.c_int32), ("shade_binned", ctypes.c_int32),
                ("bvh_pairs", ctypes.c_int32), ("tail_bounce", ctypes.c_int32)]


class Params(ctypes.Structure):
    _fields_ = [
        ("renderer_type", ctypes.c_int32),
        ("tone_mapper_type", ctypes.c_int32),
        ("tone_mapper_exposure", ctypes.c_float),
        ("pt_depth_maximum", ctypes.c_int32),
        ("pt_depth_russian_roulette", ctypes.c_int32),
        ("ao_maximum_distance", ctypes.c_float),
        ("normal_mapping", ctypes.c_int32),
        ("wireframes", ctypes.c_int32),
        ("shader_type", ctypes.c_int32),
        ("selected_primitive_offset", ctypes.c_int32),
        ("global_amplitude", ctypes.c_float),
        ("global_frequency", ctypes.c_float),
        ("global_gain", ctypes.c_float),
        ("global_lacunarity", ctypes.c_float),
        ("global_octaves", ctypes.c_int32),
        ("math_mode", ctypes.c_int32),
        ("pt_engine", ctypes.c_int32),
    ]


STAGE_NAMES = ("megakernel", "raygen", "trace", "shade", "escape", "film")

COUNTER_NAMES = ("rays", "samples", "box_tests", "tri_tests", "prim_tests", "sphere_tests", "terrain_steps",
                 "sky_evals")


class Counters(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in COUNTER_NAMES]


class SceneStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "primitives", "triangles", "spheres", "planes", "terrains", "bvh_nodes", "textures_len")]


_P = ctypes.c_void_p

HIP_PROTOTYPES = {
    "df_device_count": (ctypes.c_int, [c_int32_p]),
    "df_measure_roof": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double)]),
    "df_check_reciprocal": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(ctypes.c_uint64),
                                           ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint32)]),
    "df_check_terrain_march": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, c_float_p, ctypes.c_size_t, c_float_p,
                                              c_float_p, ctypes.POINTER(ctypes.c_uint32)]),
    "df_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int32]),
    "df_destroy": (None, [_P]),
    "df_validate_scene": (ctypes.c_int, [ctypes.POINTER(CompiledScene), ctypes.c_char_p, ctypes.c_size_t]),
    "df_load_scene": (ctypes.c_int, [_P, ctypes.POINTER(CompiledScene)]),
    "df_update": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "df_seed": (ctypes.c_int, [_P, ctypes.c_uint64]),
    "df_set_seeds": (ctypes.c_int, [_P, c_int64_p, ctypes.c_size_t]),
    "df_get_seeds": (ctypes.c_int, [_P, c_int64_p, ctypes.c_size_t]),
    "df_upload_camera": (ctypes.c_int, [_P, c_float_p, ctypes.c_size_t]),
    "df_upload_transforms": (ctypes.c_int, [_P, c_float_p, c_float_p, ctypes.c_int32]),
    "df_upload_surfaces": (ctypes.c_int, [_P, c_float_p, ctypes.c_size_t]),
    "df_upload_sky": (ctypes.c_int, [_P, ctypes.POINTER(Sky)]),
    "df_toggle_sky": (ctypes.c_int, [_P]),
    "df_toggle_sun": (ctypes.c_int, [_P]),
    "df_set_params": (ctypes.c_int, [_P, ctypes.POINTER(Params)]),
    "df_get_params": (ctypes.c_int, [_P, ctypes.POINTER(Params)]),
    "df_toggle_primitive_selection": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32]),
    "df_toggle_primitive_material": (ctypes.c_int, [_P]),
    "df_clear": (ctypes.c_int, [_P]),
    "df_execute": (ctypes.c_int, [_P, ctypes.c_int32]),
    "df_synchronize": (ctypes.c_int, [_P]),
    "df_get_engine": (ctypes.c_int, [_P, c_int32_p]),
    "df_last_execute_ms": (ctypes.c_int, [_P, c_float_p]),
    "df_get_image_bgra": (ctypes.c_int, [_P, c_uint8_p, ctypes.c_size_t]),
    "df_get_image_float": (ctypes.c_int, [_P, c_float_p, ctypes.c_size_t]),
    "df_get_primitive_offsets": (ctypes.c_int, [_P, c_int32_p, ctypes.c_size_t]),
    "df_get_film": (ctypes.c_int, [_P, c_float_p, ctypes.c_size_t, c_int32_p, ctypes.c_size_t]),
    "df_set_film": (ctypes.c_int, [_P, c_float_p, ctypes.c_size_t, c_int32_p, ctypes.c_size_t]),
    "df_get_counters": (ctypes.c_int, [_P, ctypes.POINTER(Counters)]),
    "df_reset_counters": (ctypes.c_int, [_P]),
    "df_set_stage_timing": (ctypes.c_int, [_P, ctypes.c_int32]),
    "df_set_unit_counters": (ctypes.c_int, [_P, ctypes.c_int32]),
    "df_get_stage_stats": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.POINTER(Counters),
                                          ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)]),
    "df_last_error": (ctypes.c_char_p, [_P]),
    "df_set_diff": (ctypes.c_int, [_P, ctypes.POINTER(Diff)]),
    "df_clear_diff": (ctypes.c_int, [_P]),
    "df_loss_gradient": (ctypes.c_int, [_P, c_float_p, ctypes.c_size_t, ctypes.c_int32,
                                        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                        ctypes.c_size_t, c_float_p, c_float_p]),
    "df_batch_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int32]),
    "df_batch_destroy": (None, [_P]),
    "df_batch_add": (ctypes.c_int, [_P, ctypes.POINTER(CompiledScene), ctypes.POINTER(Sky), ctypes.POINTER(Params),
                                    ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64, c_int32_p]),
    "df_batch_count": (ctypes.c_int, [_P, c_int32_p]),
    "df_batch_job": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.POINTER(_P)]),
    "df_batch_execute": (ctypes.c_int, [_P]),
    "df_batch_execute_noalloc": (ctypes.c_int, [_P]),
    "df_batch_set_graphs": (ctypes.c_int, [_P, ctypes.c_int32]),
    "df_set_graphs": (ctypes.c_int, [_P, ctypes.c_int32]),
    "df_get_schedule": (ctypes.c_int, [_P, ctypes.POINTER(Schedule)]),
    "df_set_schedule": (ctypes.c_int, [_P, ctypes.POINTER(Schedule)]),
    "df_get_bvh_pair_stack": (ctypes.c_int, [_P, c_int32_p]),
    "df_batch_last_execute_ms": (ctypes.c_int, [_P, c_float_p]),
    "df_batch_last_error": (ctypes.c_char_p, [_P]),
}

class Transform(ctypes.Structure):
    """dfs_transform: Transform(position, rotation quaternion (x, y, z, w), scale)."""
    _fields_ = [("position", ctypes.c_float * 3), ("rotation", ctypes.c_float * 4), ("scale", ctypes.c_float * 3)]


SurfaceMapper = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p)

SCENE_PROTOTYPES = {
    "dfs_build_config": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(_P)]),
    "dfs_build_batch_scene": (ctypes.c_int, [ctypes.c_uint64, ctypes.POINTER(_P)]),
    "dfs_read_scene": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_P)]),
    "dfs_write_scene": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "dfs_scene_from_arrays": (ctypes.c_int, [ctypes.POINTER(CompiledScene), ctypes.c_char_p, ctypes.POINTER(_P)]),
    "dfs_free": (None, [_P]),
    "dfs_get_compiled": (ctypes.c_int, [_P, ctypes.POINTER(CompiledScene)]),
    "dfs_get_sky": (ctypes.c_int, [_P, ctypes.POINTER(Sky)]),
    "dfs_get_render_config": (ctypes.c_int, [_P, c_int32_p, c_int32_p, c_int32_p, ctypes.POINTER(Params)]),
    "dfs_get_stats": (ctypes.c_int, [_P, ctypes.POINTER(SceneStats)]),
    "dfs_get_name": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_size_t]),
    "dfs_sky_set": (ctypes.c_int, [ctypes.POINTER(Sky), ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    "dfs_camera_setup": (ctypes.c_int, [c_float_p] + [ctypes.c_float] * 11),
    "dfs_default_params": (None, [ctypes.POINTER(Params)]),
    "dfs_camera_default": (ctypes.c_int, [c_float_p]),
    "dfs_camera_set_eye": (ctypes.c_int, [c_float_p, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    "dfs_camera_set_resolution": (ctypes.c_int, [c_float_p, ctypes.c_float, ctypes.c_float]),
    "dfs_builder_create": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_P)]),
    "dfs_builder_free": (None, [_P]),
    "dfs_builder_last_error": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_size_t]),
    "dfs_builder_constant_texture": (ctypes.c_int, [_P, c_float_p, c_int32_p]),
    "dfs_builder_checkerboard_texture": (ctypes.c_int, [_P, c_float_p, c_float_p, ctypes.c_float, ctypes.c_float,
                                                        ctypes.c_float, c_int32_p]),
    "dfs_builder_fbm_texture": (ctypes.c_int, [_P, c_float_p, c_float_p, ctypes.c_float, ctypes.c_float,
                                               ctypes.c_int32, c_int32_p]),
    "dfs_builder_blend_texture": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_float, c_int32_p]),
    "dfs_builder_bullseye_texture": (ctypes.c_int, [_P, c_float_p, c_float_p, c_int32_p]),
    "dfs_builder_uv_texture": (ctypes.c_int, [_P, c_int32_p]),
    "dfs_builder_surface_normal_texture": (ctypes.c_int, [_P, ctypes.c_int32, c_int32_p]),
    "dfs_builder_image_texture": (ctypes.c_int, [_P, c_int32_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                                 ctypes.c_float, ctypes.c_float, c_int32_p]),
    "dfs_builder_surface": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.c_float, ctypes.c_float, c_int32_p]),
    "dfs_builder_plane": (ctypes.c_int, [_P, c_float_p, c_float_p, c_float_p, ctypes.c_int32,
                                         ctypes.POINTER(Transform)]),
    "dfs_builder_sphere": (ctypes.c_int, [_P, c_float_p, ctypes.c_float, ctypes.c_int32, ctypes.POINTER(Transform)]),
    "dfs_builder_terrain": (ctypes.c_int, [_P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                           ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(Transform)]),
    "dfs_builder_triangle": (ctypes.c_int, [_P, c_float_p, ctypes.c_int32, ctypes.POINTER(Transform)]),
    "dfs_builder_mesh": (ctypes.c_int, [_P, c_float_p, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(Transform)]),
    "dfs_builder_load_obj": (ctypes.c_int, [_P, ctypes.c_char_p, SurfaceMapper, ctypes.c_void_p, ctypes.c_int32,
                                            ctypes.POINTER(Transform), c_int32_p]),
    "dfs_builder_set_camera": (ctypes.c_int, [_P, c_float_p]),
    "dfs_builder_set_sky": (ctypes.c_int, [_P, ctypes.POINTER(Sky)]),
    "dfs_builder_compile": (ctypes.c_int, [_P, ctypes.POINTER(_P)]),
}


def _load(path, prototypes):
    if not os.path.exists(path):
        raise RuntimeError(
            "%s is not built; run dayflower-path-tracer_amd/build.py (or __graft_entry__.build())" % path)
    lib = ctypes.CDLL(path)
    for name, (res, args) in prototypes.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_hip = None
_scene = None


def hip_lib():
    """The renderer library (raises if it has not been built)."""
    global _hip
    if _hip is None:
        _hip = _load(HIP_LIB_PATH, HIP_PROTOTYPES)
    return _hip


def scene_lib():
    """The host scene-compiler library (raises if it has not been built)."""
    global _scene
    if _scene is None:
        _scene = _load(SCENE_LIB_PATH, SCENE_PROTOTYPES)
    return _scene

"""ctypes mirror of include/cortex_rt.h and the loader of the in-tree libcortex_rt.so.

The library is the product: there is no Python or CPU fallback. If the shared object is missing,
importing the renderer raises immediately.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcortex_rt.so")

CRT_ABI_VERSION = 3

CRT_OK = 0
CRT_ERR_NO_DEVICE = 1
CRT_ERR_OUT_OF_MEMORY = 2
CRT_ERR_OVER_BUDGET = 3
CRT_ERR_INVALID_ARGUMENT = 4
CRT_ERR_INVALID_SCENE = 5
CRT_ERR_IO = 6
CRT_ERR_CANCELLED = 7
CRT_ERR_NOT_READY = 8
CRT_ERR_HIP = 9
CRT_ERR_INTERNAL = 10

CRT_MAT_LAMBERTIAN = 1
CRT_MAT_METAL = 2
CRT_MAT_DIELECTRIC = 3
CRT_MAT_ISOTROPIC = 4
CRT_MAT_OREN_NAYAR = 5
CRT_MAT_DIFFUSE_LIGHT = 8
CRT_MAT_SPOT_LIGHT = 9
CRT_LIGHT_OMNI = 0
CRT_LIGHT_SPOT = 1
CRT_MATF_ONE_SIDED = 1

CRT_TEX_SOLID = 0
CRT_TEX_CHECKER = 1
CRT_TEX_IMAGE = 2

CRT_RNG_PCG = 0
CRT_RNG_MINSTD = 1

CRT_RF_TIME_STAGES = 1
CRT_RF_MEGAKERNEL = 2
CRT_RF_IMAGE_ONLY = 4

CRT_BVH_HOST = 1
CRT_BVH_DEVICE = 2
CRT_SERVE_SMALL_THEN_OLDEST = 1
CRT_SERVE_ROUND_ROBIN = 2

CRT_XF_TRANSLATE = 1
CRT_XF_ROTATE = 2
CRT_XF_SCALE = 4

CRT_TRACE_CAMERA = 1
CRT_TRACE_STAGE = 2

CRT_EXTEND_DEFAULT = 0
CRT_EXTEND_GRID_STRIDE = 1
CRT_EXTEND_PERSISTENT = 2
CRT_EXTEND_WAVE_CHUNK = 3

f3 = C.c_float * 3


class crt_camera(C.Structure):
    _fields_ = [("lookfrom", f3), ("lookat", f3), ("vup", f3), ("vfov_deg", C.c_float), ("aspect_ratio", C.c_float),
                ("aperture", C.c_float), ("focus_dist", C.c_float), ("time0", C.c_float), ("time1", C.c_float),
                ("orthographic", C.c_uint32), ("ortho_height", C.c_float)]


class crt_material(C.Structure):
    _fields_ = [("type", C.c_uint32), ("texture", C.c_int32), ("albedo", f3), ("fuzz", C.c_float), ("ior", C.c_float),
                ("intensity", C.c_float), ("flags", C.c_uint32)]


class crt_texture(C.Structure):
    _fields_ = [("type", C.c_uint32), ("color", f3), ("scale", C.c_float), ("even", C.c_int32), ("odd", C.c_int32),
                ("width", C.c_uint32), ("height", C.c_uint32), ("channels", C.c_uint32),
                ("pixels", C.POINTER(C.c_uint8))]


class crt_uvmap(C.Structure):
    _fields_ = [("scale_u", C.c_float), ("scale_v", C.c_float), ("offset_u", C.c_float), ("offset_v", C.c_float),
                ("repeat_u", C.c_float), ("repeat_v", C.c_float)]


class crt_sphere(C.Structure):
    _fields_ = [("center", f3), ("radius", C.c_float), ("material", C.c_uint32), ("mapping", crt_uvmap)]


class crt_quad(C.Structure):
    _fields_ = [("corner", f3), ("u", f3), ("v", f3), ("material", C.c_uint32)]


class crt_area_light(C.Structure):
    _fields_ = [("center", f3), ("u", f3), ("v", f3), ("rgb", f3), ("intensity", C.c_float), ("invisible", C.c_uint32)]


class crt_transform(C.Structure):
    _fields_ = [("has", C.c_uint32), ("translate", f3), ("rotate_deg", f3), ("scale", f3)]


CRT_SHAPE_DISK, CRT_SHAPE_CYLINDER, CRT_SHAPE_CONE = 0, 1, 2


class crt_shape(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("center", f3), ("radius", C.c_float), ("height", C.c_float),
                ("material", C.c_uint32), ("mapping", crt_uvmap), ("transform", crt_transform)]


class crt_mesh(C.Structure):
    _fields_ = [("positions", C.POINTER(C.c_float)), ("num_vertices", C.c_uint32), ("indices", C.POINTER(C.c_uint32)),
                ("num_triangles", C.c_uint32), ("normals", C.POINTER(C.c_float)), ("uvs", C.POINTER(C.c_float)),
                ("material", C.c_uint32), ("tri_materials", C.POINTER(C.c_uint32))]


class crt_instance(C.Structure):
    _fields_ = [("mesh", C.c_uint32), ("object_to_world", C.c_float * 12), ("material_override", C.c_int32)]


class crt_sphere_light(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("center", f3), ("radius", C.c_float), ("rgb", f3), ("intensity", C.c_float),
                ("invisible", C.c_uint32), ("direction", f3), ("cutoff_deg", C.c_float), ("falloff", C.c_float)]




class crt_scene_desc(C.Structure):
    _fields_ = [("abi_version", C.c_uint32), ("camera", crt_camera), ("background", f3),
                ("materials", C.POINTER(crt_material)), ("num_materials", C.c_uint32),
                ("textures", C.POINTER(crt_texture)), ("num_textures", C.c_uint32),
                ("lights", C.POINTER(crt_area_light)), ("num_lights", C.c_uint32),
                ("spheres", C.POINTER(crt_sphere)), ("num_spheres", C.c_uint32),
                ("quads", C.POINTER(crt_quad)), ("num_quads", C.c_uint32),
                ("meshes", C.POINTER(crt_mesh)), ("num_meshes", C.c_uint32),
                ("instances", C.POINTER(crt_instance)), ("num_instances", C.c_uint32),
                ("sphere_lights", C.POINTER(crt_sphere_light)), ("num_sphere_lights", C.c_uint32),
                ("shapes", C.POINTER(crt_shape)), ("num_shapes", C.c_uint32)]


class crt_render_desc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32), ("max_depth", C.c_uint32),
                ("seed", C.c_uint64), ("rng_mode", C.c_uint32), ("sample_begin", C.c_uint32),
                ("sample_count", C.c_uint32), ("flags", C.c_uint32)]


class crt_image(C.Structure):
    _fields_ = [("rgb8", C.POINTER(C.c_uint8)), ("accum", C.POINTER(C.c_float))]


class crt_stats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("rays", C.c_uint64), ("retraces", C.c_uint64), ("light_pdf_tests", C.c_uint64),
                ("node_visits", C.c_uint64), ("prim_tests", C.c_uint64), ("gpu_ms", C.c_double),
                ("extend_ms", C.c_double), ("host_ms", C.c_double)]


class crt_config(C.Structure):
    _fields_ = [("device", C.c_int32), ("hbm_budget_bytes", C.c_uint64), ("max_paths", C.c_uint32),
                ("block_size", C.c_uint32), ("count_work", C.c_uint32), ("leaf_size", C.c_uint32),
                ("bvh_width", C.c_uint32), ("extend_kernel", C.c_uint32), ("refill_min", C.c_uint32),
                ("pass_paths", C.c_uint32), ("grid_per_cu", C.c_uint32), ("shade_grid_per_cu", C.c_uint32),
                ("sub_passes", C.c_uint32), ("graphs", C.c_uint32), ("stack_lds", C.c_uint32),
                ("ray_sort", C.c_uint32), ("bvh_builder", C.c_uint32), ("batch_group_paths", C.c_uint32),
                ("serve_order", C.c_uint32), ("serve_small_paths", C.c_uint32)]


class crt_scene_info(C.Structure):
    _fields_ = [("num_nodes", C.c_uint64), ("num_leaves", C.c_uint64), ("num_prims", C.c_uint64),
                ("device_bytes", C.c_uint64), ("max_depth", C.c_uint32), ("num_instances", C.c_uint32),
                ("bvh_width", C.c_uint32), ("node_bytes", C.c_uint32), ("leaf_size", C.c_uint32),
                ("builder", C.c_uint32), ("build_ms", C.c_double), ("upload_ms", C.c_double), ("sah_cost", C.c_double)]


class crt_batch_item(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("rd", crt_render_desc), ("out", crt_image)]


class crt_request_timing(C.Structure):
    _fields_ = [("submit_s", C.c_double), ("admit_s", C.c_double), ("first_pass_s", C.c_double), ("done_s", C.c_double),
                ("rays", C.c_uint64), ("passes", C.c_uint32), ("status", C.c_int32)]


class crt_work_counters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("rays", "retraces", "light_pdf_tests", "node_visits", "leaf_visits", "tri_tests",
                                          "sphere_tests", "quad_tests", "inst_visits", "extend_launches", "passes",
                                          "iterations", "exact_queries")]


class crt_memory_info(C.Structure):
    _fields_ = [("budget", C.c_uint64), ("reserved", C.c_uint64), ("scenes", C.c_uint64), ("requests", C.c_uint64)]


class crt_obj_mesh(C.Structure):
    _fields_ = [("positions", C.POINTER(C.c_float)), ("num_vertices", C.c_uint32), ("indices", C.POINTER(C.c_uint32)),
                ("num_triangles", C.c_uint32), ("normals", C.POINTER(C.c_float)), ("uvs", C.POINTER(C.c_float))]


class crt_scenefile_info(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32), ("max_depth", C.c_uint32),
                ("aspect_ratio", C.c_float), ("unsupported", C.c_uint32), ("skipped", C.c_uint32),
                ("output_path", C.c_char * 512)]


class crt_tune_result(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("max_paths", "grid_per_cu", "shade_grid_per_cu", "sub_passes", "extend_kernel",
                                          "refill_min", "stack_lds", "bvh_width", "leaf_size", "trials")] + \
               [("best_mrays_s", C.c_double), ("initial_mrays_s", C.c_double)]


STRUCTS_IN_ABI_ORDER = [crt_camera, crt_material, crt_texture, crt_uvmap, crt_sphere, crt_quad, crt_area_light,
                        crt_mesh, crt_instance, crt_scene_desc, crt_render_desc, crt_image, crt_stats, crt_config,
                        crt_scene_info, crt_batch_item, crt_request_timing, crt_work_counters,
                        crt_memory_info, crt_obj_mesh, crt_scenefile_info, crt_tune_result, crt_sphere_light,
                        crt_transform]

# Every symbol declared in include/cortex_rt.h.
EXPORTED = ["crt_create", "crt_destroy", "crt_last_error", "crt_device_count", "crt_scene_upload", "crt_scene_release",
            "crt_scene_get_info", "crt_scene_verify", "crt_scene_build_check", "crt_render", "crt_trace_rays", "crt_trace_primary", "crt_render_batch",
            "crt_submit", "crt_poll", "crt_wait", "crt_cancel", "crt_request_get_timing", "crt_last_counters",
            "crt_memory_get_info", "crt_obj_load", "crt_obj_free", "crt_scenefile_load",
            "crt_scenefile_desc", "crt_scenefile_close", "crt_struct_sizes", "crt_autotune", "crt_debug_fail_alloc", "crt_debug_valu_probe",
            "crt_render_noalloc", "crt_builder_create", "crt_builder_destroy", "crt_builder_camera",
            "crt_builder_background", "crt_builder_solid_texture", "crt_builder_checker_texture",
            "crt_builder_image_texture", "crt_builder_image_texture_pixels", "crt_builder_lambertian", "crt_builder_metal",
            "crt_builder_dielectric", "crt_builder_diffuse_light", "crt_builder_has_material",
            "crt_builder_directional_light", "crt_builder_omni_light", "crt_builder_spot_light", "crt_builder_sphere",
            "crt_builder_quad", "crt_builder_box", "crt_builder_shape", "crt_builder_isotropic", "crt_builder_oren_nayar", "crt_builder_obj_mesh", "crt_builder_note_unsupported",
            "crt_builder_desc"]

_P = C.POINTER
_SIGS = {
    "crt_create": ([_P(crt_config), _P(C.c_void_p)], C.c_int),
    "crt_destroy": ([C.c_void_p], C.c_int),
    "crt_last_error": ([], C.c_char_p),
    "crt_device_count": ([_P(C.c_int32)], C.c_int),
    "crt_scene_upload": ([C.c_void_p, _P(crt_scene_desc), _P(C.c_uint32)], C.c_int),
    "crt_scene_release": ([C.c_void_p, C.c_uint32], C.c_int),
    "crt_scene_get_info": ([C.c_void_p, C.c_uint32, _P(crt_scene_info)], C.c_int),
    "crt_scene_verify": ([C.c_void_p, C.c_uint32], C.c_int),
    "crt_scene_build_check": ([_P(crt_scene_desc), C.c_uint32, C.c_uint32, _P(crt_scene_info)], C.c_int),
    "crt_render": ([C.c_void_p, C.c_uint32, _P(crt_render_desc), _P(crt_image), _P(crt_stats)], C.c_int),
    "crt_trace_rays": ([C.c_void_p, C.c_uint32, _P(C.c_float), C.c_uint64, C.c_int32, _P(C.c_int32), _P(C.c_int32),
                        _P(C.c_float), _P(C.c_float)], C.c_int),
    "crt_trace_primary": ([C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, _P(C.c_int32), _P(C.c_float)], C.c_int),
    "crt_render_batch": ([C.c_void_p, _P(crt_batch_item), C.c_uint32, _P(crt_stats)], C.c_int),
    "crt_submit": ([C.c_void_p, C.c_uint32, _P(crt_render_desc), _P(C.c_uint64)], C.c_int),
    "crt_poll": ([C.c_void_p, C.c_uint64, _P(C.c_int32)], C.c_int),
    "crt_wait": ([C.c_void_p, C.c_uint64, _P(crt_image), _P(crt_stats)], C.c_int),
    "crt_cancel": ([C.c_void_p, C.c_uint64], C.c_int),
    "crt_request_get_timing": ([C.c_void_p, C.c_uint64, _P(crt_request_timing)], C.c_int),
    "crt_last_counters": ([C.c_void_p, _P(crt_work_counters)], C.c_int),
    "crt_memory_get_info": ([C.c_void_p, _P(crt_memory_info)], C.c_int),
    "crt_obj_load": ([C.c_char_p, _P(crt_obj_mesh)], C.c_int),
    "crt_obj_free": ([_P(crt_obj_mesh)], C.c_int),
    "crt_scenefile_load": ([C.c_char_p, _P(C.c_void_p)], C.c_int),
    "crt_scenefile_desc": ([C.c_void_p, _P(crt_scene_desc), _P(crt_scenefile_info)], C.c_int),
    "crt_scenefile_close": ([C.c_void_p], C.c_int),
    "crt_builder_create": ([_P(C.c_void_p)], C.c_int),
    "crt_builder_destroy": ([C.c_void_p], C.c_int),
    "crt_builder_camera": ([C.c_void_p, _P(crt_camera)], C.c_int),
    "crt_builder_background": ([C.c_void_p, _P(C.c_float)], C.c_int),
    "crt_builder_solid_texture": ([C.c_void_p, C.c_char_p, _P(C.c_float)], C.c_int),
    "crt_builder_checker_texture": ([C.c_void_p, C.c_char_p, C.c_float, C.c_char_p, C.c_char_p, _P(C.c_float),
                                     _P(C.c_float)], C.c_int),
    "crt_builder_image_texture": ([C.c_void_p, C.c_char_p, C.c_char_p], C.c_int),
    "crt_builder_image_texture_pixels": ([C.c_void_p, C.c_char_p, C.c_uint32, C.c_uint32, C.c_uint32, _P(C.c_uint8)],
                                         C.c_int),
    "crt_builder_lambertian": ([C.c_void_p, C.c_char_p, _P(C.c_float), C.c_char_p], C.c_int),
    "crt_builder_metal": ([C.c_void_p, C.c_char_p, _P(C.c_float), C.c_float, C.c_char_p], C.c_int),
    "crt_builder_dielectric": ([C.c_void_p, C.c_char_p, C.c_float], C.c_int),
    "crt_builder_isotropic": ([C.c_void_p, C.c_char_p, _P(C.c_float), C.c_char_p], C.c_int),
    "crt_builder_oren_nayar": ([C.c_void_p, C.c_char_p, _P(C.c_float), C.c_char_p, C.c_float, C.c_float], C.c_int),
    "crt_builder_diffuse_light": ([C.c_void_p, C.c_char_p, _P(C.c_float), C.c_float, C.c_int32], C.c_int),
    "crt_builder_has_material": ([C.c_void_p, C.c_char_p], C.c_int),
    "crt_builder_directional_light": ([C.c_void_p, C.c_char_p, _P(C.c_float), _P(C.c_float), _P(C.c_float), C.c_float,
                                       _P(C.c_float), C.c_int32], C.c_int),
    "crt_builder_omni_light": ([C.c_void_p, C.c_char_p, _P(C.c_float), C.c_float, C.c_float, _P(C.c_float), C.c_int32],
                               C.c_int),
    "crt_builder_spot_light": ([C.c_void_p, C.c_char_p, _P(C.c_float), _P(C.c_float), C.c_float, C.c_float, C.c_float,
                                C.c_float, _P(C.c_float), C.c_int32], C.c_int),
    "crt_builder_sphere": ([C.c_void_p, C.c_char_p, _P(C.c_float), C.c_float, C.c_char_p, _P(crt_uvmap),
                            _P(crt_transform)], C.c_int),
    "crt_builder_quad": ([C.c_void_p, C.c_char_p, _P(C.c_float), _P(C.c_float), _P(C.c_float), C.c_char_p,
                          _P(crt_transform)], C.c_int),
    "crt_builder_box": ([C.c_void_p, C.c_char_p, _P(C.c_float), _P(C.c_float), C.c_char_p, _P(crt_transform)], C.c_int),
    "crt_builder_shape": ([C.c_void_p, C.c_uint32, C.c_char_p, _P(C.c_float), C.c_float, C.c_float, C.c_char_p,
                           _P(crt_uvmap), _P(crt_transform)], C.c_int),
    "crt_builder_obj_mesh": ([C.c_void_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int32], C.c_int),
    "crt_builder_note_unsupported": ([C.c_void_p, C.c_uint32], C.c_int),
    "crt_builder_desc": ([C.c_void_p, _P(crt_scene_desc), _P(C.c_uint32), _P(C.c_uint32)], C.c_int),
    "crt_struct_sizes": ([_P(C.c_uint64), C.c_uint32], C.c_int),
    "crt_autotune": ([C.c_void_p, C.c_uint32, _P(crt_scene_desc), _P(crt_render_desc), _P(crt_tune_result)], C.c_int),
    "crt_debug_fail_alloc": ([C.c_void_p, C.c_int32], C.c_int),
    "crt_debug_valu_probe": ([C.c_void_p, C.c_int32, C.c_uint32, C.c_uint32, _P(C.c_double), _P(C.c_uint64)], C.c_int),
    "crt_render_noalloc": ([C.c_void_p, C.c_uint32, _P(crt_render_desc), C.c_void_p, C.c_void_p, _P(crt_stats)], C.c_int),
}

_lib = None


class CrtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("crt error %d: %s" % (code, msg))
        self.code = code


def load():
    """Loads libcortex_rt.so (built by __graft_entry__.build()); raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libcortex_rt.so is not built at %s; run __graft_entry__.build()" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def check(rc):
    if rc != CRT_OK:
        msg = load().crt_last_error()
        raise CrtError(rc, msg.decode() if msg else "")
    return rc

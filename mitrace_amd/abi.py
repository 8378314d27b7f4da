"""ctypes mirror of ``include/mitrace_gpu.h`` (the C-ABI of the GPU Tracer).

Every structure here has the same field order and types as its C counterpart; ``tests``
check the sizes against the compiled library.
"""
import ctypes as C

MT_OK = 0
(MT_DEBUG_NONE, MT_DEBUG_PRIMARY_BVH_TESTS, MT_DEBUG_TOTAL_BVH_TESTS,
 MT_DEBUG_PRIMARY_TRIANGLE_TESTS, MT_DEBUG_TOTAL_TRIANGLE_TESTS, MT_DEBUG_ALBEDO,
 MT_DEBUG_METALLIC_ROUGHNESS, MT_DEBUG_GEOMETRIC_NORMAL, MT_DEBUG_SHADING_NORMAL,
 MT_DEBUG_DIRECT_ONLY, MT_DEBUG_INDIRECT_ONLY, MT_DEBUG_EMISSION, MT_DEBUG_COLOR_PER_MESH,
 MT_DEBUG_COLOR_PER_TRIANGLE, MT_DEBUG_COLOR_PER_BOUNDING_VOLUME, MT_DEBUG_BOUNCES,
 MT_DEBUG_DEPTH, MT_DEBUG_FIRST_HIT_FRESNEL, MT_DEBUG_FIRST_HIT_BRDF, MT_DEBUG_FIRST_HIT_PDF,
 MT_DEBUG_REFLECTED_DIRECTION, MT_DEBUG_PIXEL_STANDARD_DEVIATION,
 MT_DEBUG_FIREFLY_ELIMINATION) = range(23)
MT_E_INVALID = -1
MT_E_HIP = -2
MT_E_NOMEM = -3
MT_E_UNSUPPORTED = -4
MT_E_KILLED = -5

MT_LIGHT_POINT, MT_LIGHT_DIRECTIONAL, MT_LIGHT_SPOT, MT_LIGHT_AREA = 0, 1, 2, 3
MT_CAMERA_PERSPECTIVE, MT_CAMERA_ORTHOGONAL = 0, 1
MT_FILTER_NEAREST, MT_FILTER_LINEAR = 0, 1
MT_ENV_NONE, MT_ENV_RGB8, MT_ENV_RGB32F = 0, 1, 2
MT_MODE_PARITY, MT_MODE_FIXED = 0, 1
MT_OUTPUT_TONEMAPPED, MT_OUTPUT_LINEAR = 0, 1
MT_TRAVERSAL_REFERENCE, MT_TRAVERSAL_FAST = 0, 1
MT_MATH_SIN, MT_MATH_COS, MT_MATH_ATAN2, MT_MATH_ACOS, MT_MATH_POW5 = 0, 1, 2, 3, 4


class mt_mesh_desc(C.Structure):
    _fields_ = [("tri_count", C.c_uint32),
                ("positions", C.POINTER(C.c_float)),
                ("normals", C.POINTER(C.c_float)),
                ("texcoord0", C.POINTER(C.c_float))]


class mt_image_desc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgb", C.POINTER(C.c_uint8))]


class mt_texture_desc(C.Structure):
    _fields_ = [("image", C.c_int32), ("filter", C.c_int32),
                ("offset", C.c_float * 2), ("scale", C.c_float * 2)]


class mt_material_desc(C.Structure):
    _fields_ = [("base_color_factor", C.c_float * 4),
                ("emissive_factor", C.c_float * 3),
                ("normal_scale", C.c_float),
                ("metallic_factor", C.c_float),
                ("roughness_factor", C.c_float),
                ("occlusion_strength", C.c_float),
                ("base_color_tex", mt_texture_desc),
                ("metallic_roughness_tex", mt_texture_desc),
                ("normal_tex", mt_texture_desc),
                ("occlusion_tex", mt_texture_desc),
                ("emissive_tex", mt_texture_desc)]


class mt_instance_desc(C.Structure):
    _fields_ = [("mesh", C.c_uint32), ("material", C.c_uint32), ("transform", C.c_float * 16)]


class mt_light_desc(C.Structure):
    _fields_ = [("type", C.c_int32),
                ("position", C.c_float * 3),
                ("color", C.c_float * 3),
                ("point_size", C.c_float),
                ("spot_inner_cos", C.c_float),
                ("spot_outer_cos", C.c_float),
                ("direction", C.c_float * 3)]


class mt_camera_desc(C.Structure):
    _fields_ = [("type", C.c_int32),
                ("camera_to_world", C.c_float * 16),
                ("yfov", C.c_float),
                ("ortho_size", C.c_float * 2),
                ("aspect", C.c_float)]


class mt_env_desc(C.Structure):
    _fields_ = [("format", C.c_int32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("texels", C.c_void_p), ("filter", C.c_int32)]


class mt_bvh_options(C.Structure):
    _fields_ = [("disable", C.c_int32), ("max_triangles", C.c_uint32), ("max_depth", C.c_uint32),
                ("fast", C.c_int32), ("device_build_min_tris", C.c_uint32),
                ("fast_device_build_min_tris", C.c_uint32)]


class mt_scene_desc(C.Structure):
    _fields_ = [("meshes", C.POINTER(mt_mesh_desc)), ("mesh_count", C.c_uint32),
                ("images", C.POINTER(mt_image_desc)), ("image_count", C.c_uint32),
                ("materials", C.POINTER(mt_material_desc)), ("material_count", C.c_uint32),
                ("instances", C.POINTER(mt_instance_desc)), ("instance_count", C.c_uint32),
                ("lights", C.POINTER(mt_light_desc)), ("light_count", C.c_uint32),
                ("env", mt_env_desc),
                ("bvh", mt_bvh_options)]


class mt_settings(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32),
                ("samples", C.c_uint32), ("bounces", C.c_uint32),
                ("ev_exposure", C.c_float), ("hdri_rotation", C.c_float),
                ("hdri_secondary_intensity", C.c_float), ("hdri_primary_intensity", C.c_float),
                ("emission_base_intensity", C.c_float),
                ("firefly_enable", C.c_int32), ("firefly_threshold", C.c_float),
                ("terminate_energy", C.c_float),
                ("image_block_size", C.c_uint32),
                ("seed", C.c_uint64),
                ("mode", C.c_int32), ("rr_depth", C.c_uint32), ("output", C.c_int32),
                ("hdri_rotation_apply", C.c_int32),
                ("paths_in_flight", C.c_uint32),
                ("count_work", C.c_int32), ("debug_mode", C.c_int32),
                ("traversal", C.c_int32)]


class mt_stats(C.Structure):
    _fields_ = [("progress", C.c_float), ("time_elapsed", C.c_float),
                ("estimated_time_remaining", C.c_float),
                ("rays", C.c_uint64), ("samples", C.c_uint64),
                ("inst_aabb_tests", C.c_uint64), ("inst_entered", C.c_uint64),
                ("node_pops", C.c_uint64), ("tri_tests", C.c_uint64),
                ("hits_shaded", C.c_uint64), ("misses", C.c_uint64),
                ("kernel_ms_extend", C.c_float), ("kernel_ms_shadow", C.c_float),
                ("kernel_ms_shade", C.c_float), ("kernel_ms_total", C.c_float),
                ("extend_launches", C.c_uint32), ("shadow_launches", C.c_uint32),
                ("diag", C.c_uint64 * 9)]


class mt_hit(C.Structure):
    _fields_ = [("instance", C.c_int32), ("triangle", C.c_uint32), ("bv_index", C.c_uint32),
                ("u", C.c_float), ("v", C.c_float), ("distance", C.c_float),
                ("world_pos", C.c_float * 3),
                ("bvh_tests", C.c_uint32), ("triangle_tests", C.c_uint32)]


class mt_diff_params(C.Structure):
    """device pointers (integers from torch's data_ptr()); 0 keeps the scene's value"""
    _fields_ = [("base_texture_dev", C.c_void_p), ("tex_width", C.c_uint32),
                ("tex_height", C.c_uint32), ("tex_material_first", C.c_uint32),
                ("tex_material_count", C.c_uint32), ("roughness_dev", C.c_void_p),
                ("metallic_dev", C.c_void_p), ("emission_dev", C.c_void_p),
                ("material_count", C.c_uint32)]


class mt_diff_grads(C.Structure):
    _fields_ = [("base_texture_dev", C.c_void_p), ("roughness_dev", C.c_void_p),
                ("metallic_dev", C.c_void_p), ("emission_dev", C.c_void_p),
                ("detach_sampling", C.c_int32)]


class mt_cache_options(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("train_batch", C.c_uint32), ("learning_rate", C.c_float),
                ("train_every", C.c_uint32), ("query_depth", C.c_uint32),
                ("warmup_passes", C.c_uint32), ("steps_per_pass", C.c_uint32),
                ("spp_per_pass", C.c_uint32), ("query_after_light", C.c_uint32),
                ("train_nee", C.c_uint32), ("loss", C.c_uint32)]


class mt_batch_options(C.Structure):
    _fields_ = [("paths_per_iteration", C.c_uint32), ("tile_pixels", C.c_uint32),
                ("iteration_tiles", C.c_uint32), ("trace_blocks_per_cu", C.c_uint32),
                ("streams", C.c_uint32), ("mode", C.c_uint32)]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class mt_batch_job_info(C.Structure):
    _fields_ = [("done", C.c_int32), ("status", C.c_int32), ("tiles", C.c_uint32),
                ("pad", C.c_uint32), ("samples", C.c_uint64), ("submit_ms", C.c_double),
                ("complete_ms", C.c_double)]


class mt_batch_stats(C.Structure):
    _fields_ = [("iterations", C.c_uint64), ("graphs_captured", C.c_uint64),
                ("samples", C.c_uint64), ("jobs", C.c_uint32), ("scenes", C.c_uint32),
                ("options", mt_batch_options), ("pad", C.c_uint32), ("launches", C.c_uint64)]


class mt_autotune_trial(C.Structure):
    _fields_ = [("options", mt_batch_options), ("seconds", C.c_double),
                ("msamples_per_s", C.c_double)]


# every exported symbol of the library and its signature: (restype, argtypes)
SIGNATURES = {
    "mt_abi_version": (C.c_int, []),
    "mt_last_error": (C.c_char_p, []),
    "mt_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "mt_pool_trim": (C.c_int, []),
    "mt_scene_create": (C.c_int, [C.POINTER(mt_scene_desc), C.POINTER(C.c_void_p)]),
    "mt_scene_destroy": (C.c_int, [C.c_void_p]),
    "mt_scene_mesh_bvh": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_uint32)]),
    "mt_render": (C.c_int, [C.c_void_p, C.POINTER(mt_camera_desc), C.POINTER(mt_settings),
                            C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(mt_stats)]),
    "mt_render_start": (C.c_int, [C.c_void_p, C.POINTER(mt_camera_desc), C.POINTER(mt_settings),
                                  C.POINTER(C.c_float), C.POINTER(C.c_void_p)]),
    "mt_render_wait": (C.c_int, [C.c_void_p]),
    "mt_render_kill": (C.c_int, [C.c_void_p]),
    "mt_render_stats": (C.c_int, [C.c_void_p, C.POINTER(mt_stats)]),
    "mt_render_is_done": (C.c_int, [C.c_void_p]),
    "mt_job_destroy": (C.c_int, [C.c_void_p]),
    "mt_trace_closest": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_float),
                                   C.POINTER(mt_hit)]),
    "mt_trace_closest_ex": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_float),
                                      C.POINTER(mt_hit), C.c_int32]),
    "mt_trace_occluded": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_float),
                                    C.POINTER(C.c_uint32), C.c_int32]),
    "mt_brdf_eval": (C.c_int, [C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "mt_trace_closest_noalloc": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                           C.c_int32, C.c_uint64]),
    "mt_math_eval": (C.c_int, [C.c_int32, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "mt_fast_bvh_build": (C.c_int, [C.POINTER(C.c_float), C.c_uint32, C.POINTER(C.c_float),
                                    C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_uint32), C.POINTER(C.c_float)]),
    "mt_diff_create": (C.c_int, [C.c_void_p, C.POINTER(mt_settings), C.POINTER(C.c_void_p)]),
    "mt_diff_forward": (C.c_int, [C.c_void_p, C.POINTER(mt_camera_desc),
                                  C.POINTER(mt_diff_params), C.POINTER(mt_diff_params),
                                  C.c_uint64, C.c_void_p, C.c_uint64]),
    "mt_diff_backward": (C.c_int, [C.c_void_p, C.POINTER(mt_camera_desc),
                                   C.POINTER(mt_diff_params), C.c_uint64, C.c_void_p,
                                   C.POINTER(mt_diff_grads), C.c_uint64]),
    "mt_diff_destroy": (C.c_int, [C.c_void_p]),
    "mt_cache_create": (C.c_int, [C.c_void_p, C.POINTER(mt_cache_options), C.POINTER(C.c_void_p)]),
    "mt_cache_destroy": (C.c_int, [C.c_void_p]),
    "mt_render_cached": (C.c_int, [C.c_void_p, C.POINTER(mt_camera_desc), C.POINTER(mt_settings),
                                   C.c_void_p, C.POINTER(C.c_float), C.POINTER(mt_stats)]),
    "mt_cache_get_weights": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "mt_cache_set_weights": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "mt_cache_infer": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_float),
                                 C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "mt_cache_train": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_float),
                                 C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                 C.POINTER(C.c_float)]),
    "mt_batch_create": (C.c_int, [C.POINTER(mt_batch_options), C.POINTER(C.c_void_p)]),
    "mt_batch_submit": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(mt_camera_desc),
                                  C.POINTER(mt_settings), C.POINTER(C.c_float),
                                  C.POINTER(C.c_uint32)]),
    "mt_batch_start": (C.c_int, [C.c_void_p]),
    "mt_batch_finish": (C.c_int, [C.c_void_p]),
    "mt_batch_get_job_info": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(mt_batch_job_info)]),
    "mt_batch_get_stats": (C.c_int, [C.c_void_p, C.POINTER(mt_batch_stats)]),
    "mt_batch_destroy": (C.c_int, [C.c_void_p]),
    "mt_batch_autotune": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(mt_camera_desc),
                                    C.POINTER(mt_settings), C.c_uint32,
                                    C.POINTER(mt_batch_options), C.POINTER(mt_autotune_trial),
                                    C.c_uint32, C.POINTER(C.c_uint32)]),
}


def fptr(a):
    """float32 numpy array -> POINTER(c_float) (the array must stay alive)."""
    return a.ctypes.data_as(C.POINTER(C.c_float))

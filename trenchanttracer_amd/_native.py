"""ctypes binding of libtt.so (the C ABI declared in include/tt.h).

The library is the product: there is no Python or CPU fallback for any device
entry point. If libtt.so is missing, importing this module raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtt.so")
ABI_VERSION = 5   # tt_abi_version() of the library this binding's structs describe

TT_OK = 0
STATUS_NAMES = {0: "TT_OK", 1: "TT_ERR_INVALID", 2: "TT_ERR_NOMEM", 3: "TT_ERR_DEVICE",
                4: "TT_ERR_IO", 5: "TT_ERR_UNSUPPORTED"}


class TTError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class tt_ray(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("ox", "oy", "oz", "tmin", "dx", "dy", "dz", "tmax")]


class tt_hit(C.Structure):
    _fields_ = [("tri_id", C.c_int32), ("t", C.c_float), ("u", C.c_float), ("v", C.c_float)]


class tt_bvh_stats(C.Structure):
    _fields_ = [("num_tris", C.c_int32), ("num_refs", C.c_int32), ("num_duplicates", C.c_int32),
                ("num_inner", C.c_int32), ("num_leaves", C.c_int32), ("max_depth", C.c_int32),
                ("sah", C.c_float), ("root_lo", C.c_float * 3), ("root_hi", C.c_float * 3),
                ("build_seconds", C.c_double)]


class tt_compact_sizes(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("nodes", "vertices", "normals", "tri_indices", "mat_indices",
                                          "leaf_count", "tri_count")]


class tt_material(C.Structure):
    _fields_ = [("type", C.c_int32), ("albedo", C.c_float * 3), ("emission", C.c_float * 3),
                ("roughness", C.c_float), ("ior", C.c_float), ("specular", C.c_float * 3)]


class tt_geometry(C.Structure):
    _fields_ = [("type", C.c_int32), ("material", C.c_int32), ("a", C.c_float * 3), ("b", C.c_float * 3),
                ("c", C.c_float * 3), ("pad", C.c_float)]


class tt_trace_cfg(C.Structure):
    _fields_ = [("variant", C.c_int32), ("lds_stack", C.c_int32), ("block", C.c_int32),
                ("xcd_order", C.c_int32), ("refill", C.c_int32), ("reorder", C.c_int32)]


class tt_camera(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("view", C.c_float * 3), ("up", C.c_float * 3),
                ("fov", C.c_float * 2), ("aperture", C.c_float), ("focal", C.c_float),
                ("width", C.c_int32), ("height", C.c_int32)]


class tt_render_params(C.Structure):
    _fields_ = [("spp", C.c_int32), ("max_depth", C.c_int32), ("rr_depth", C.c_int32),
                ("seed", C.c_uint32), ("mode", C.c_int32), ("guided", C.c_int32),
                ("sample_offset", C.c_int32), ("accumulate_fp64", C.c_int32),
                ("pi", C.c_float), ("two_pi", C.c_float), ("tmin", C.c_float), ("use_graph", C.c_int32),
                ("paths_in_flight", C.c_int32), ("lds_stack", C.c_int32), ("block", C.c_int32),
                ("variant", C.c_int32), ("refill_camera", C.c_int32), ("refill", C.c_int32), ("tile", C.c_int32)]


class tt_render_stats(C.Structure):
    _fields_ = [("rays_extension", C.c_int64), ("rays_shadow", C.c_int64), ("samples", C.c_int64),
                ("seconds", C.c_double), ("iterations", C.c_int32), ("paths_in_flight", C.c_int32)]


class tt_config_info(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("samples", C.c_int32), ("bvh_reload", C.c_int32),
                ("material_reload", C.c_int32), ("num_materials", C.c_int32), ("num_geometries", C.c_int32),
                ("num_tris", C.c_int32), ("position", C.c_float * 3), ("yaw", C.c_float), ("pitch", C.c_float),
                ("radius", C.c_float), ("focal", C.c_float), ("aperture", C.c_float), ("fovx", C.c_float),
                ("work_dir", C.c_char * 256), ("hdr_file", C.c_char * 256), ("material_file", C.c_char * 256)]


class tt_guide_params(C.Structure):
    _fields_ = [("alpha", C.c_float), ("min_roughness", C.c_float), ("train", C.c_int32), ("train_paths", C.c_int32),
                ("train_batch", C.c_int32), ("train_spp", C.c_int32), ("max_bounce", C.c_int32),
                ("pass_growth", C.c_int32), ("learn_alpha", C.c_int32)]


class tt_guide_stats(C.Structure):
    _fields_ = [("train_steps", C.c_int64), ("train_rows", C.c_int64), ("train_seconds", C.c_double),
                ("last_loss", C.c_float), ("first_loss", C.c_float)]


P = C.c_void_p
I32, I64, U32, U64, F32, F64 = C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_float, C.c_double
PF = C.POINTER(C.c_float)
PI32 = C.POINTER(C.c_int32)
PI64 = C.POINTER(C.c_int64)
PD = C.POINTER(C.c_double)

# name -> (restype, argtypes)
PROTOS = {
    "tt_last_error": (C.c_char_p, []),
    "tt_abi_version": (C.c_int, []),
    "tt_mesh_create": (C.c_int, [P, I32, P, I32, P, I32, P, P, C.POINTER(P)]),
    "tt_mesh_destroy": (None, [P]),
    "tt_mesh_info": (C.c_int, [P, PI32, PI32, PI32]),
    "tt_mesh_get": (C.c_int, [P, P, P, P, P, P]),
    "tt_mesh_append": (C.c_int, [P, P, I32]),
    "tt_mesh_gen_soup": (C.c_int, [I32, U64, F32, F32, C.POINTER(P)]),
    "tt_mesh_gen_icosphere": (C.c_int, [I32, F32, PF, F32, I32, U64, I32, C.POINTER(P)]),
    "tt_mesh_gen_quad": (C.c_int, [PF, PF, PF, I32, C.POINTER(P)]),
    "tt_mesh_assign_materials": (C.c_int, [P, I32, PF, PI32, U64]),
    "tt_rays_primary": (C.c_int, [I32, I32, I32, PF, PF, F32, F32, F32, U64, P]),
    "tt_rays_incoherent": (C.c_int, [P, I64, F32, F32, U64, P]),
    "tt_bvh_build": (C.c_int, [P, F32, I32, C.POINTER(P)]),
    "tt_bvh_destroy": (None, [P]),
    "tt_bvh_get_stats": (C.c_int, [P, C.POINTER(tt_bvh_stats)]),
    "tt_bvh_export_compact": (C.c_int, [P, C.POINTER(tt_compact_sizes), P, P, P, P, P]),
    "tt_bvh_export_device_layout": (C.c_int, [P, PI64, PI64, P, P]),
    "tt_bvh_save": (C.c_int, [P, C.c_char_p]),
    "tt_bvh_load": (C.c_int, [C.c_char_p, C.POINTER(P)]),
    "tt_ctx_create": (C.c_int, [I32, C.POINTER(P)]),
    "tt_ctx_destroy": (None, [P]),
    "tt_ctx_sync": (C.c_int, [P]),
    "tt_scene_create": (C.c_int, [P, P, P, P, I32, P, I32, I32, F32, C.POINTER(P)]),
    "tt_scene_destroy": (None, [P]),
    "tt_scene_info": (C.c_int, [P, PI64, PI64, PI64]),
    "tt_env_export_cdf": (C.c_int, [P, P, P, PD]),
    "tt_dev_alloc": (C.c_int, [P, C.c_size_t, C.POINTER(P)]),
    "tt_dev_free": (C.c_int, [P, P]),
    "tt_memcpy_h2d": (C.c_int, [P, P, P, C.c_size_t]),
    "tt_memcpy_d2h": (C.c_int, [P, P, P, C.c_size_t]),
    "tt_memset_d": (C.c_int, [P, P, C.c_int, C.c_size_t]),
    "tt_bench_gather64": (C.c_int, [P, P, C.c_size_t, I64, U32, C.POINTER(C.c_float)]),
    "tt_dev_copy": (C.c_int, [P, P, P, C.c_size_t]),
    "tt_display_frame": (C.c_int, [P, P, P, I64, I32, P, P]),
    "tt_timer_create": (C.c_int, [P, I32, C.POINTER(P)]),
    "tt_timer_record": (C.c_int, [P, I32]),
    "tt_timer_elapsed": (C.c_int, [P, I32, I32, C.POINTER(C.c_float)]),
    "tt_timer_destroy": (None, [P]),
    "tt_ctx_stream": (P, [P]),
    "tt_camera_from_orbit": (C.c_int, [PF, F32, F32, F32, F32, I32, I32, F32, F32, C.POINTER(tt_camera)]),
    "tt_render": (C.c_int, [P, C.POINTER(tt_camera), C.POINTER(tt_render_params), P, P, C.POINTER(tt_render_stats)]),
    "tt_debug_math": (C.c_int, [P, I32, P, P, P, I64]),
    "tt_guide_create": (C.c_int, [P, U32, F32, C.POINTER(P)]),
    "tt_guide_destroy": (None, [P]),
    "tt_guide_param_count": (I64, []),
    "tt_guide_get_params": (C.c_int, [P, P]),
    "tt_guide_set_params": (C.c_int, [P, P]),
    "tt_guide_forward": (C.c_int, [P, P, I64, P]),
    "tt_guide_grad": (C.c_int, [P, P, P, P, I64, PF, P]),
    "tt_debug_guide_grad_unfused": (C.c_int, [P, P, P, P, I64, PF, P]),
    "tt_guide_train_step": (C.c_int, [P, P, P, P, I64, PF]),
    "tt_debug_guide_sample": (C.c_int, [P, P, I64, U32, P]),
    "tt_debug_guide_forward_unfused": (C.c_int, [P, P, I64, P]),
    "tt_debug_guide_infer_slots": (C.c_int, [P, P, P, I64, P, PI64]),
    "tt_scene_set_geometries": (C.c_int, [P, P, I32]),
    "tt_hdr_read": (C.c_int, [C.c_char_p, PI32, PI32, P, I64]),
    "tt_hdr_write": (C.c_int, [C.c_char_p, I32, I32, P]),
    "tt_obj_load": (C.c_int, [C.c_char_p, C.POINTER(P), P, I32, PI32]),
    "tt_config_load": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(P)]),
    "tt_config_destroy": (None, [P]),
    "tt_config_get": (C.c_int, [P, C.POINTER(tt_config_info), P, P, C.POINTER(P)]),
    "tt_guide_render": (C.c_int, [P, P, C.POINTER(tt_camera), C.POINTER(tt_render_params), C.POINTER(tt_guide_params),
                                  P, P, C.POINTER(tt_render_stats), C.POINTER(tt_guide_stats)]),
    "tt_batch_create": (C.c_int, [P, I32, P, P, P, P, P, I32, I32, F32, C.POINTER(P)]),
    "tt_batch_destroy": (None, [P]),
    "tt_batch_render": (C.c_int, [P, P, C.POINTER(tt_render_params), P, C.POINTER(tt_render_stats)]),
    "tt_debug_bsdf_sample": (C.c_int, [P, C.POINTER(tt_material), PF, I32, I64, U32, P, P]),
    "tt_debug_env_sample": (C.c_int, [P, I64, U32, P]),
    "tt_debug_env_invert": (C.c_int, [P, P, I64, P]),
    "tt_debug_ray_order": (C.c_int, [P, P, I64, I64, P]),
    "tt_trace_closest": (C.c_int, [P, P, I64, P, C.POINTER(tt_trace_cfg), P]),
    "tt_trace_closest_noalloc": (C.c_int, [P, P, I64, P, C.POINTER(tt_trace_cfg), P]),
    "tt_bench_fetch_replay": (C.c_int, [P, P, I64, C.POINTER(tt_trace_cfg), C.c_int32, C.POINTER(C.c_float),
                                        C.POINTER(I64), C.POINTER(I64), P]),
    "tt_trace_any": (C.c_int, [P, P, I64, P, C.POINTER(tt_trace_cfg), P]),
    "tt_trace_closest_host": (C.c_int, [P, P, I64, P, C.POINTER(tt_trace_cfg)]),
    "tt_trace_any_host": (C.c_int, [P, P, I64, P, C.POINTER(tt_trace_cfg)]),
}

OPTIONAL = set()


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
                          " (the product has no fallback path)")
    lib = C.CDLL(LIB_PATH)
    lib.tt_abi_version.restype = C.c_int
    if lib.tt_abi_version() != ABI_VERSION:
        raise ImportError(f"{LIB_PATH} has ABI version {lib.tt_abi_version()}, this binding expects {ABI_VERSION}:"
                          " rebuild it")
    for name, (res, args) in PROTOS.items():
        try:
            fn = getattr(lib, name)
        except AttributeError:
            if name in OPTIONAL:
                continue
            raise
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def last_error():
    return lib.tt_last_error().decode(errors="replace")


def check(status):
    if status != TT_OK:
        raise TTError(status, last_error())
    return status

"""ctypes binding of libviennals_gpu.so (the C ABI in include/viennals_gpu.h).

The library is built in-tree by viennals_amd/build.py.  Loading fails loudly
when it is missing: there is no CPU fallback for the advection path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libviennals_gpu.so")
_lib = None


class LsgError(RuntimeError):
    pass


# every symbol declared in include/viennals_gpu.h
EXPORTED = [
    "lsg_last_error", "lsg_domain_create", "lsg_domain_destroy", "lsg_domain_deep_copy",
    "lsg_domain_insert_points", "lsg_domain_set_runs", "lsg_domain_num_runs", "lsg_domain_get_runs",
    "lsg_domain_num_points", "lsg_domain_width", "lsg_domain_set_width", "lsg_make_sphere", "lsg_make_plane", "lsg_make_box",
    "lsg_make_sphere_ignore_bc", "lsg_make_box_ignore_bc",
    "lsg_expand", "lsg_expand_noalloc", "lsg_reduce", "lsg_prune", "lsg_boolean", "lsg_rebuild", "lsg_advect_create",
    "lsg_advect_destroy", "lsg_advect_insert_next_level_set", "lsg_advect_clear_level_sets",
    "lsg_advect_set_velocity_builtin", "lsg_advect_set_velocity_builtin_material",
    "lsg_advect_set_dissipation_alpha_builtin", "lsg_advect_set_spatial_scheme",
    "lsg_advect_set_temporal_scheme",
    "lsg_advect_set_advection_time", "lsg_advect_set_time_step_ratio", "lsg_advect_set_single_step",
    "lsg_advect_set_dissipation_alpha", "lsg_advect_set_check_dissipation",
    "lsg_advect_set_calculate_normal_vectors", "lsg_advect_set_ignore_voids",
    "lsg_advect_set_adaptive_time_stepping", "lsg_advect_set_update_point_data", "lsg_advect_prepare_ls",
    "lsg_advect_apply", "lsg_advect_get_advected_time", "lsg_advect_get_current_time_step",
    "lsg_advect_get_number_of_time_steps", "lsg_advect_get_point_updates", "lsg_advect_set_profiling",
    "lsg_advect_get_rate_kernel_stats", "lsg_advect_get_star_kernel_stats", "lsg_advect_get_expand2_kernel_stats",
    "lsg_weno_eval", "lsg_device_synchronize",
    "lsg_advect_set_learned_weights", "lsg_weno5_learned_eval", "lsg_mlp_create", "lsg_mlp_destroy",
    "lsg_mlp_set_samples", "lsg_mlp_loss_grad", "lsg_mlp_train", "lsg_mlp_get_params",
    "lsg_batch_create", "lsg_batch_destroy", "lsg_batch_set_velocity_builtin", "lsg_batch_set_schemes",
    "lsg_batch_apply", "lsg_batch_get_advected_times", "lsg_batch_get_point_updates", "lsg_batch_write_back",
    "lsg_batch_set_scan_blocks", "lsg_batch_autotune", "lsg_advect_set_velocity_host",
    "lsg_advect_set_dissipation_alpha_host",
    "lsg_advect_set_velocity_update_callback",
    "lsg_advect_set_save_advection_velocities", "lsg_domain_point_data_insert", "lsg_domain_point_data_count",
    "lsg_domain_point_data_label", "lsg_domain_point_data_size", "lsg_domain_point_data_get",
    "lsg_domain_point_data_clear", "lsg_domain_serialize", "lsg_domain_deserialize", "lsg_domain_get_grid",
    "lsg_mark_void_points", "lsg_surface_points", "lsg_compare_chamfer",
    "lsg_compare_volume", "lsg_set_single_pass_band_ops",
]


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LsgError(f"{LIB_PATH} is missing: build it with `python -m viennals_amd.build` "
                       "(the advection path has no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    PP = ctypes.POINTER(ctypes.c_void_p)
    i32p = ctypes.POINTER(ctypes.c_int32)
    i64p = ctypes.POINTER(ctypes.c_int64)
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    fp = ctypes.POINTER(ctypes.c_float)
    c_int, c_double = ctypes.c_int, ctypes.c_double
    sig = {
        "lsg_last_error": ([], ctypes.c_char_p),
        "lsg_domain_create": ([c_int, i32p, i32p, ip, c_double, PP], c_int),
        "lsg_domain_destroy": ([P], c_int),
        "lsg_domain_deep_copy": ([P, P], c_int),
        "lsg_domain_insert_points": ([P, ctypes.c_int64, i32p, dp, c_int], c_int),
        "lsg_domain_set_runs": ([P, ctypes.c_int64, i32p, dp, c_int], c_int),
        "lsg_domain_num_runs": ([P, i64p], c_int),
        "lsg_domain_get_runs": ([P, i32p, dp, ctypes.c_int64], c_int),
        "lsg_domain_num_points": ([P, i64p], c_int),
        "lsg_domain_width": ([P, ip], c_int),
        "lsg_domain_set_width": ([P, c_int], c_int),
        "lsg_make_sphere": ([P, dp, c_double], c_int),
        "lsg_make_sphere_ignore_bc": ([P, dp, c_double, c_int], c_int),
        "lsg_make_plane": ([P, dp, dp], c_int),
        "lsg_make_box": ([P, dp, dp], c_int),
        "lsg_make_box_ignore_bc": ([P, dp, dp, c_int], c_int),
        "lsg_expand": ([P, c_int], c_int),
        "lsg_expand_noalloc": ([P, c_int], c_int),
        "lsg_reduce": ([P, c_int], c_int),
        "lsg_prune": ([P, c_int], c_int),
        "lsg_boolean": ([P, P, c_int], c_int),
        "lsg_rebuild": ([P], c_int),
        "lsg_advect_create": ([PP], c_int),
        "lsg_advect_destroy": ([P], c_int),
        "lsg_advect_insert_next_level_set": ([P, P], c_int),
        "lsg_advect_clear_level_sets": ([P], c_int),
        "lsg_advect_set_velocity_builtin": ([P, c_int, c_int, c_int, c_double, c_double, c_double, dp], c_int),
        "lsg_advect_set_velocity_builtin_material": ([P, c_int, c_int, c_int, c_int, c_double, c_double, c_double, dp],
                                                     c_int),
        "lsg_advect_set_dissipation_alpha_builtin": ([P, c_int, dp], c_int),
        "lsg_advect_set_spatial_scheme": ([P, c_int], c_int),
        "lsg_advect_set_temporal_scheme": ([P, c_int], c_int),
        "lsg_advect_set_advection_time": ([P, c_double], c_int),
        "lsg_advect_set_time_step_ratio": ([P, c_double], c_int),
        "lsg_advect_set_single_step": ([P, c_int], c_int),
        "lsg_advect_set_dissipation_alpha": ([P, c_double], c_int),
        "lsg_advect_set_check_dissipation": ([P, c_int], c_int),
        "lsg_advect_set_calculate_normal_vectors": ([P, c_int], c_int),
        "lsg_advect_set_ignore_voids": ([P, c_int], c_int),
        "lsg_advect_set_adaptive_time_stepping": ([P, c_int, ctypes.c_uint], c_int),
        "lsg_advect_set_update_point_data": ([P, c_int], c_int),
        "lsg_advect_set_save_advection_velocities": ([P, c_int], c_int),
        "lsg_domain_point_data_insert": ([P, ctypes.c_char_p, c_int, ctypes.c_void_p, ctypes.c_int64, c_int], c_int),
        "lsg_domain_point_data_count": ([P, c_int], c_int),
        "lsg_domain_point_data_label": ([P, c_int, c_int, ctypes.c_char_p, c_int], c_int),
        "lsg_domain_point_data_size": ([P, c_int, c_int], ctypes.c_int64),
        "lsg_domain_point_data_get": ([P, c_int, c_int, ctypes.c_void_p, ctypes.c_int64], c_int),
        "lsg_domain_point_data_clear": ([P], c_int),
        "lsg_domain_serialize": ([P, ctypes.c_void_p, ctypes.c_int64, i64p], c_int),
        "lsg_domain_deserialize": ([P, ctypes.c_void_p, ctypes.c_int64], c_int),
        "lsg_domain_get_grid": ([P, ip, i32p, i32p, ip, dp], c_int),
        "lsg_mark_void_points": ([P], c_int),
        "lsg_surface_points": ([P, ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)], c_int),
        "lsg_compare_chamfer": ([P, P, ctypes.c_void_p, ctypes.c_void_p], c_int),
        "lsg_compare_volume": ([P, P, ctypes.c_void_p, ctypes.c_void_p], c_int),
        "lsg_set_single_pass_band_ops": ([c_int], c_int),
        "lsg_advect_prepare_ls": ([P], c_int),
        "lsg_advect_apply": ([P], c_int),
        "lsg_advect_get_advected_time": ([P, dp], c_int),
        "lsg_advect_get_current_time_step": ([P, dp], c_int),
        "lsg_advect_get_number_of_time_steps": ([P, ctypes.POINTER(ctypes.c_uint)], c_int),
        "lsg_advect_get_point_updates": ([P, dp], c_int),
        "lsg_advect_set_profiling": ([P, c_int], c_int),
        "lsg_advect_get_rate_kernel_stats": ([P, dp, i64p, dp], c_int),
        "lsg_advect_get_star_kernel_stats": ([P, dp, i64p, dp], c_int),
        "lsg_advect_get_expand2_kernel_stats": ([P, dp, i64p, dp], c_int),
        "lsg_weno_eval": ([c_int, ctypes.c_int64, dp, c_double, c_int, dp], c_int),
        "lsg_advect_set_learned_weights": ([P, fp, c_int], c_int),
        "lsg_weno5_learned_eval": ([ctypes.c_int64, dp, c_double, c_int, fp, dp], c_int),
        "lsg_mlp_create": ([ctypes.POINTER(P), fp], c_int),
        "lsg_mlp_destroy": ([P], c_int),
        "lsg_mlp_set_samples": ([P, ctypes.c_int64, fp, fp], c_int),
        "lsg_mlp_loss_grad": ([P, ctypes.c_int64, ctypes.c_int64, dp, fp], c_int),
        "lsg_mlp_train": ([P, c_int, ctypes.c_int64, c_double, dp], c_int),
        "lsg_mlp_get_params": ([P, fp], c_int),
        "lsg_batch_create": ([ctypes.POINTER(P), ctypes.POINTER(P), c_int], c_int),
        "lsg_batch_destroy": ([P], c_int),
        "lsg_batch_set_velocity_builtin": ([P, c_int, c_int, c_int, c_int, c_double, c_double, c_double, dp], c_int),
        "lsg_batch_set_schemes": ([P, c_int, c_int, c_double], c_int),
        "lsg_batch_apply": ([P], c_int),
        "lsg_batch_get_advected_times": ([P, dp], c_int),
        "lsg_batch_get_point_updates": ([P, dp, ctypes.POINTER(ctypes.c_uint)], c_int),
        "lsg_batch_write_back": ([P], c_int),
        "lsg_batch_set_scan_blocks": ([P, c_int], c_int),
        "lsg_batch_autotune": ([P, ip, dp], c_int),
        "lsg_advect_set_velocity_host": ([P, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p], c_int),
        "lsg_advect_set_dissipation_alpha_host": ([P, ctypes.c_void_p, ctypes.c_void_p], c_int),
        "lsg_advect_set_velocity_update_callback": ([P, ctypes.c_void_p, ctypes.c_void_p], c_int),
        "lsg_device_synchronize": ([], c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return _lib


def check(code):
    """Raise LsgError with lsg_last_error() when a C-ABI call failed."""
    if code != 0:
        msg = lib().lsg_last_error()
        raise LsgError(f"viennals_gpu error {code}: {msg.decode() if msg else ''}")
    return code


# Host velocity callbacks (lsg_scalar_velocity_fn / lsg_vector_velocity_fn).
SCALAR_VELOCITY_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_double), ctypes.c_uint64, ctypes.c_void_p)
VECTOR_VELOCITY_FN = ctypes.CFUNCTYPE(None, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_double), ctypes.c_uint64, ctypes.c_void_p,
                                      ctypes.POINTER(ctypes.c_double))
VELOCITY_UPDATE_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p)
DISSIPATION_ALPHA_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                        ctypes.c_void_p)

"""ctypes binding of libles_hip.so (include/les_hip.h).  The library is the only compute path:
importing works without a GPU (for the host-only generators and field metadata), while
creating a handle without a HIP device fails loudly with LES_E_NODEV."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libles_hip.so")

c_int, c_float, c_double, c_void_p = ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_void_p
c_int64, c_uint64 = ctypes.c_int64, ctypes.c_uint64
P_float = ctypes.POINTER(ctypes.c_float)

LES_OK, LES_E_INVALID, LES_E_HIP, LES_E_NOMEM, LES_E_STATE, LES_E_NODEV = range(6)
LES_PREC_FP32, LES_PREC_MIXED = 0, 1
LES_SUM_EXACT, LES_SUM_REFERENCE = 0, 1
ERR_NAMES = {1: "LES_E_INVALID", 2: "LES_E_HIP", 3: "LES_E_NOMEM", 4: "LES_E_STATE",
             5: "LES_E_NODEV"}


class LesConfig(ctypes.Structure):
    _fields_ = [("ip", c_int), ("jp", c_int), ("kp", c_int), ("dt", c_float), ("ro", c_float),
                ("vn", c_float), ("alpha", c_float), ("beta", c_float), ("cs0", c_float),
                ("omega", c_float), ("pjuge", c_float), ("nmaxp", c_int), ("precision", c_int),
                ("refine_tol", c_double), ("refine_max", c_int), ("device", c_int),
                ("sub_i", c_int), ("sub_j", c_int), ("refine_omega", c_float),
                ("refine_solver", c_int), ("sum_order", c_int)]


class LesScalars(ctypes.Structure):
    _fields_ = [("uout", c_float), ("rhsav", c_float), ("area", c_float), ("pav", c_float),
                ("pco", c_float), ("sor", c_float), ("sor_iters", c_int),
                ("resid_l2", c_double), ("refine_passes", c_int), ("cg_iters", c_int)]


class LesTuning(ctypes.Structure):
    _fields_ = [("sor_tile_j", c_int), ("sor_tile_k", c_int), ("sor_iter_ms", c_float),
                ("stencil_bx", c_int * 3), ("stencil_by", c_int * 3), ("stencil_ms", c_float * 3)]


# every symbol declared in include/les_hip.h: name -> (restype, argtypes)
H = c_void_p
PROTOS = {
    "les_field_bounds": (c_int64, [c_int, c_int, c_int, c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int)]),
    "les_default_config": (c_int, [ctypes.POINTER(LesConfig), c_int, c_int, c_int]),
    "les_create": (c_int, [ctypes.POINTER(LesConfig), ctypes.POINTER(c_void_p)]),
    "les_destroy": (c_int, [H]),
    "les_last_error": (ctypes.c_char_p, [H]),
    "les_sync": (c_int, [H]),
    "les_get_dims": (c_int, [H, ctypes.POINTER(c_int), ctypes.POINTER(c_int), ctypes.POINTER(c_int)]),
    "les_write_restart": (c_int, [H, ctypes.c_char_p, ctypes.c_char_p, c_int, c_float, c_int]),
    "les_read_restart": (c_int, [H, ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(c_int),
                                 ctypes.POINTER(c_float)]),
    "les_netcdf_create": (c_int, [ctypes.c_char_p, c_int, c_int, c_int, c_int,
                                  ctypes.POINTER(c_void_p)]),
    "les_netcdf_write": (c_int, [c_void_p, H, c_int]),
    "les_netcdf_write_arrays": (c_int, [c_void_p, c_int] + [c_void_p] * 7),
    "les_netcdf_close": (c_int, [c_void_p]),
    "les_anime": (c_int, [H, c_int, c_int, c_int, ctypes.c_char_p, ctypes.c_char_p]),
    "les_tiled_create": (c_int, [ctypes.POINTER(LesConfig), ctypes.POINTER(c_void_p)]),
    "les_tiled_destroy": (c_int, [c_void_p]),
    "les_tiled_last_error": (ctypes.c_char_p, [c_void_p]),
    "les_tiled_sync": (c_int, [c_void_p]),
    "les_tiled_upload": (c_int, [c_void_p, c_int, c_void_p]),
    "les_tiled_download": (c_int, [c_void_p, c_int, c_void_p]),
    "les_tiled_init_coefficients": (c_int, [c_void_p]),
    "les_tiled_ifdata": (c_int, [c_void_p]),
    "les_tiled_step": (c_int, [c_void_p, c_int]),
    "les_tiled_run": (c_int, [c_void_p, c_int, c_int]),
    "les_adj_create": (c_int, [ctypes.POINTER(LesConfig), c_int, ctypes.POINTER(c_void_p)]),
    "les_adj_destroy": (c_int, [c_void_p]),
    "les_adj_last_error": (ctypes.c_char_p, [c_void_p]),
    "les_adj_upload": (c_int, [c_void_p, c_int, c_void_p]),
    "les_adj_download": (c_int, [c_void_p, c_int, c_void_p]),
    "les_adj_upload_adjoint": (c_int, [c_void_p, c_int, c_void_p]),
    "les_adj_download_adjoint": (c_int, [c_void_p, c_int, c_void_p]),
    "les_adj_set_sums": (c_int, [c_void_p, c_double, c_double]),
    "les_adj_zero_adjoints": (c_int, [c_void_p]),
    "les_adj_step": (c_int, [c_void_p, c_int]),
    "les_adj_reverse_step": (c_int, [c_void_p]),
    "les_adj_window": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int, ctypes.POINTER(c_double)]),
    "les_adj_set_slots": (c_int, [c_void_p, c_int]),
    "les_adj_set_transposes": (c_int, [c_void_p, c_int]),
    "les_anime_arrays": (c_int, [c_int] * 6 + [c_void_p] * 7 + [ctypes.c_char_p] * 2),
    "les_upload": (c_int, [H, c_int, c_void_p]),
    "les_download": (c_int, [H, c_int, c_void_p]),
    "les_device_ptr": (c_int, [H, c_int, ctypes.POINTER(c_void_p), ctypes.POINTER(c_int64)]),
    "les_init_coefficients": (c_int, [H]),
    "les_velnw": (c_int, [H]),
    "les_bondv1_init_uvw": (c_int, [H, c_int]),
    "les_velnw__bondv1_init_uvw": (c_int, [H, c_int]),
    "les_bondv1_calc_uout": (c_int, [H]),
    "les_bondv1_calc_uvw": (c_int, [H]),
    "les_velfg": (c_int, [H]),
    "les_feedbf": (c_int, [H]),
    "les_les_calc_sm": (c_int, [H]),
    "les_velfg__feedbf__les_calc_sm": (c_int, [H]),
    "les_les_bound_sm": (c_int, [H]),
    "les_les_calc_visc": (c_int, [H]),
    "les_adam": (c_int, [H]),
    "les_les_calc_visc__adam": (c_int, [H]),
    "les_les": (c_int, [H]),
    "les_press_rhsav": (c_int, [H]),
    "les_press_sor": (c_int, [H, c_int]),
    "les_press_sor_solve": (c_int, [H]),
    "les_press_pav": (c_int, [H]),
    "les_press_adj": (c_int, [H]),
    "les_press_boundp": (c_int, [H]),
    "les_press": (c_int, [H]),
    "les_bondfg": (c_int, [H]),
    "les_boundp1": (c_int, [H]),
    "les_boundp2": (c_int, [H]),
    "les_step": (c_int, [H, c_int]),
    "les_run": (c_int, [H, c_int, c_int]),
    "les_ifdata": (c_int, [H]),
    "les_get_scalars": (c_int, [H, ctypes.POINTER(LesScalars)]),
    "les_halo_size": (c_int64, [c_int, c_int, c_int, c_int, c_int, c_int]),
    "les_exchange_halo_read": (c_int, [H, ctypes.POINTER(c_int), c_int, c_int, c_int, c_void_p]),
    "les_exchange_halo_write": (c_int, [H, ctypes.POINTER(c_int), c_int, c_int, c_void_p]),
    "les_halo_read_arrays": (c_int, [H, ctypes.POINTER(c_void_p), c_int, c_int, c_int, c_int, c_int,
                                     c_int, c_void_p]),
    "les_halo_write_arrays": (c_int, [H, ctypes.POINTER(c_void_p), c_int, c_int, c_int, c_int, c_int,
                                      c_void_p]),
    "les_reduce_sum_exact": (c_int, [H, c_void_p, c_int64, P_float]),
    "les_reduce_sum_exact_noalloc": (c_int, [c_void_p, c_int64, c_void_p, c_void_p]),
    "les_time_kernel": (c_int, [H, c_int, c_int, P_float]),
    "les_autotune": (c_int, [H, c_int, ctypes.POINTER(LesTuning)]),
    "les_set_sor_tile": (c_int, [H, c_int, c_int]),
    "les_ensemble_create": (c_int, [ctypes.POINTER(c_void_p), c_int, ctypes.POINTER(c_void_p)]),
    "les_ensemble_set_mode": (c_int, [c_void_p, c_int]),
    "les_ensemble_batched": (c_int, [c_void_p]),
    "les_ensemble_run": (c_int, [c_void_p, c_int, c_int]),
    "les_ensemble_sync": (c_int, [c_void_p]),
    "les_ensemble_destroy": (c_int, [c_void_p]),
    "les_initialise_LES_kernel": (c_int, [c_void_p] * 42 + [c_float, c_int, c_int, c_int]),
    "les_run_LES_kernel": (c_int, [c_int, c_int, c_void_p]),
    "les_default_handle": (c_void_p, []),
    "les_finalise_LES_kernel": (c_int, []),
    "les_gen_grid": (c_int, [c_int, c_int, c_int, c_float, c_float, c_float, c_float, c_float]
                     + [c_void_p] * 7),
    "les_gen_terrain": (c_int, [c_int, c_int, c_float, c_float, c_uint64, c_float, c_float,
                                c_float, c_float, c_void_p]),
    "les_gen_masks": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                              c_void_p, c_void_p, c_void_p]),
    "les_gen_inflow": (c_int, [c_int, c_int, c_void_p, c_void_p, c_float, c_float, c_float,
                               c_uint64, c_void_p]),
    "les_initial_wind_profile": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "les_gis_read_zbm": (c_int, [ctypes.c_char_p, c_int, c_int, c_void_p]),
    "les_aveflow": (c_int, [H, c_int, c_int, c_int]),
    "les_aveflow_download": (c_int, [H, c_int, c_void_p]),
    "les_set_terrain": (c_int, [H, c_void_p]),
}

_lib = None


def _bind_to_torch_runtime():
    """PyTorch-ROCm wheels ship their own HIP runtime under the same SONAME
    (libamdhip64.so.7).  Loading torch first makes libles_hip.so bind to that copy, so torch
    tensors and this library share one runtime and one device context; loaded the other way
    round, the process would hold two HIP runtimes and the second to initialise finds no device.
    Without torch installed, libles_hip.so uses the system ROCm runtime."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """Loads libles_hip.so (built in-tree); raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libles_hip.so is not built (%s); run `python -m les_halo_amd.build`"
                               % LIB_PATH)
        _bind_to_torch_runtime()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in PROTOS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ptr(a):
    """Address of a contiguous numpy array (host) as c_void_p."""
    return ctypes.c_void_p(a.ctypes.data)

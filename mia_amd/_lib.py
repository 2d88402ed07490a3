"""ctypes binding of libmia_hip.so (include/mia_hip.h).

The library is the only compute path: if it is missing this module raises,
there is no fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmia_hip.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "mia_hip.h")

OK = 0
ERR_ARG = -1
ERR_SIZE = -2
ERR_STATE = -3
ERR_DEVICE = -4
ERR_LIMIT = -5

GRAD_REFERENCE = 0
GRAD_EXACT = 1

_c_int_p = ctypes.POINTER(ctypes.c_int)
_u32_p = ctypes.POINTER(ctypes.c_uint32)
_dbl_p = ctypes.POINTER(ctypes.c_double)
_flt_p = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p

# name -> (restype, argtypes); must match include/mia_hip.h
SIGNATURES = {
    "mia_hip_last_error": (ctypes.c_char_p, []),
    "mia_hip_spline_csize": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_float, ctypes.c_int]),
    "mia_hip_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "mia_hip_ctx_destroy": (ctypes.c_int, [_vp]),
    "mia_hip_setup": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _u32_p, _u32_p, ctypes.c_int]),
    "mia_hip_setup_grid": (ctypes.c_int, [_vp, _u32_p]),
    "mia_hip_set_images": (ctypes.c_int, [_vp, ctypes.c_int, _flt_p, _flt_p]),
    "mia_hip_set_images_device": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp]),
    "mia_hip_get_stage1_path": (ctypes.c_int, [_vp, ctypes.c_int, _c_int_p]),
    "mia_hip_set_cost": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_int]),
    "mia_hip_n_params": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t),
                                        ctypes.POINTER(ctypes.c_size_t)]),
    "mia_hip_cost_grad": (ctypes.c_int, [_vp, _dbl_p, _dbl_p, _dbl_p]),
    "mia_hip_cost_grad_device": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "mia_hip_device_buffers": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "mia_hip_warp": (ctypes.c_int, [_vp, ctypes.c_int, _dbl_p, _flt_p]),
    "mia_hip_graph_capture": (ctypes.c_int, [_vp]),
    "mia_hip_graph_params": (ctypes.c_int, [_vp, ctypes.POINTER(_dbl_p)]),
    "mia_hip_graph_outputs": (ctypes.c_int, [_vp, ctypes.POINTER(_dbl_p), ctypes.POINTER(_dbl_p)]),
    "mia_hip_graph_run": (ctypes.c_int, [_vp]),
    "mia_hip_optimize": (ctypes.c_int, [_vp, _dbl_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                        _vp, _vp, _vp, _dbl_p]),
    "mia_hip_optimize_many": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(_dbl_p), ctypes.c_int,
                                             ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.POINTER(_dbl_p),
                                             ctypes.POINTER(_dbl_p), ctypes.POINTER(_c_int_p), _dbl_p]),
    "mia_hip_set_optimize_launch": (ctypes.c_int, [_vp, ctypes.c_int]),
    "mia_hip_synchronize": (ctypes.c_int, [_vp]),
    "mia_hip_get_stream": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "mia_hip_profile_enable": (ctypes.c_int, [_vp, ctypes.c_int]),
    "mia_hip_profile_read": (ctypes.c_int, [_vp, _dbl_p, _c_int_p, ctypes.c_int, _c_int_p]),
    "mia_hip_stage_name": (ctypes.c_char_p, [ctypes.c_int]),
    "mia_hip_device_arch": (ctypes.c_int, [ctypes.c_int, ctypes.c_char_p, ctypes.c_int]),
    "mia_hip_warp_variants": (ctypes.c_int, [_c_int_p, ctypes.c_int, _c_int_p]),
    "mia_hip_warp2d_variants": (ctypes.c_int, [_c_int_p, ctypes.c_int, _c_int_p]),
    "mia_hip_set_stage1_path": (ctypes.c_int, [_vp, ctypes.c_int]),
    "mia_hip_autotune": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_char_p, _vp]),
    "mia_hip_set_warp_variant": (ctypes.c_int, [_vp, ctypes.c_int]),
    "mia_hip_get_warp_variant": (ctypes.c_int, [_vp, ctypes.c_int, _c_int_p]),
    "mia_hip_autotune_warp": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_char_p, _vp]),
    "mia_hip_load_warp_tuning": (ctypes.c_int, [ctypes.c_int, ctypes.c_char_p]),
    "mia_hip_warp_tuning_source": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
    "mia_hip_scale_image": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _u32_p, _flt_p, _u32_p, _flt_p]),
    "mia_hip_image_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _u32_p, _flt_p, ctypes.POINTER(_vp)]),
    "mia_hip_image_scale": (ctypes.c_int, [_vp, _u32_p, ctypes.POINTER(_vp)]),
    "mia_hip_image_read": (ctypes.c_int, [_vp, _flt_p]),
    "mia_hip_image_destroy": (ctypes.c_int, [_vp]),
    "mia_hip_set_images_image": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp]),
    "mia_hip_divcurl_create": (ctypes.c_int, [ctypes.c_int, _u32_p, _flt_p, ctypes.c_double, ctypes.c_double,
                                              ctypes.POINTER(_vp)]),
    "mia_hip_divcurl_evaluate": (ctypes.c_int, [_vp, _flt_p, _dbl_p, _vp]),
    "mia_hip_divcurl_destroy": (ctypes.c_int, [_vp]),
    "mia_hip_set_penalty": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_double]),
    "mia_hip_cost_grad_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _u32_p, _u32_p, ctypes.c_int,
                                                 ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_int,
                                                 _vp, _vp, _vp, _vp, _vp, _vp]),
}

NOALLOC_SCRATCH_BYTES = 64 << 20  # MIA_HIP_NOALLOC_SCRATCH_BYTES

_lib = None


class MiaHipError(RuntimeError):
    """A failed C-ABI call; `code` is the library's status code."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libmia_hip.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status):
    """Map a status to the reference's exception types: invalid_argument ->
    ValueError, size mismatch (runtime_error) -> RuntimeError, others ->
    MiaHipError."""
    if status == OK:
        return
    msg = load().mia_hip_last_error().decode(errors="replace")
    if status == ERR_ARG:
        raise ValueError(msg)
    if status == ERR_SIZE:
        raise RuntimeError(msg)
    raise MiaHipError(status, msg)

"""ctypes binding of libdensecrf_ho_amd.so (declarations: include/densecrf_ho_amd.h)."""

from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_NAME = "libdensecrf_ho_amd.so"

_vp = C.c_void_p
_i = C.c_int
_f = C.c_float
_d = C.c_double
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


class QPStats(C.Structure):
    _fields_ = [("iterations", _i), ("updates", _i), ("capacity", _i), ("energy", C.POINTER(_d)), ("num", C.POINTER(_d)),
                ("denom", C.POINTER(_d)), ("step", C.POINTER(_d))]


class LPParams(C.Structure):
    _fields_ = [("prox_reg_const", _f), ("dual_gap_tol", _f), ("prox_energy_tol", _f), ("prox_max_iter", _i),
                ("fw_max_iter", _i), ("qp_max_iter", _i), ("qp_tol", _f), ("qp_const", _f), ("best_int", _i),
                ("accel_prox", _i), ("work_set_size", _i), ("approx_fw_iter", _i),
                ("less_confident_percent", _f), ("confidence_tol", _f)]


class LPStats(C.Structure):
    _fields_ = [("outer", _i), ("inner", _i), ("capacity", _i), ("qp_sweeps", _ip), ("dual_gap", C.POINTER(_d)),
                ("delta", C.POINTER(_d)), ("int_energy", C.POINTER(_d)), ("init_lp_energy", _d),
                ("init_int_energy", _d)]


# name -> (restype, argtypes); every exported symbol of the header
SIGNATURES = {
    "dcrf_last_error": (C.c_char_p, []),
    "dcrf_version": (C.c_char_p, []),
    "dcrf_set_reference_path": (_i, [_i]),
    "dcrf_reference_path": (_i, []),
    "dcrf_lattice_create": (_i, [_vp, _i, _i, _i, C.POINTER(_vp)]),
    "dcrf_lattice_destroy": (_i, [_vp]),
    "dcrf_lattice_info": (_i, [_vp, _ip, _ip]),
    "dcrf_lattice_export": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "dcrf_lattice_compute": (_i, [_vp, _vp, _vp, _i, _i, _i]),
    "dcrf_lattice_compute_upper_minus_lower_ord": (_i, [_vp, _vp, _vp, _i, _i]),
    "dcrf_lattice_compute_upper_minus_lower_ord_restricted": (_i, [_vp, _vp, _vp, _vp, _i, _vp, _i, _i, _i]),
    "dcrf_pairwise_create": (_i, [_vp, _i, _i, _f, _i, _i, C.POINTER(_vp)]),
    "dcrf_pairwise_destroy": (_i, [_vp]),
    "dcrf_pairwise_apply": (_i, [_vp, _vp, _vp, _i, _i, _i]),
    "dcrf_pairwise_apply_upper_minus_lower_ord": (_i, [_vp, _vp, _vp, _i, _i]),
    "dcrf_lattice_gradient": (_i, [_vp, _vp, _vp, _vp, _i, _i]),
    "dcrf_pairwise_set_kernel_type": (_i, [_vp, _i]),
    "dcrf_pairwise_kernel_nparams": (_i, [_vp]),
    "dcrf_pairwise_kernel_parameters": (_i, [_vp, _vp]),
    "dcrf_pairwise_set_kernel_parameters": (_i, [_vp, _vp, _i]),
    "dcrf_pairwise_kernel_gradient": (_i, [_vp, _vp, _vp, _i, _vp, _i]),
    "dcrf_crf_pairwise_kernel_gradient": (_i, [_vp, _i, _vp, _vp, _vp, _i]),
    "dcrf_crf_pairwise_set_kernel_type": (_i, [_vp, _i, _i]),
    "dcrf_crf_pairwise_kernel_nparams": (_i, [_vp, _i, _ip]),
    "dcrf_lattice_compute_upper_minus_lower_dc": (_i, [_vp, _vp, _i, _i, _i, _i, _i]),
    "dcrf_pairwise_apply_upper_minus_lower_dc": (_i, [_vp, _vp, _vp, _i, _i]),
    "dcrf_crf_apply_upper_minus_lower_dc": (_i, [_vp, _i, _vp, _vp, _i]),
    "dcrf_crf_unary_energy": (_i, [_vp, _vp, _vp, _i]),
    "dcrf_crf_energy_higher_order": (_i, [_vp, _vp, _i, C.POINTER(C.c_double)]),
    "dcrf_crf_pairwise_energy_true": (_i, [_vp, _vp, _i, _vp, _i]),
    "dcrf_pairwise_apply_bf": (_i, [_vp, _vp, _vp, _i, _i]),
    "dcrf_pairwise_apply_upper_minus_lower_bf": (_i, [_vp, _vp, _vp, _i, _i]),
    "dcrf_pairwise_apply_upper_minus_lower_bf_ord": (_i, [_vp, _vp, _vp, _vp, _i, _i]),
    "dcrf_pairwise_apply_upper_minus_lower_ord_restricted": (_i, [_vp, _vp, _vp, _vp, _i, _vp, _i, _i, _i]),
    "dcrf_crf_create": (_i, [_i, _i, C.POINTER(_vp)]),
    "dcrf_crf_destroy": (_i, [_vp]),
    "dcrf_crf_set_unary": (_i, [_vp, _vp, _i]),
    "dcrf_crf_set_unary_logistic": (_i, [_vp, _vp, _vp, _i, _i]),
    "dcrf_logistic_unary": (_i, [_vp, _vp, _i, _i, _i, _vp, _i]),
    "dcrf_logistic_unary_gradient": (_i, [_vp, _vp, _i, _i, _i, _vp, _i]),
    "dcrf_crf_add_pairwise_potential": (_i, [_vp, _vp]),
    "dcrf_pairwise_set_compatibility": (_i, [_vp, _i, _vp, _i]),
    "dcrf_crf_add_pairwise": (_i, [_vp, _vp, _i, _f, _i, _i]),
    "dcrf_crf_add_pairwise_compat": (_i, [_vp, _vp, _i, _i, _vp, _i, _i]),
    "dcrf_crf_pairwise_nparams": (_i, [_vp, _i, _ip]),
    "dcrf_crf_pairwise_parameters": (_i, [_vp, _i, _vp]),
    "dcrf_crf_pairwise_set_parameters": (_i, [_vp, _i, _vp]),
    "dcrf_crf_pairwise_gradient": (_i, [_vp, _i, _vp, _vp, _vp, _i]),
    "dcrf_crf_meanfield": (_i, [_vp, _i, _vp, _i]),
    "dcrf_crf_meanfield_backward": (_i, [_vp, _vp, _vp, _vp, _i]),
    "dcrf_crf_meanfield_id": (_i, [_vp, C.POINTER(C.c_longlong)]),
    "dcrf_crf_cross_entropy": (_i, [_vp, _vp, _vp, _vp, C.POINTER(_d), _i]),
    "dcrf_crf_add_pairwise_gaussian": (_i, [_vp, _i, _i, _f, _f, _f]),
    "dcrf_crf_add_pairwise_bilateral": (_i, [_vp, _i, _i, _f, _f, _f, _f, _f, _vp, _f]),
    "dcrf_crf_add_superpixel_layer": (_i, [_vp, _vp, _i, _vp, _vp]),
    "dcrf_crf_apply": (_i, [_vp, _i, _vp, _vp, _i, _i, _i]),
    "dcrf_crf_apply_upper_minus_lower_ord": (_i, [_vp, _i, _vp, _vp, _i]),
    "dcrf_crf_apply_bf": (_i, [_vp, _i, _vp, _vp, _i]),
    "dcrf_crf_apply_upper_minus_lower_bf": (_i, [_vp, _i, _vp, _vp, _i]),
    "dcrf_crf_apply_upper_minus_lower_bf_ord": (_i, [_vp, _i, _vp, _vp, _vp, _i]),
    "dcrf_crf_apply_upper_minus_lower_ord_restricted": (_i, [_vp, _i, _vp, _vp, _vp, _i, _vp, _i, _i]),
    "dcrf_crf_unary_init": (_i, [_vp, _vp, _i]),
    "dcrf_crf_qp": (_i, [_vp, _vp, _vp, _i, _i, _i, C.POINTER(QPStats)]),
    "dcrf_lp_params_default": (None, [C.POINTER(LPParams)]),
    "dcrf_crf_lp": (_i, [_vp, _vp, _vp, C.POINTER(LPParams), _i, C.POINTER(LPStats)]),
    "dcrf_crf_current_map": (_i, [_vp, _vp, _vp, _i]),
    "dcrf_crf_assignment_energy_true": (_i, [_vp, _vp, _i, C.POINTER(_d)]),
    "dcrf_crf_assignment_energy_higher_order": (_i, [_vp, _vp, _i, C.POINTER(_d)]),
    "dcrf_crf_lp_energy": (_i, [_vp, _vp, _i, C.POINTER(_d)]),
    "dcrf_crf_qp_value": (_i, [_vp, _vp, _i, C.POINTER(_d)]),
    "dcrf_crf_ucondgrad": (_i, [_vp, _vp, _vp, _i]),
    "dcrf_batch_create": (_i, [_i, _i, _vp, C.POINTER(_vp)]),
    "dcrf_batch_create_2d": (_i, [_i, _i, _vp, _vp, C.POINTER(_vp)]),
    "dcrf_batch_destroy": (_i, [_vp]),
    "dcrf_batch_pixel_offsets": (_i, [_vp, _vp]),
    "dcrf_batch_set_graphs": (_i, [_vp, _i]),
    "dcrf_batch_graph_captures": (_i, [_vp, _vp]),
    "dcrf_batch_set_value_type": (_i, [_vp, _i]),
    "dcrf_crf_set_value_type": (_i, [_vp, _i]),
    "dcrf_batch_set_unary": (_i, [_vp, _i, _vp, _i]),
    "dcrf_batch_add_pairwise": (_i, [_vp, _vp, _i, _f, _i, _i]),
    "dcrf_batch_add_pairwise_gaussian": (_i, [_vp, _f, _f, _f]),
    "dcrf_batch_add_pairwise_bilateral": (_i, [_vp, _f, _f, _f, _f, _f, _vp, _f]),
    "dcrf_batch_unary_init": (_i, [_vp, _vp, _i]),
    "dcrf_batch_qp": (_i, [_vp, _vp, _vp, _i, _i, _i, _vp]),
    "dcrf_batch_lp": (_i, [_vp, _vp, _vp, C.POINTER(LPParams), _i, _vp, _i]),
    "dcrf_batch_assignment_energy_true": (_i, [_vp, _vp, _i, _vp]),
    "dcrf_io_load_unary": (_i, [C.c_char_p, _vp, C.c_size_t, _ip, _ip, _ip]),
    "dcrf_io_load_image": (_i, [C.c_char_p, _vp, C.c_size_t, _ip, _ip]),
    "dcrf_io_save_map": (_i, [_vp, _i, _i, _i, C.c_char_p]),
    "dcrf_superpixel_terms": (_i, [_vp, _i, _vp, _i, _f, _f, _vp, _vp]),
    "dcrf_less_confident_pixels": (_i, [_vp, _i, _i, _f, _vp, _ip, _i]),
    "dcrf_less_confident_pixels_noalloc": (_i, [_vp, _i, _i, _f, _vp, _ip]),
    "dcrf_malloc": (_i, [C.POINTER(_vp), C.c_size_t]),
    "dcrf_free": (_i, [_vp]),
    "dcrf_memcpy": (_i, [_vp, _vp, C.c_size_t, _i]),
    "dcrf_device_synchronize": (_i, []),
    "dcrf_profile_enable": (_i, [_i]),
    "dcrf_profile_filter": (_i, [C.c_char_p]),
    "dcrf_profile_report": (_i, [C.c_char_p, _i]),
    "dcrf_autotune_enable": (_i, [_i]),
    "dcrf_autotune_report": (_i, [C.c_char_p, _i]),
}

_lib = None


def lib_path() -> str:
    return os.environ.get("DCRF_LIB", os.path.join(HERE, _LIB_NAME))


def load():
    """Load the engine library; raises if it has not been built."""
    global _lib
    if _lib is None:
        p = lib_path()
        if not os.path.exists(p):
            raise RuntimeError(f"densecrf_ho_amd: engine library not built: {p} "
                               "(run `python -m densecrf_ho_amd.build` or __graft_entry__.build())")
        L = C.CDLL(p)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib

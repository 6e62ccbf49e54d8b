"""ctypes binding of libsimkit_amd.so (the C ABI declared in include/simkit_amd.h).

The library is loaded from this package directory only. If it is missing or
no GPU is visible, every entry point raises: there is no CPU fallback.
"""

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsimkit_amd.so")

SK_OK = 0
SK_ERR_INVALID = -1
SK_ERR_HIP = -2
SK_ERR_LS_FAIL = -3
SK_ERR_PCG_BREAKDOWN = -4
SK_ERR_PCG_MAXIT = -5
SK_ERR_NONFINITE = -6
SK_ERR_INVERTED = -7
SK_ERR_NEWTON_MAXIT = -8
SK_ERR_UNSUPPORTED = -9
SK_ERR_EIG_MAXIT = -10

SK_SNH = 0
SK_ARAP = 1
SK_LINEAR = 2
SK_MMNH = 3
SK_FCR = 4
SK_PREC_FP64 = 0
SK_PREC_MIXED = 1

SK_PSD_NONE = 0
SK_PSD_PROJ = 1
SK_PSD_ABS = 2

MATERIALS = {
    "snh": SK_SNH, "stable_neo_hookean": SK_SNH, "stable-neo-hookean": SK_SNH,
    "arap": SK_ARAP,
    "linear": SK_LINEAR, "linear-elasticity": SK_LINEAR,
    "mmnh": SK_MMNH, "macklin-mueller-neo-hookean": SK_MMNH,
    "fcr": SK_FCR,
}


class SimkitAmdError(RuntimeError):
    """A non-zero status from the HIP library."""

    def __init__(self, status, message):
        super().__init__(f"simkit_amd status {status}: {message}")
        self.status = status


class SkSimParams(ctypes.Structure):
    _fields_ = [
        ("material", ctypes.c_int),
        ("psd", ctypes.c_int),
        ("h", ctypes.c_double),
        ("newton_rtol", ctypes.c_double),
        ("newton_max_iter", ctypes.c_int),
        ("pcg_rtol", ctypes.c_double),
        ("pcg_max_iter", ctypes.c_int),
        ("ls_alpha", ctypes.c_double),
        ("ls_beta", ctypes.c_double),
        ("ls_max_iter", ctypes.c_int),
        ("ls_threshold", ctypes.c_double),
        ("timing", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("inner_rtol", ctypes.c_double),
        ("max_refine", ctypes.c_int),
        ("spmv_blocks", ctypes.c_int),
        ("update_blocks", ctypes.c_int),
        ("pcg_chunk", ctypes.c_int),
        ("max_inverted", ctypes.c_int),
        ("pcg_variant", ctypes.c_int),
        ("psd_jacobi", ctypes.c_int),
    ]


class SkStepInfo(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int),
        ("newton_iters", ctypes.c_int),
        ("pcg_iters", ctypes.c_int64),
        ("ls_trials", ctypes.c_int),
        ("g0_norm", ctypes.c_double),
        ("g_norm", ctypes.c_double),
        ("ms_gradient", ctypes.c_double),
        ("ms_hessian", ctypes.c_double),
        ("ms_pcg", ctypes.c_double),
        ("ms_line_search", ctypes.c_double),
        ("refine_steps", ctypes.c_int),
        ("guess_used", ctypes.c_int),
        ("guess_dE", ctypes.c_double),
        ("guess_g_norm", ctypes.c_double),
        ("stalled", ctypes.c_int),
        ("inverted_elems", ctypes.c_int),
        ("polar_fp64_elems", ctypes.c_int64),
        ("start_dE", ctypes.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class SkBatchParams(ctypes.Structure):
    _fields_ = [
        ("material", ctypes.c_int),
        ("psd", ctypes.c_int),
        ("h", ctypes.c_double),
        ("newton_rtol", ctypes.c_double),
        ("newton_max_iter", ctypes.c_int),
        ("pcg_rtol", ctypes.c_double),
        ("pcg_max_iter", ctypes.c_int),
        ("inner_rtol", ctypes.c_double),
        ("ls_alpha", ctypes.c_double),
        ("ls_beta", ctypes.c_double),
        ("ls_max_iter", ctypes.c_int),
        ("ls_threshold", ctypes.c_double),
        ("block", ctypes.c_int),
        ("lanes", ctypes.c_int),
        ("lds", ctypes.c_int),
        ("use_graph", ctypes.c_int),
    ]


class SkBatchInfo(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int),
        ("newton_iters", ctypes.c_int),
        ("pcg_iters", ctypes.c_int64),
        ("ls_trials", ctypes.c_int),
        ("refine_steps", ctypes.c_int),
        ("g0_norm", ctypes.c_double),
        ("g_norm", ctypes.c_double),
        ("stalled", ctypes.c_int),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class SkModalInfo(ctypes.Structure):
    _fields_ = [
        ("sweeps", ctypes.c_int),
        ("spmvs", ctypes.c_int64),
        ("block", ctypes.c_int),
        ("max_rel_residual", ctypes.c_double),
        ("spectrum_bound", ctypes.c_double),
        ("ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_D = ctypes.c_double
_INT = ctypes.c_int

# name -> argtypes (restype is int unless listed in _RESTYPES)
SIGNATURES = {
    "sk_last_error": [],
    "sk_status_string": [_INT],
    "sk_version": [],
    "sk_tet_operator": [_P, _P, _I64, _I64, _P, _P, _P],
    "sk_deformation_gradient": [_P, _P, _P, _I64, _P, _P],
    "sk_element_F": [_INT, _P, _P, _P, _I64, _P, _P, _P, _P],
    "sk_psd_project": [_P, _I64, _INT, _INT, _P, _P],
    "sk_polar_svd": [_P, _I64, _P, _P, _P, _P, _P],
    "sk_rotation_gradient": [_P, _I64, _P, _P],
    "sk_mesh_create": [_P, _P, _I64, _I64, ctypes.POINTER(_P)],
    "sk_mesh_destroy": [_P],
    "sk_mesh_info": [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I64), ctypes.POINTER(_I64), ctypes.POINTER(_I64)],
    "sk_mesh_csr_pattern": [_P, _P, _P, _P],
    "sk_mesh_element_nnz": [_P, _P, _P],
    "sk_mesh_lumped_mass": [_P, _P, _P, _P, _P],
    "sk_mesh_energy": [_P, _INT, _P, _P, _P, _P, _P, ctypes.POINTER(_D), _P, _P],
    "sk_mesh_gradient": [_P, _INT, _P, _P, _P, _P, _P, _P, _P],
    "sk_mesh_hessian": [_P, _INT, _P, _P, _P, _P, _P, _INT, _P, _P],
    "sk_csr_spmv": [_I64, _P, _P, _P, _P, _P, _P],
    "sk_kinetic": [_I64, _P, _P, _P, _D, _D, _P, _P, _P, ctypes.POINTER(_D), _P, _P, _P, _P],
    "sk_kinetic_bdf2": [_I64, _P, _P, _P, _P, _P, _D, _P, _P, _P, ctypes.POINTER(_D), _P, _P, _P, _P],
    "sk_pcg_csr": [_I64, _P, _P, _P, _P, _P, _D, _I64, ctypes.POINTER(_I64), ctypes.POINTER(_D), _P],
    "sk_pcg_csr_noalloc": [_I64, _P, _P, _P, _P, _P, _D, _I64, ctypes.POINTER(_I64), ctypes.POINTER(_D), _P],
    "sk_sim_create": [_P, ctypes.POINTER(SkSimParams), _P, _P, _P, _P, _P, _P, ctypes.POINTER(_P)],
    "sk_sim_destroy": [_P],
    "sk_sim_step": [_P, _P, _P, _P, _P, ctypes.POINTER(SkStepInfo), _P],
    "sk_modal_basis": [_P, _INT, _P, _P, _P, _P, _P, _P, _INT, _D, _D, _INT, _INT, ctypes.c_uint64, _P, _P,
                       ctypes.POINTER(SkModalInfo), _P],
    "sk_batch_create": [_P, _INT, _P, _P, _P, ctypes.POINTER(SkBatchParams), _P, _P, _P, _P, _P, _P, _P],
    "sk_batch_destroy": [_P],
    "sk_batch_step": [_P, _P, _P, _P, _P, _P, _P, _P],
    "sk_batch_set_shape": [_P, _INT, _INT, _INT],
    "sk_batch_set_coarse": [_P, _P, _INT, _INT],
    "sk_batch_autotune": [_P, _P, _P, _P, _INT, _P, _P, _P, _P, _P],
    "sk_dense_project": [_I64, _INT, _P, _P, _P, _P, _P, _P],
    "sk_dense_expand": [_I64, _INT, _P, _P, _P, _P, _P],
    "sk_axpby": [_I64, _D, _P, _D, _P, _P, _P],
    "sk_gemm_bf16": [_INT, _INT, _INT, _P, _I64, _I64, _P, _I64, _I64, _P, _I64, _P],
    "sk_mlp_create": [_INT, _P, _INT, _P],
    "sk_mlp_destroy": [_P],
    "sk_mlp_num_params": [_P],
    "sk_mlp_set_params": [_P, _P, _P],
    "sk_mlp_get_params": [_P, _P, _P],
    "sk_mlp_get_grad": [_P, _P, _P],
    "sk_mlp_forward": [_P, _P, _INT, _P, _P],
    "sk_mlp_predict": [_P, _P, _INT, _P, _P, _P, _P, _P, _P],
    "sk_mlp_backward": [_P, _P, _P, _INT, ctypes.POINTER(_D), _P],
    "sk_mlp_adam": [_P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, _P],
    "sk_sim_step_guess": [_P, _P, _P, _P, _P, _P, ctypes.POINTER(SkStepInfo), _P],
    "sk_sim_step_guesses": [_P, _P, _P, _P, ctypes.c_int, _P, _P, ctypes.POINTER(SkStepInfo), _P],
    "sk_sim_step_bdf2": [_P, _P, _P, _P, _P, _P, _P, ctypes.POINTER(SkStepInfo), _P],
    "sk_sim_spmv_reps": [_P, _INT, _P],
    "sk_sim_hess_reps": [_P, _INT, _P],
    "sk_sim_set_coarse": [_P, _P, _INT],
    "sk_coarse_order": [_P, _I64, _INT, _P, _P, _P],
    "sk_sell_order": [_P, _I64, _I64, _P],
    "sk_sim_coarse_info": [_P, _P, _P, _P, _P],
    "sk_sim_set_order": [_P, _P],
}
_RESTYPES = {"sk_last_error": ctypes.c_char_p, "sk_status_string": ctypes.c_char_p, "sk_mlp_num_params": ctypes.c_int64}

_lock = threading.Lock()
_handle = None


def load():
    """Load the library (once). Raises if it has not been built."""
    global _handle
    with _lock:
        if _handle is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    f"simkit_amd: HIP library not found at {LIB_PATH}; build it with "
                    "`python -m simkit_amd.build` (there is no CPU fallback)")
            h = ctypes.CDLL(LIB_PATH)
            for name, args in SIGNATURES.items():
                fn = getattr(h, name)
                fn.argtypes = args
                fn.restype = _RESTYPES.get(name, ctypes.c_int)
            _handle = h
    return _handle


def check(status):
    if status != SK_OK:
        msg = load().sk_last_error().decode(errors="replace")
        raise SimkitAmdError(status, msg)
    return status


def call(name, *args):
    """Call an entry point and raise SimkitAmdError on a non-zero status."""
    return check(getattr(load(), name)(*args))

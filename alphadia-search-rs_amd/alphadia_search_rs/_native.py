"""ctypes binding of libadsr.so, the C ABI declared in include/adsr.h.

The shared object is built in-tree next to this file.  Importing this module
fails with ImportError when it is missing: there is no Python or CPU
fallback for the selection path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ADSR_LIBRARY selects another build of the same library (the sanitizer
# build of the CPU tests); by default the in-tree libadsr.so.
LIB_PATH = os.environ.get("ADSR_LIBRARY") or os.path.join(_HERE, "libadsr.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is not built; build it with `make -C "
        f"{os.path.normpath(os.path.join(_HERE, '..', 'csrc'))}` "
        "(or __graft_entry__.build())"
    )

# One HIP runtime per process.  libadsr.so needs libamdhip64.so.7; PyTorch
# ships its own copy under that soname.  Loading torch first makes the
# dynamic loader bind libadsr.so to torch's runtime, so device pointers,
# streams and device initialisation are shared.  Loaded the other way round,
# a process holds two runtimes, and the second one to initialise can find
# no device.
try:
    import torch  # noqa: F401
except ImportError:
    pass

lib = C.CDLL(LIB_PATH)

c_size = C.c_size_t
c_u64 = C.c_uint64
c_f32 = C.c_float
vp = C.c_void_p

ADSR_OK = 0
ADSR_ERR_INVALID_ARGUMENT = 1
ADSR_ERR_DEVICE = 2
ADSR_ERR_OUT_OF_MEMORY = 3
ADSR_ERR_UNSUPPORTED = 4


class ScoringParamsC(C.Structure):
    """struct adsr_scoring_params (include/adsr.h)."""

    _fields_ = [("mass_tolerance", c_f32), ("top_k_fragments", c_u64)]


class QuantParamsC(C.Structure):
    """struct adsr_quant_params (include/adsr.h)."""

    _fields_ = [("tolerance_ppm", c_f32), ("top_k_fragments", c_u64)]


class SelectionParamsC(C.Structure):
    """struct adsr_selection_params (include/adsr.h)."""

    _fields_ = [
        ("fwhm_rt", c_f32),
        ("kernel_size", c_u64),
        ("peak_length", c_u64),
        ("mass_tolerance", c_f32),
        ("rt_tolerance", c_f32),
        ("candidate_count", c_u64),
        ("top_k_fragments", c_u64),
        ("rt_tolerance_per_precursor", vp),
        ("sigma_per_window", vp),
        ("n_sigma_per_window", c_u64),
    ]


def _sig(name, restype, *argtypes):
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = list(argtypes)
    return fn


I = C.c_int
_sig("adsr_last_error", C.c_char_p)
_sig("adsr_selection_params_default", None, C.POINTER(SelectionParamsC))
_sig("adsr_device_available", I)
_sig("adsr_backend_set", I, C.c_char_p)
_sig("adsr_backend_clear", None)
_sig("adsr_backend_current", I, C.c_char_p, c_size)
_sig("adsr_backend_optimal", I, C.c_char_p, c_size)
_sig("adsr_dia_create", I, vp, vp, vp, vp, vp, vp, vp, c_size, vp, vp, c_size,
     vp, vp, C.POINTER(vp))
_sig("adsr_dia_create_device", I, vp, vp, vp, vp, vp, vp, vp, c_size, vp, vp,
     c_size, vp, vp, C.POINTER(vp))
_sig("adsr_dia_build_timings", I, vp, C.POINTER(C.c_double),
     C.POINTER(C.c_double), C.POINTER(C.c_double))
_sig("adsr_dia_upload", I, vp)
_sig("adsr_dia_free", None, vp)
_sig("adsr_dia_num_observations", c_size, vp)
_sig("adsr_dia_valid_observations", I, vp, c_f32, vp, c_size, C.POINTER(c_size))
_sig("adsr_dia_memory_footprint_bytes", c_u64, vp)
_sig("adsr_dia_rt_values", c_size, vp, vp, c_size)
_sig("adsr_dia_cycle_shape", None, vp, vp)
_sig("adsr_dia_cycle_copy", None, vp, vp)
_sig("adsr_dia_rt_index", c_size, vp, vp, c_size, C.POINTER(c_f32))
_sig("adsr_dia_observation", I, vp, c_size, vp, C.POINTER(c_u64), vp,
     C.POINTER(c_size), vp, vp, C.POINTER(c_size))
_sig("adsr_mz_grid", c_size, vp, c_size)
_sig("adsr_lib_create", I, c_size, vp, vp, vp, vp, vp, vp, vp, vp, c_size, vp,
     vp, vp, vp, vp, vp, vp, vp, vp, C.POINTER(vp))
_sig("adsr_lib_upload", I, vp)
_sig("adsr_lib_free", None, vp)
_sig("adsr_lib_num_precursors", c_size, vp)
_sig("adsr_lib_num_fragments", c_size, vp)
_sig("adsr_select", I, vp, vp, C.POINTER(SelectionParamsC), C.POINTER(vp))
_sig("adsr_select_timed", I, vp, vp, C.POINTER(SelectionParamsC), I, I,
     C.POINTER(I), C.POINTER(I), C.POINTER(C.c_double), C.POINTER(C.c_double),
     C.POINTER(C.c_double), C.POINTER(vp))
_sig("adsr_select_algorithmic_bytes", I, vp, vp, C.POINTER(SelectionParamsC),
     C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double))
_sig("adsr_select_scores", I, vp, vp, C.POINTER(SelectionParamsC), c_size,
     c_size, c_size, vp, vp, vp)
_sig("adsr_select_plan", I, vp, vp, C.POINTER(SelectionParamsC), c_size, c_size,
     vp, vp, vp, vp, c_size)
_sig("adsr_diff_create", I, vp, vp, C.POINTER(SelectionParamsC), C.POINTER(vp))
_sig("adsr_diff_forward", I, vp, vp, vp, vp, vp)
_sig("adsr_diff_backward", I, vp, vp, vp, vp)
_sig("adsr_diff_timings", I, vp, C.POINTER(C.c_double), C.POINTER(C.c_double))
_sig("adsr_hinge_loss", I, vp, c_size, vp, vp, c_size, c_f32, vp, vp, vp)
_sig("adsr_diff_forward_async", I, vp, vp, vp, vp, vp, vp)
_sig("adsr_diff_free", None, vp)
_sig("adsr_candidates_len", c_size, vp)
_sig("adsr_candidates_export", I, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp)
_sig("adsr_candidates_from_arrays", I, c_size, vp, vp, vp, vp, vp, vp, vp, vp,
     vp, C.POINTER(vp))
_sig("adsr_candidates_free", None, vp)
_sig("adsr_select_device", I, vp, vp, C.POINTER(SelectionParamsC), C.POINTER(vp))
_sig("adsr_feature_name", C.c_char_p, c_size)
_sig("adsr_scoring_params_default", None, C.POINTER(ScoringParamsC))
_sig("adsr_score", I, vp, vp, vp, C.POINTER(ScoringParamsC), C.POINTER(vp))
_sig("adsr_score_device", I, vp, vp, vp, C.POINTER(ScoringParamsC), C.POINTER(vp))
_sig("adsr_features_len", c_size, vp)
_sig("adsr_features_export", I, vp, vp, vp, vp)
_sig("adsr_features_timings", I, vp, C.POINTER(C.c_double), C.POINTER(C.c_double))
_sig("adsr_features_free", None, vp)
_sig("adsr_quant_params_default", None, C.POINTER(QuantParamsC))
_sig("adsr_quantify", I, vp, vp, vp, C.POINTER(QuantParamsC), C.POINTER(vp))
_sig("adsr_quantified_num_precursors", c_size, vp)
_sig("adsr_quantified_num_fragments", c_size, vp)
_sig("adsr_quantified_export_precursors", I, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp)
_sig("adsr_quantified_export_fragments", I, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
     vp, vp, vp)
_sig("adsr_quantified_timings", I, vp, C.POINTER(C.c_double), C.POINTER(C.c_double))
_sig("adsr_quantified_free", None, vp)
_sig("adsr_device_candidates_len", c_size, vp)
_sig("adsr_device_candidates_export", I, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp)
_sig("adsr_select_rows_noalloc", I, vp, vp, C.POINTER(SelectionParamsC), c_size, c_size,
     vp, vp, vp, vp, vp, vp, vp, vp, vp, c_size, C.POINTER(c_size), vp)
_sig("adsr_device_candidates_to_host", I, vp, C.POINTER(vp))
_sig("adsr_device_candidates_free", None, vp)
_sig("adsr_gaussian_kernel", c_size, c_f32, c_f32, c_u64, c_f32, vp, c_size)
_sig("adsr_gaussian_kernel_sigma", c_size, c_f32, c_u64, vp, c_size)
_sig("adsr_convolution", I, vp, c_size, vp, c_size, c_size, vp)
_sig("adsr_axis_log_dot_product", I, vp, c_size, c_size, vp, vp)
_sig("adsr_find_local_maxima", I, vp, c_size, c_u64, vp, vp, C.POINTER(c_size))
_sig("adsr_dense_xic", I, vp, c_f32, c_u64, c_u64, c_f32, vp, c_size, vp)
_sig("adsr_device_logf", I, vp, c_size, vp)
_sig("adsr_logf_host", c_f32, c_f32)


class AdsrError(RuntimeError):
    pass


def check(status: int) -> None:
    """Raise the Python exception that corresponds to an adsr_status."""
    if status == ADSR_OK:
        return
    msg = lib.adsr_last_error().decode("utf-8", "replace")
    if status == ADSR_ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    if status == ADSR_ERR_OUT_OF_MEMORY:
        raise MemoryError(msg)
    raise AdsrError(msg)


def ptr(a: np.ndarray | None):
    """Data pointer of a C-contiguous array (None -> NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(vp)


def device_available() -> bool:
    return bool(lib.adsr_device_available())

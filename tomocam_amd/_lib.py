"""ctypes binding of libtomocam_hip.so (the C-ABI in include/tomocam_hip.h).

The shared library is the product path: nothing here computes on the CPU.
``lib()`` raises if the library has not been built (``__graft_entry__.build()``
or ``make -C tomocam_amd/csrc``).
"""

from __future__ import annotations

import ctypes as C
import os
import threading

LIBNAME = "libtomocam_hip.so"
_HERE = os.path.dirname(os.path.abspath(__file__))
_lock = threading.Lock()
_lib = None

c_int_p = C.POINTER(C.c_int)
c_long_p = C.POINTER(C.c_long)
c_double_p = C.POINTER(C.c_double)
c_float_p = C.POINTER(C.c_float)
vp = C.c_void_p

# name -> argument types (every function returns int status)
SIGNATURES = {
    "tc_version": [c_int_p, c_int_p, c_int_p],
    "tc_set_num_of_gpus": [C.c_int],
    "tc_psf_create": [C.POINTER(vp), vp, C.c_int, C.c_int, vp],
    "tc_psf_create_from_array": [C.POINTER(vp), vp, C.c_int, vp],
    "tc_psf_get_array": [vp, vp, vp],
    "tc_psf_destroy": [vp],
    "tc_psf_info": [vp, c_int_p, c_int_p],
    "tc_toeplitz_fft_length": [C.c_int, c_int_p],
    "tc_graph_replays": [C.POINTER(C.c_long)],
    "tc_toeplitz_apply": [vp, vp, vp, C.c_int, c_double_p, vp],
    "tc_toeplitz_apply_mixed": [vp, vp, vp, C.c_int, c_double_p, vp],
    "tc_toeplitz_apply_noalloc": [vp, vp, vp, C.c_int, c_double_p, vp],
    "tc_psf_tune": [vp, C.c_int, C.c_int, c_int_p, vp],
    "tc_gradient2": [vp, vp, vp, vp, C.c_int, vp],
    "tc_function_value2": [vp, vp, vp, C.c_int, C.c_double, c_double_p, vp],
    "tc_add_total_var2": [vp, vp, C.c_int, C.c_int, C.c_int, C.c_float,
                          C.c_float, vp],
    "tc_add_total_var": [vp, vp, C.c_int, C.c_int, C.c_int, C.c_float,
                         C.c_float, vp],
    "tc_add_tv_hessian": [c_float_p, C.c_long, C.c_float, c_float_p],
    "tc_qggmrf_grad_free": [vp, vp, C.c_int, C.c_int, C.c_int, C.c_float,
                            C.c_float, vp],
    "tc_qggmrf_value_free": [vp, C.c_int, C.c_int, C.c_int, C.c_float,
                             C.c_float, c_double_p, vp],
    "tc_qggmrf_value_free_dev": [vp, C.c_int, C.c_int, C.c_int, C.c_float,
                                 C.c_float, vp, vp],
    "tc_run2": [vp, vp, vp, vp, C.c_int, C.c_int, C.c_double, C.c_float,
                C.c_float, C.c_double, C.c_int, C.c_int, c_double_p,
                c_double_p, c_int_p, vp],
    "tc_project": [vp, vp, C.c_int, C.c_int, vp, C.c_int, vp],
    "tc_backproject": [vp, vp, C.c_int, C.c_int, C.c_int, vp, C.c_int, vp],
    "tc_gradient": [vp, vp, vp, C.c_int, C.c_int, vp, C.c_int, vp],
    "tc_function_value": [vp, vp, C.c_int, C.c_int, vp, C.c_int, c_double_p,
                          vp],
    "tc_xerror": [vp, vp, C.c_long, c_double_p, vp],
    "tc_preproc": [vp, vp, C.c_int, C.c_int, C.c_int, C.c_float, c_int_p, vp],
    "tc_postproc": [vp, vp, C.c_int, C.c_int, C.c_int, vp],
    "tc_pad2d": [vp, vp, C.c_int, C.c_int, C.c_int, vp],
    "tc_mbir2": [vp, vp, vp, C.c_int, C.c_int, C.c_int, vp, C.c_float,
                 C.c_int, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int,
                 c_double_p, vp],
    "tc_mbir2_mixed": [vp, vp, vp, C.c_int, C.c_int, C.c_int, vp, C.c_float,
                       C.c_int, C.c_float, C.c_int, C.c_int, C.c_int,
                       c_double_p, vp],
    "tc_mbir": [vp, vp, vp, C.c_int, C.c_int, C.c_int, vp, C.c_float,
                C.c_int, C.c_float, C.c_float, C.c_float, C.c_int, c_double_p,
                vp],
    "tc_mbir2_batch": [C.c_int, vp, vp, c_int_p, c_int_p, c_int_p, vp, vp,
                       C.c_int, C.c_float, C.c_int, c_double_p, vp],
    "tc_projections_to_sinogram": [vp, vp, C.c_int, C.c_int, C.c_int,
                                   C.c_int, vp],
    "tc_mbir2_host": [vp, vp, C.c_int, C.c_int, C.c_int, vp, C.c_float,
                      C.c_int, C.c_float, C.c_float, C.c_float, C.c_int],
    "tc_project_f64": [vp, vp, C.c_int, C.c_int, vp, C.c_int, vp],
    "tc_backproject_f64": [vp, vp, C.c_int, C.c_int, C.c_int, vp, C.c_int,
                           vp],
    "tc_gradient_f64": [vp, vp, vp, C.c_int, C.c_int, vp, C.c_int, vp],
    "tc_function_value_f64": [vp, vp, C.c_int, C.c_int, vp, C.c_int,
                              c_double_p, vp],
    "tc_psf_create_f64": [C.POINTER(vp), vp, C.c_int, C.c_int, vp],
    "tc_psf_destroy_f64": [vp],
    "tc_toeplitz_apply_f64": [vp, vp, vp, C.c_int, vp],
    "tc_gradient2_f64": [vp, vp, vp, vp, C.c_int, vp],
    "tc_function_value2_f64": [vp, vp, vp, C.c_int, C.c_double, c_double_p,
                               vp],
    "tc_add_total_var2_f64": [vp, vp, C.c_int, C.c_int, C.c_int, C.c_double,
                              C.c_double, vp],
    "tc_add_total_var_f64": [vp, vp, C.c_int, C.c_int, C.c_int, C.c_double,
                             C.c_double, vp],
    "tc_qggmrf_grad_free_f64": [vp, vp, C.c_int, C.c_int, C.c_int,
                                C.c_double, C.c_double, vp],
    "tc_qggmrf_value_free_f64": [vp, C.c_int, C.c_int, C.c_int, C.c_double,
                                 C.c_double, c_double_p, vp],
    "tc_mbir2_f64": [vp, vp, vp, C.c_int, C.c_int, C.c_int, vp, C.c_double,
                     C.c_int, C.c_double, c_double_p, vp],
    "tc_mbir_f64": [vp, vp, vp, C.c_int, C.c_int, C.c_int, vp, C.c_double,
                    C.c_int, C.c_double, c_double_p, vp],
    "tc_psf_tune_length": [vp, C.c_int, C.POINTER(C.c_int), vp],
    "tc_xerror_f64": [vp, vp, C.c_long, c_double_p, vp],
    "tc_preproc_f64": [vp, vp, C.c_int, C.c_int, C.c_int, C.c_double,
                       C.POINTER(C.c_int), vp],
    "tc_postproc_f64": [vp, vp, C.c_int, C.c_int, C.c_int, vp],
    "tc_pad2d_f64": [vp, vp, C.c_int, C.c_int, C.c_int, vp],
    "tc_alloc_stats": [c_long_p, c_long_p, c_double_p],
    "tc_release_cached": [],
    "tc_ktimer_enable": [C.c_int, C.c_int],
    "tc_ktimer_stride": [C.c_int],
    "tc_ktimer_collect": [C.c_int, c_double_p, c_long_p, c_double_p],
}


class TomocamError(RuntimeError):
    """Raised when a C-ABI call returns a non-zero status."""

    def __init__(self, func: str, status: int, msg: str):
        super().__init__(f"{func} failed (status {status}): {msg}")
        self.status = status


def lib_path() -> str:
    """The in-tree build, or the file named by TOMOCAM_HIP_LIB (another
    build of the same library, e.g. with different compile-time options)."""
    return os.environ.get("TOMOCAM_HIP_LIB") or os.path.join(_HERE, LIBNAME)


def lib():
    """Loads (once) and returns the configured ctypes library."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError(
                f"{LIBNAME} is not built ({path} missing); run "
                "__graft_entry__.build() or make -C tomocam_amd/csrc")
        L = C.CDLL(path)
        L.tc_last_error.restype = C.c_char_p
        L.tc_last_error.argtypes = []
        _lib = L
    return _lib


_bound = {}


def fn(name: str):
    """Returns C-ABI function ``name`` with its signature applied."""
    f = _bound.get(name)
    if f is None:
        f = getattr(lib(), name)
        f.restype = C.c_int
        f.argtypes = SIGNATURES[name]
        _bound[name] = f
    return f


def check(func: str, status: int) -> None:
    if status != 0:
        msg = lib().tc_last_error().decode(errors="replace")
        if status == 1 or status == 5:
            raise ValueError(f"{func}: {msg}")
        raise TomocamError(func, status, msg)


def call(name: str, *args) -> None:
    """Calls C-ABI function ``name`` and raises on a non-zero status."""
    check(name, fn(name)(*args))

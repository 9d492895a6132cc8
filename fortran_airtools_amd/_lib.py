"""ctypes binding of libfair_hip.so (the C ABI declared in include/fair_hip.h).

The library is loaded on first use.  There is no fallback: if the shared object is missing or
no GPU is visible, every call raises FairError.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libfair_hip.so")

FAIR_OK, FAIR_ERR_ARG, FAIR_ERR_HIP, FAIR_ERR_OVERFLOW, FAIR_ERR_INTERNAL, FAIR_ERR_METHOD = range(6)


class FairError(RuntimeError):
    """A non-zero status from the library; `code` is the status, the text is the library's."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class Geom(C.Structure):
    _fields_ = [("n", C.c_int32), ("p", C.c_int32), ("nth", C.c_int32),
                ("r", C.c_double), ("dw", C.c_double), ("sd", C.c_double)]


class Opts(C.Structure):
    _fields_ = [("relaxpar", C.c_double), ("ubound", C.c_double), ("lbound", C.c_double),
                ("damp", C.c_double), ("verbose", C.c_int32), ("upb", C.c_int32),
                ("lwrb", C.c_int32), ("rowprobability", C.c_int32)]


_P = C.c_void_p
_I32, _I64, _D = C.c_int32, C.c_int64, C.c_double

# name -> argtypes (restype int unless listed in _RESTYPES)
SIGNATURES = {
    "fair_last_error": [],
    "fair_init": [C.c_int],
    "fair_set_stream": [_P],
    "fair_synchronize": [],
    "fair_version": [],
    "fair_fanlineartomo": [C.POINTER(Geom), _P, C.POINTER(_P)],
    "fair_paralleltomo": [C.POINTER(Geom), _P, C.POINTER(_P)],
    "fair_linspace": [_D, _D, _I32, _P],
    "fair_csr_create": [_I32, _I32, _I64, _P, _P, _P, C.POINTER(_P)],
    "fair_csr_info": [_P, C.POINTER(_I32), C.POINTER(_I32), C.POINTER(_I64)],
    "fair_csr_download": [_P, _P, _P, _P],
    "fair_csr_device_arrays": [_P, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)],
    "fair_csr_free": [_P],
    "fair_matrix_store": [C.c_char_p, _I32, _I32, _I64, _P, _P, _P, _I32],
    "fair_matrix_read_header": [C.c_char_p, C.POINTER(_I32), C.POINTER(_I32), C.POINTER(_I64), C.POINTER(_I32)],
    "fair_matrix_read": [C.c_char_p, _I64, _I32, _P, _P, _P],
    "fair_absorbance": [_P, _I64, _I64, _I64, _P],
    "fair_absorbance_dev": [_P, _I64, _I64, _I64, _P],
    "fair_phantom_ellipses": [_I32, _P, _P, _I32, _P, _P],
    "fair_phantom_bumps": [_I32, _P, _P, _I32, _P, _I32, _P],
    "fair_phantom_nearest": [_I32, _P, _P, _I32, _P, _P],
    "fair_phantom_last_disk": [_I32, _P, _P, _I32, _P, _P],
    "fair_phantom_stamp_disks": [_I32, _P, _P, _I32, _P, _D, _P],
    "fair_phantom_last_band": [_I32, _P, _P, _I32, _P, _P],
    "fair_normal_newton": [_I64, _P, _P],
    "fair_matvecmul": [_P, _P, _I32, _P, _I32],
    "fair_transvecmul": [_P, _P, _I32, _P, _I32],
    "fair_rowvecmult": [_P, _I32, _P, _I32, _P],
    "fair_row_norms": [_P, _I32, _P, _I32],
    "fair_column_norms": [_P, _I32, _P, _I32],
    "fair_transpose": [_P],
    "fair_multiply_by_scalar": [_P, _D],
    "fair_add_scalar_mult_row": [_P, _I32, _D, _P, _I32, _P, _P],
    "fair_matvecmul_dev": [_P, _P, _P],
    "fair_transvecmul_dev": [_P, _P, _P],
    "fair_spmv_rows_dev": [_P, _P, _P],
    "fair_sirt": [_P, _I32, _P, _I32, _P, C.POINTER(Opts), _P],
    "fair_sirt_dev": [_P, _I32, _P, _I32, _P, C.POINTER(Opts)],
    "fair_sirt_dev_noalloc": [_P, _I32, _P, _I32, _P, C.POINTER(Opts)],
    "fair_sirt_prepare": [_P, _I32],
    "fair_sirt_f32": [_P, _I32, _P, _I32, _P, C.POINTER(Opts), _P],
    "fair_sirt_f32_dev": [_P, _I32, _P, _I32, _P, C.POINTER(Opts)],
    "fair_sirt_batch_dev": [_P, _I32, _P, _I32, _I32, _P, C.POINTER(Opts), _I32],
    "fair_batch_relerr_dev": [_P, _P, _I32, _I32, _P],
    "fair_sirt_batch": [_P, _I32, _P, _I32, _I32, _P, C.POINTER(Opts), _I32, _P],
    "fair_autotune_batch": [_P, _I32, _P],
    "fair_autotune_spmv": [_P, _P],
    "fair_normest": [_P, _P, _D, _I32, _P],
    "fair_gradient_descent": [_P, _P, _I32, _P, _D, _D, _D, C.POINTER(Opts), _P, _P],
    "fair_art": [_P, _I32, _P, _I32, _P, _P, _P, _I32, C.POINTER(Opts), _P, _I64, _P],
    "fair_art_oracle": [_P, _I32, _P, _I32, _P, _P, _P, _P, _I32, C.POINTER(Opts), _P, _I64, _P,
                        C.POINTER(_I32), C.POINTER(_I64)],
    "fair_art_uniform_count": [_P, _I32, _I32, _P, _P, _I32, C.POINTER(Opts), C.POINTER(_I64)],
    "fair_art_steps": [_P, _P, _I64, _P, _P, C.POINTER(Opts), _P],
    "fair_art_steps_dev": [_P, _P, _I64, _P, _P, C.POINTER(Opts), _P],
    "fair_level_schedule": [_I32, _I32, _P, _P, _P, _I64, _P, _P, C.POINTER(C.c_int32)],
    "fair_twin_art": [_P, _I32, _I32, _P, _I32, _P, _P, _P, _P, _I32, _I32, C.POINTER(Opts), _P, _I64, _P,
                      C.POINTER(_I32), _P, _P, _P, C.POINTER(_I64)],
    "fair_twin_art_uniform_bound": [_P, _I32, _I32, _P, _P, _I32, C.POINTER(Opts), C.POINTER(_I64)],
    "fair_group_create": [_I32, _P, _I32, C.POINTER(_P)],
    "fair_group_free": [_P],
    "fair_group_sirt_dev": [_P, _P, _I32, _P, C.POINTER(Opts)],
    "fair_group_autotune": [_P, _P],
    "fair_mf_create": [C.POINTER(Geom), _P, C.POINTER(_P)],
    "fair_mf_free": [_P],
    "fair_mf_forward": [_P, _P, _P],
    "fair_mf_adjoint": [_P, _P, _P],
    "fair_mf_forward_dev": [_P, _P, _P],
    "fair_mf_adjoint_dev": [_P, _P, _P],
    "fair_mf_segments": [_P, C.POINTER(_I64)],
    "fair_mf_residual_dev": [_P, _P, _P, _P],
    "fair_mf_normest": [_P, _P, _D, _I32, _P],
    "fair_mf_gd_dev": [_P, _P, _I32, _D, _D, _D, C.POINTER(Opts), _P, _P],
    "fair_mf_gd": [_P, _P, _I32, _P, _D, _D, _D, C.POINTER(Opts), _P, _P],
}
_RESTYPES = {"fair_last_error": C.c_char_p, "fair_version": C.c_char_p}

_lib = None
_lock = threading.Lock()
_ready = False


def load(path: str | None = None) -> C.CDLL:
    """Load the shared library (no device initialisation).  FAIR_HIP_LIB names another build of it
    (the tests' sanitizer build of the host code)."""
    global _lib
    path = path or os.environ.get("FAIR_HIP_LIB") or LIB_PATH
    with _lock:
        if _lib is None:
            if not os.path.exists(path):
                raise FairError(FAIR_ERR_HIP, f"{path} is not built (run __graft_entry__.build())")
            # torch ships its own HIP runtime under a file name that does not match the soname
            # libamdhip64.so.7.  Loaded first, it satisfies this library's dependency and the
            # process has one runtime; loaded after /opt/rocm's, a second HSA runtime starts
            # and finds no device.  So torch, when present, is imported before the library.
            if not os.environ.get("FAIR_NO_TORCH"):
                try:
                    import torch  # noqa: F401
                except ImportError:
                    pass
            lib = C.CDLL(path)
            for name, args in SIGNATURES.items():
                f = getattr(lib, name)
                f.argtypes = args
                f.restype = _RESTYPES.get(name, C.c_int)
            _lib = lib
    return _lib


def lib() -> C.CDLL:
    """The library, with device 0 initialised on first use."""
    global _ready
    L = load()
    if not _ready:
        check(L.fair_init(int(os.environ.get("FAIR_DEVICE", "0"))))
        _ready = True
    return L


def check(status: int) -> None:
    if status != FAIR_OK:
        raise FairError(status, load().fair_last_error().decode())

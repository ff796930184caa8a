"""ctypes binding of libbia_hip.so (the gfx950 HIP library).

There is no fallback: if the library has not been built (``python -c "import
__graft_entry__ as g; g.build()"`` or ``make -C bia_3d_gpu_amd/csrc``) importing a
function that needs it raises ``RuntimeError``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbia_hip.so")

BIA_OK, BIA_ERR_ARG, BIA_ERR_HIP, BIA_ERR_OOM, BIA_ERR_STATE, BIA_ERR_RKF, BIA_ERR_IO = 0, 1, 2, 3, 4, 5, 6
METHOD_RKF, METHOD_ALG = 0, 1
MODE_PLAIN, MODE_LEARN, MODE_GUESS, MODE_PLAIN_LEARN = 0, 1, 2, 3
STATS_LEN = 16
COUPLING_SCALAR, COUPLING_GRAVITY, COUPLING_COULOMB = 0, 1, 2

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_ip = C.POINTER(C.c_int)
_dptr = C.POINTER(C.c_double)
_vp = C.c_void_p

# name -> (restype, argtypes); every exported symbol declared in include/bia_hip.h
SIGNATURES = {
    "find_dxyuv_kernel_": (None, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                  _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "invariance_gpu_": (None, [C.c_int, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp] + [_dptr] * 11),
    "rkf_c_": (None, [C.c_int, C.c_double, _dptr, _dptr, _dp, _ip, _dptr, _dptr, _dp, _dp, _ip, _ip, _dptr,
                      _dp, _dp, _dp, _dp]),
    "sincos_": (None, [_dptr, _dptr, _dptr]),
    "bia_version": (C.c_int, []),
    "bia_error_string": (C.c_char_p, [C.c_int]),
    "bia_last_error": (C.c_char_p, [_vp]),
    "bia_create": (C.c_int, [C.POINTER(_vp), C.c_int]),
    "bia_destroy": (C.c_int, [_vp]),
    "bia_set_stream": (C.c_int, [_vp, _vp]),
    "bia_synchronize": (C.c_int, [_vp]),
    "bia_set_profiling": (C.c_int, [_vp, C.c_int]),
    "bia_last_kernel_ms": (C.c_int, [_vp, C.POINTER(C.c_float)]),
    "bia_kernel_ms": (C.c_int, [_vp, C.POINTER(C.c_float), C.c_int, _ip]),
    "bia_reduce_kernel_ms": (C.c_int, [_vp, C.POINTER(C.c_float), C.c_int, _ip]),
    "bia_step_dev": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _vp, C.c_int,
                               _vp, _vp]),
    "bia_step_noalloc_dev": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _vp,
                                       C.c_int, _vp]),
    "bia_step_host": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _vp]),
    "bia_step_batched_dev": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _ip32, _dp, C.c_double, C.c_double,
                                       C.c_int, _vp, C.c_int, _vp, _vp, _vp, _vp]),
    "bia_autotune_dev": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _ip32, _dp, C.c_double, C.c_double,
                                   C.c_int, _vp, C.c_int, _vp, _vp, _vp, C.POINTER(C.c_float), _ip]),
    "bia_run_steps_dev": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _ip32, _dp, C.c_double, C.c_double,
                                    C.c_int, _vp, C.c_int, _vp, _vp, _vp, C.c_int, C.c_int, _vp]),
    "bia_set_workspace_limit": (C.c_int, [_vp, C.c_ulonglong]),
    "bia_plan_info": (C.c_int, [_vp, _ip, C.POINTER(C.c_ulonglong)]),
    "bia_set_rkf_tolerance": (C.c_int, [_vp, C.c_double, C.c_double]),
    "bia_set_failure_log": (C.c_int, [_vp, _vp, C.c_int, _vp]),
    "bia_invariants_dev": (C.c_int, [_vp, C.c_int, C.c_double, _vp, C.c_int, _vp]),
    "bia_invariants_host": (C.c_int, [_vp, C.c_int, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "bia_invariants_batched_dev": (C.c_int, [_vp, C.c_int, _ip32, _dp, _vp, C.c_int, _vp]),
    "bia_invariants_coupled_batched_dev": (C.c_int, [_vp, C.c_int, C.c_double, C.c_int, _ip32, _vp, C.c_int, _vp,
                                                     _vp, _vp]),
    "bia_mlp_nparam": (C.c_int, []),
    "bia_mlp_select": (C.c_int, [_vp, C.c_int]),
    "bia_mlp_init": (C.c_int, [_vp, C.c_ulonglong]),
    "bia_mlp_set_weights": (C.c_int, [_vp, _fp, C.c_int]),
    "bia_mlp_get_weights": (C.c_int, [_vp, _fp, C.c_int]),
    "bia_mlp_forward_dev": (C.c_int, [_vp, _vp, C.c_int, _vp]),
    "bia_mlp_grad_dev": (C.c_int, [_vp, _vp, C.c_int, _vp]),
    "bia_mlp_train_dev": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_ulonglong, _vp]),
    "bia_mlp_build_table": (C.c_int, [_vp, C.c_int]),
    "bia_record_samples_dev": (C.c_int, [_vp, C.c_int, C.c_double, C.c_double, _vp, C.c_int, C.c_int,
                                         C.c_ulonglong, _vp]),
    "bia_dim_rkf45_host": (C.c_int, [_vp, C.c_int, C.c_double, _dptr, C.c_double, _dp, _ip, C.c_double,
                                     C.c_double, _dptr, _vp, _ip, _ip, _ip, _ip]),
    "bia_dim_rkf45_coupled_host": (C.c_int, [_vp, C.c_int, C.c_double, C.c_int, _vp, _vp, _dptr, C.c_double, _dp,
                                             _ip, C.c_double, C.c_double, _dptr, _vp, _ip, _ip, _ip, _ip]),
    "bia_records_scan": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "bia_record_read": (C.c_int, [C.c_char_p, C.c_int, C.c_longlong, _vp, _vp, C.c_longlong,
                                  C.POINTER(C.c_longlong)]),
    "bia_record_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int, _vp, _vp, C.c_longlong, C.c_longlong]),
}

_lib = None


def load():
    """Load libbia_hip.so (raises RuntimeError when it is missing or incomplete)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `make -C bia_3d_gpu_amd/csrc` (hipcc, gfx950)")
    # PyTorch-ROCm ships its own HIP runtime under the same SONAME (libamdhip64.so.7).  Loading
    # torch first makes this library bind to that runtime, so a process holds one HIP runtime
    # whichever of the two touches the GPU first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)  # AttributeError here means an export is missing
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def declared_symbols() -> list[str]:
    return list(SIGNATURES)


class BiaError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str = ""):
        msg = load().bia_error_string(status).decode()
        super().__init__(f"{where}: {msg}" + (f" ({detail})" if detail else ""))
        self.status = status


def check(status: int, where: str, ctx=None):
    if status != BIA_OK:
        detail = load().bia_last_error(ctx).decode() if ctx else ""
        raise BiaError(status, where, detail)

"""ctypes binding of the C ABI declared in ``include/pfc_hip.h``.

The shared library ``openpfc_amd/libopenpfc_hip.so`` is built in-tree by
``openpfc_amd.build`` (hipcc, gfx950).  Loading fails loudly when it is
missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libopenpfc_hip.so")

PFC_HIP_OK = 0
PFC_HIP_EINVAL = 1
PFC_HIP_EUNSUPPORTED = 2
PFC_HIP_EDEVICE = 3
PFC_HIP_ENOMEM = 4
PFC_HIP_ESTATE = 5
PFC_HIP_EIO = 6
IO_REAL = 0
IO_COMPLEX = 1

F64 = 0
F32 = 1
# plan flag: FP64 butterflies on FP32 fields (pfc_hip.h PFC_HIP_FLAG_F64_BUTTERFLIES)
FLAG_F64_BUTTERFLIES = 1
MODEL_PFC = 0
MODEL_TUNGSTEN = 1
TUNGSTEN_NPARAMS = 17

# Every symbol include/pfc_hip.h declares, with its ctypes signature.
_vp = ctypes.c_void_p
_sz = ctypes.POINTER(ctypes.c_size_t)
_ip3 = ctypes.POINTER(ctypes.c_int)
SIGNATURES = {
    "pfc_hip_abi_version": (ctypes.c_int, []),
    "pfc_hip_last_error": (ctypes.c_char_p, []),
    "pfc_hip_plan_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_void_p)]),
    "pfc_hip_plan_destroy": (ctypes.c_int, [_vp]),
    "pfc_hip_sizes": (ctypes.c_int, [_vp, _sz, _sz, _sz]),
    "pfc_hip_allocated_bytes": (ctypes.c_int, [_vp, _sz]),
    "pfc_hip_r2c": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "pfc_hip_c2r": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "pfc_hip_pfc_set_operators": (ctypes.c_int, [_vp, ctypes.c_int,
                                                 ctypes.POINTER(ctypes.c_double),
                                                 ctypes.c_int, ctypes.c_double]),
    "pfc_hip_pfc_step": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp]),
    "pfc_hip_pfc_step_timed": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp,
                                              ctypes.POINTER(ctypes.c_double),
                                              ctypes.c_int]),
    "pfc_hip_ic_constant": (ctypes.c_int, [_vp, _vp, ctypes.c_double, _vp]),
    "pfc_hip_ic_single_seed": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(ctypes.c_double),
                                              ctypes.POINTER(ctypes.c_double),
                                              ctypes.c_double, ctypes.c_double,
                                              ctypes.c_double, ctypes.c_int, _vp]),
    "pfc_hip_ic_add_noise": (ctypes.c_int, [_vp, _vp, ctypes.c_ulonglong, ctypes.c_double,
                                            ctypes.c_double, _vp]),
    "pfc_hip_pfc_operators": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "pfc_hip_tungsten_operators": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "pfc_hip_field_write": (ctypes.c_int, [ctypes.c_char_p, _vp, ctypes.c_int, ctypes.c_int,
                                           _ip3, _ip3, _ip3, ctypes.c_int, _vp]),
    "pfc_hip_field_read": (ctypes.c_int, [ctypes.c_char_p, _vp, ctypes.c_int, ctypes.c_int,
                                          _ip3, _ip3, _ip3, _vp]),
    "pfc_hip_field_write_host": (ctypes.c_int, [ctypes.c_char_p, _vp, ctypes.c_int,
                                                ctypes.c_int, _ip3, _ip3, _ip3,
                                                ctypes.c_int]),
    "pfc_hip_field_read_host": (ctypes.c_int, [ctypes.c_char_p, _vp, ctypes.c_int,
                                               ctypes.c_int, _ip3, _ip3, _ip3]),
    "pfc_hip_tungsten_step": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, _vp]),
    "pfc_hip_pfc_adjoint_step": (ctypes.c_int, [_vp, _vp, _vp, ctypes.POINTER(ctypes.c_double),
                                                _vp]),
    "pfc_hip_bc_fixed": (ctypes.c_int, [_vp, _vp, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, _vp]),
    "pfc_hip_pfc_set_mean": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double]),
    "pfc_hip_variant_count": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "pfc_hip_variant_name": (ctypes.c_char_p, [_vp, ctypes.c_int, ctypes.c_int]),
    "pfc_hip_variant_get": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "pfc_hip_variant_set": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "pfc_hip_pfc_autotune": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp,
                                            ctypes.POINTER(ctypes.c_double), ctypes.c_int]),
    "pfc_hip_plan_create_batched": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                   ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                   ctypes.POINTER(ctypes.c_void_p)]),
    "pfc_hip_pfc_set_members": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double),
                                               ctypes.POINTER(ctypes.c_double), ctypes.c_int]),
    "pfc_hip_ensemble_layout": (ctypes.c_int, [ctypes.c_int, _ip3,
                                               ctypes.POINTER(ctypes.c_double), _sz, _sz]),
    "pfc_hip_ensemble_create": (ctypes.c_int, [ctypes.c_int, _ip3,
                                               ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                               ctypes.c_int, _vp,
                                               ctypes.POINTER(ctypes.c_void_p)]),
    "pfc_hip_ensemble_destroy": (ctypes.c_int, [_vp]),
    "pfc_hip_ensemble_groups": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "pfc_hip_ensemble_group": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p),
                                              ctypes.POINTER(ctypes.c_int), _sz]),
    "pfc_hip_ensemble_autotune": (ctypes.c_int, [_vp, ctypes.c_int, _vp]),
    "pfc_hip_ensemble_step": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp]),
    "pfc_hip_pfc_free_energy": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(ctypes.c_double), _vp]),
    "pfc_hip_stream_copy": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t, _vp]),
    "pfc_hip_ic_seeds": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                        ctypes.c_int, _vp]),
    "pfc_hip_column_max": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(ctypes.c_double), _vp]),
    "pfc_hip_column_max_noalloc": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "pfc_hip_bc_moving": (ctypes.c_int, [_vp, _vp] + [ctypes.c_double] * 7 + [_vp]),
}

_lib = None


class PfcHipError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[pfc_hip error {code}] {msg}")
        self.code = code


def load(path: str = LIB_PATH):
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} not found: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            " (there is no CPU fallback)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != PFC_HIP_OK:
        msg = load().pfc_hip_last_error().decode(errors="replace")
        raise PfcHipError(rc, msg)

"""The C-ABI library loads without a GPU and exports exactly what the header
declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pfc_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pfc_hip_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ["pfc_hip_plan_create", "pfc_hip_plan_destroy", "pfc_hip_sizes",
                     "pfc_hip_r2c", "pfc_hip_c2r", "pfc_hip_pfc_set_operators",
                     "pfc_hip_pfc_step", "pfc_hip_last_error"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    import openpfc_amd._lib as L
    if not os.path.exists(L.LIB_PATH):
        pytest.fail(f"{L.LIB_PATH} missing: run __graft_entry__.build()")
    lib = ctypes.CDLL(L.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    import openpfc_amd._lib as L
    assert sorted(L.SIGNATURES) == declared_functions()


def test_abi_version_and_error_string_without_device():
    import openpfc_amd as P
    lib = P.load()
    assert lib.pfc_hip_abi_version() == 1
    # argument checks run before any device call
    assert lib.pfc_hip_plan_create(0, 4, 4, 0, 0, ctypes.byref(ctypes.c_void_p())) == 1
    assert b"extents" in lib.pfc_hip_last_error()
    assert lib.pfc_hip_plan_create(6, 4, 4, 0, 0, ctypes.byref(ctypes.c_void_p())) == 2
    assert b"unsupported" in lib.pfc_hip_last_error()
    assert lib.pfc_hip_plan_create(8, 8, 8, 7, 0, ctypes.byref(ctypes.c_void_p())) == 1
    assert lib.pfc_hip_sizes(None, None, None, None) == 1
    assert lib.pfc_hip_plan_destroy(None) == 0


@pytest.mark.parametrize("src,dst,nbytes,why", [
    (0, 0x20000, 16, "null"),
    (0x10000, 0x20000, 24, "16-byte"),
    (0x10008, 0x20000, 16, "16-byte"),
    (0x10000, 0x10010, 64, "overlap"),
])
def test_stream_copy_rejects_bad_arguments_before_any_device_call(src, dst, nbytes, why):
    """pfc_hip_stream_copy checks its arguments on the host: these calls
    return PFC_HIP_EINVAL without touching a device (none is needed here)."""
    from openpfc_amd._lib import load
    lib = load()
    rc = lib.pfc_hip_stream_copy(ctypes.c_void_p(src), ctypes.c_void_p(dst), nbytes, None)
    assert rc == 1
    assert why in lib.pfc_hip_last_error().decode()

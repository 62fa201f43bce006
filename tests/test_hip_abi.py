"""The thin HIP C ABI (include/ptk_hip.h) loads and exports every declared entry point.

CPU-only: no compute call is made.  With no device present the library must still load, report
zero devices without error and give readable status strings.
"""
import os

import pytest

from ptask_amd import hip as H
from ptask_amd import paths


def test_every_declared_symbol_is_exported_and_bound():
    declared = H.declared_symbols()
    assert len(declared) >= 35
    L = H.lib()
    for name in declared:
        assert hasattr(L, name), name
    # the Python binding covers exactly the header
    assert set(declared) == set(H.SIGNATURES)


def test_status_strings_and_zero_devices_are_not_errors():
    L = H.lib()
    assert L.ptk_hip_status_string(0) == b"ok"
    assert b"not found" in L.ptk_hip_status_string(5)
    assert H.device_count() >= 0


def test_invalid_arguments_are_rejected_without_touching_hip():
    import ctypes as C
    L = H.lib()
    assert L.ptk_hip_malloc(None, 16) == 2
    assert L.ptk_hip_launch(None, 1, 1, 1, 64, 1, 1, 0, None, None, 0) == 2
    assert b"invalid" in L.ptk_hip_last_error_string()
    # oversize block is rejected before reaching the runtime
    assert L.ptk_hip_launch(C.c_void_p(1), 1, 1, 1, 2048, 1, 1, 0, None, None, 0) == 2


def test_missing_code_object_reports_not_found():
    import ctypes as C
    m = C.c_void_p()
    rc = H.lib().ptk_hip_module_load(C.byref(m), b"/nonexistent/fdtd.hsaco")
    assert rc == 5


@pytest.mark.parametrize("name", ["fdtd.hsaco"])
def test_code_objects_are_built_for_gfx950(name):
    path = paths.kernel_path(name)
    assert os.path.exists(path), path
    data = open(path, "rb").read()
    assert b"gfx950" in data
    for k in (b"HxComputation", b"HyComputation", b"HzComputation", b"ExComputation",
              b"EyComputation", b"EzComputation"):
        assert k in data

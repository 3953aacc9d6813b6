"""C-ABI checks that need no GPU: the library loads, exports every symbol
declared in include/gputils_hip.h, and rejects invalid arguments with the
LAPACK-style negative index before touching the device."""
import ctypes
import re
import subprocess

import pytest

from gputils_amd import _lib


def test_library_loads_and_exports_header():
    L = _lib.lib()
    decls = _lib.parse_header()
    assert len(decls) >= 39
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (gputils_\w+)", out))
    missing = sorted(set(decls) - exported)
    assert not missing, f"declared but not exported: {missing}"
    for name in decls:
        assert hasattr(L, name)


def test_version_and_error_string():
    L = _lib.lib()
    assert b"gfx950" in L.gputils_version()
    assert isinstance(L.gputils_last_error(), bytes)


def test_device_count_without_gpu_is_not_an_error():
    c = ctypes.c_int(-1)
    assert _lib.lib().gputils_device_count(ctypes.byref(c)) == 0
    assert c.value >= 0


def _msg():
    return _lib.lib().gputils_last_error().decode()


@pytest.mark.parametrize("fn,args,code", [
    ("gputils_dpotrf_batched", (-1, None, 1, 1, None, 1, None), -1),
    ("gputils_dpotrf_batched", (4, None, 3, 16, None, 1, None), -3),
    ("gputils_spotrf_batched", (4, None, 4, 16, None, -2, None), -6),
    ("gputils_dpotrf_batched", (4, None, 4, 16, None, 1, None), -2),   # A NULL
    ("gputils_dpotrs_batched", (4, -1, None, 4, 16, None, 4, 4, None, 1, None), -2),
    ("gputils_dpotrs_batched", (4, 1, None, 4, 16, None, 3, 4, None, 1, None), -7),
    ("gputils_dposv_batched", (-1, 1, None, 1, 0, None, 1, 0, None, 1, None), -1),
    ("gputils_dposv_batched", (129, 1, None, 129, 0, None, 129, 0, None, 1, None), -3),  # any n; A NULL
    ("gputils_dposv_mixed_batched", (8, 1, None, 8, 64, None, 8, 8, None, 8, 8, 0.0, 3, None, None, None, 1, None), -12),
    ("gputils_dposv_mixed_batched", (8, 1, None, 8, 64, None, 8, 8, None, 8, 8, 1e-12, -1, None, None, None, 1, None), -13),
    ("gputils_sgemm_strided_batched", (2, 2, 2, 1.0, None, 1, 4, None, 2, 4, 0.0, None, 2, 4, 1, 0, None), -6),
    ("gputils_sgemm_strided_batched", (2, 2, 2, 1.0, None, 2, 4, None, 2, 4, 0.0, None, 2, 4, 1, 7, None), -16),
    ("gputils_dgemm_strided_batched", (2, 2, 2, 1.0, None, 2, 4, None, 2, 4, 0.0, None, 1, 4, 1, None), -13),
    ("gputils_dnrm2", (-5, None, None, None, None), -1),
    ("gputils_stranspose_batched", (3, 2, None, 2, 6, None, 2, 6, 1, None), -4),
    ("gputils_gels_ragged_plan_configure", (None, 0, 0, 4, 1, 0), -1),
    ("gputils_drot", (-1, None, 1, None, 1, 1.0, 0.0, None), -1),
    ("gputils_gels_ragged_plan_set_concurrency", (None, 1), -1),
    ("gputils_gels_ragged_plan_get_concurrency", (None, None), -1),
    ("gputils_srot", (4, None, 0, None, 1, 1.0, 0.0, None), -3),
    ("gputils_dgivens_annihilate", (4, 3, None, 4, 0, 4, 0, None), -6),
    ("gputils_sgivens_annihilate", (4, 3, None, 4, 0, 1, 3, None), -7),
    ("gputils_gels_ragged_plan_describe", (None, None, 0, None), -1),
    ("gputils_gels_ragged_plan_tune", (None, None, None, 1, None), -1),
    ("gputils_snullspace_project_batched", (0, None, 1, 0, None, None, 0, 1, None), -1),
    ("gputils_dnullspace_project_batched", (8, None, 7, 64, None, None, 8, 1, None), -3),
    ("gputils_snullspace_project_batched", (8, None, 8, 64, None, None, 8, 1, None), -2),  # N NULL
])
def test_invalid_arguments_are_rejected(fn, args, code):
    assert getattr(_lib.lib(), fn)(*args) == code
    assert f"invalid argument {-code}" in _msg()


def test_empty_problems_are_noops():
    L = _lib.lib()
    assert L.gputils_dpotrf_batched(0, None, 1, 0, None, 5, None) == 0
    assert L.gputils_dpotrf_batched(4, None, 4, 16, None, 0, None) == 0
    assert L.gputils_sgemm_strided_batched(0, 3, 3, 1.0, None, 1, 0, None, 3, 9, 0.0, None, 1, 0, 4, 0, None) == 0
    assert L.gputils_dposv_mixed_batched(8, 0, None, 8, 64, None, 8, 0, None, 8, 0, 1e-12, 3, None, None, None,
                                          9, None) == 0
    assert L.gputils_dposv_mixed_workspace_bytes(128, 10) == 0
    assert L.gputils_dposv_mixed_workspace_bytes(129, 10) == 129 * 129 * 8 * 10
    assert L.gputils_dposv_mixed_workspace_bytes(-1, 10) == -1


def test_empty_ragged_plan_shape_calls():
    """A plan of zero systems has no buckets: describe reports none, tuning
    is a no-op and configuring a bucket is out of range (no device memory
    is touched, so this runs without a GPU)."""
    import ctypes
    L = _lib.lib()
    plan = ctypes.c_void_p()
    assert L.gputils_gels_ragged_plan_create(ctypes.byref(plan), 0, None, None, None, None) == 0
    try:
        nb = ctypes.c_int(-1)
        assert L.gputils_gels_ragged_plan_describe(plan, None, 0, ctypes.byref(nb)) == 0
        assert nb.value == 0
        assert L.gputils_gels_ragged_plan_tune(plan, None, None, 1, None) == 0
        assert L.gputils_gels_ragged_plan_configure(plan, 0, 0, 8, 1, 0) == -2
        assert "invalid argument 2" in _msg()
        assert L.gputils_gels_ragged_plan_configure(plan, 0, 5, 3, 1, 7) == -2  # the bucket is checked first
    finally:
        assert L.gputils_gels_ragged_plan_destroy(plan) == 0


def test_no_oracle_in_the_product():
    """The shipped library must not depend on the host oracle."""
    out = subprocess.run(["ldd", _lib.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in out
    syms = subprocess.run(["nm", "-D", _lib.LIB_PATH], capture_output=True, text=True).stdout
    assert "orc_" not in syms

"""The C-ABI library loads and exports every symbol include/ses3d_hip.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "ses3d_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ses3d_hip_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_header_symbols():
    from ses3d_amd.solver import LIB_PATH
    if not os.path.exists(LIB_PATH):
        from ses3d_amd import build
        build.build()
    lib = ctypes.CDLL(LIB_PATH)
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_solver_binding_covers_header():
    from ses3d_amd import solver
    lib = solver.load_library()
    for n in _declared():
        assert getattr(lib, n).argtypes is not None or n == "ses3d_hip_last_error", n


def test_errors_without_device_are_reported():
    """Creating a context with invalid parameters fails with a message, never a crash."""
    from ses3d_amd import solver
    lib = solver.load_library()
    cfg = solver.Config(3, 3, 3, 5, 2, 0, 0, 1, 10, 0.1, 0.7, 0.9, 0.2, 0.4, 5.4e6, 6.371e6, 6.371e6,
                        (ctypes.c_float * 3)(1, 1, 1), (ctypes.c_float * 3)(0, 0, 0))
    h = ctypes.c_void_p()
    rc = lib.ses3d_hip_create(ctypes.byref(cfg), 1, ctypes.byref(h))
    assert rc != 0
    assert b"lpd" in lib.ses3d_hip_last_error()


def test_missing_library_fails_loudly(tmp_path):
    from ses3d_amd import solver
    with pytest.raises(solver.Ses3dError):
        solver._lib_saved = solver._lib
        solver._lib = None
        try:
            solver.load_library(str(tmp_path / "nope.so"))
        finally:
            solver._lib = solver._lib_saved


def test_group_and_schedule_entry_points_reject_null():
    """The group / schedule entry points validate their handles before any
    device call (no GPU needed)."""
    from ses3d_amd import solver
    lib = solver.load_library()
    h = ctypes.c_void_p()
    assert lib.ses3d_hip_group_create(None, 1, ctypes.byref(h)) != 0
    assert b"null" in lib.ses3d_hip_last_error()
    assert lib.ses3d_hip_group_run(None, 1, 1) != 0
    assert lib.ses3d_hip_group_autotune(None, 1, 6) != 0
    sched = (ctypes.c_int32 * 4)(2, 256, 1, 1)
    assert lib.ses3d_hip_set_schedule(None, sched) != 0
    assert lib.ses3d_hip_autotune(None, 1, 6) != 0
    assert lib.ses3d_hip_set_streaming(None, 1) != 0
    assert lib.ses3d_hip_get_streaming(None, None, None) != 0
    assert lib.ses3d_hip_group_set_streaming(None, 0) != 0
    assert lib.ses3d_hip_group_get_streaming(None, None, None) != 0
    assert lib.ses3d_hip_set_kernel_precision(None, 32) != 0
    assert b"null" in lib.ses3d_hip_last_error()
    assert lib.ses3d_hip_set_checkpoint_keep(None, 2) != 0
    assert lib.ses3d_hip_prepare_checkpoints(None, 100) != 0
    assert b"null" in lib.ses3d_hip_last_error()
    lib.ses3d_hip_group_destroy(None)

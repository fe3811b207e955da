"""The C-ABI library loads without a GPU and exports every entry point include/lvhip.h declares."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "lvhip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(lv_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "lv_create" in names and "lv_remesh" in names and "lv_step_fused" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol(lv):
    L = ctypes.CDLL(lv.library_path)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_package_binds_every_declared_symbol(lv):
    bound = set(lv._SIGNATURES) | {"lv_last_error"}
    assert set(declared_functions()) <= bound


def test_null_context_is_rejected_without_gpu_work(lv):
    L = lv.lib()
    assert L.lv_remesh(None) == lv.LV_E_INVALID
    assert L.lv_create(None, None) == lv.LV_E_INVALID

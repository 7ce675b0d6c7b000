"""CPU checks of the drop-in boundary: the built library loads, exports every entry point that
include/arrowmetal.h declares, and answers the NULL-handle conventions of the reference's C ABI
(Sources/ArrowMetalC/ArrowMetalC.swift:9-34) without touching a GPU."""
import ctypes
import os
import re

import pytest

import arrowmetal_amd as am

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "arrowmetal.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-zA-Z_][\w\s\*]*?\b(am_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_path():
    names = declared_functions()
    for n in ["am_argsort", "am_argsort_ex", "am_argsort_ex2", "am_take", "am_group_by", "am_group_by_keys",
              "am_group_by_group_count", "am_group_by_keys_result", "am_group_by_ids", "am_group_by_release",
              "am_group_agg_ex", "am_import", "am_import_device", "am_export", "am_export_device", "am_release",
              "am_batch_begin", "am_batch_end", "am_last_error", "am_version"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(am.library_path())
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_declares_every_symbol():
    assert set(declared_functions()) <= set(am._SIGNATURES)


def test_null_handles_return_2_without_a_gpu():
    L = am.lib()
    out = ctypes.c_void_p()
    assert L.am_argsort(None, 0, ctypes.byref(out)) == 2
    assert L.am_argsort_ex2(None, 0, 0, 0, ctypes.byref(out)) == 2
    assert L.am_take(None, None, ctypes.byref(out)) == 2
    assert L.am_group_by(None, 4, 0, None, ctypes.byref(out)) == 2
    assert L.am_group_by_keys(None, 1, ctypes.byref(out)) == 2
    assert L.am_group_by_group_count(None) == -1
    assert L.am_group_agg_ex(None, None, 0, 0.0, ctypes.byref(out)) == 2
    assert L.am_length(None) == -1
    L.am_release(None)
    L.am_group_by_release(None)
    assert L.am_version().decode().startswith("0.")


def test_library_is_built_for_gfx950():
    data = open(am.library_path(), "rb").read()
    assert b"gfx950" in data

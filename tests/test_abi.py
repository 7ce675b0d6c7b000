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


def test_plan_null_handles_and_parse_errors_without_a_gpu():
    """am_plan_*: NULL arguments return 2 (or -1 / NULL) with no message (ArrowMetalC_Plan.swift:52-160);
    a plan that does not parse is refused before any device work, with PlanJSON.swift's texts."""
    L = am.lib()
    out = ctypes.c_void_p()
    assert L.am_plan_run(None, None, 0, 1, ctypes.byref(out)) == 2
    assert L.am_plan_run(b"{}", None, 0, 1, None) == 2
    assert L.am_plan_run(b"{}", None, -1, 1, ctypes.byref(out)) == 2
    assert L.am_plan_source_create(None, None, None, 0, ctypes.byref(out)) == 2
    assert L.am_plan_column_count(None) == -1
    assert L.am_plan_row_count(None) == -1
    assert L.am_plan_column_name(None, 0) is None
    assert L.am_plan_column(None, 0, ctypes.byref(out)) == 2
    assert L.am_plan_explain(None, None, 0, 1) is None
    L.am_plan_result_release(None)
    L.am_plan_source_release(None)
    cases = {
        "{": "Invalid ArrowArray: plan: not a JSON object",
        "[1, 2]": "Invalid ArrowArray: plan: not a JSON object",
        '{"x": 1}': 'Invalid ArrowArray: plan: node has no "op"',
        '{"op": "sort"}': 'Invalid ArrowArray: plan: sort has no "input"',
        '{"op": "scan", "source": "t"}': 'Invalid ArrowArray: plan: unknown source "t"',
        '{"op": "nope"}': 'Invalid ArrowArray: plan: unknown operator "nope"',
        '{"op": "join", "left": {}, "right": {}}': 'Unsupported Arrow type: plan: operator "join" is not supported '
                                                   'by this build',
        '{"op": "limit", "count": 2, "input": {"op": "scan", "source": "s\\u00e9"}}':
            'Invalid ArrowArray: plan: unknown source "sé"',
    }
    for text, msg in cases.items():
        assert L.am_plan_run(text.encode(), None, 0, 1, ctypes.byref(out)) == 1, text
        assert L.am_last_error().decode() == msg, text
        with pytest.raises(am.ArrowMetalError) as e:
            am.plan_explain(text)
        assert str(e.value) == msg


def test_plan_expressions_outside_the_subset_are_refused():
    """Only column references are expressions here; the s-expression forms of ExprText.swift:14-45
    ("(col "x")", "(col x)") both name a column. A table with no columns needs no device."""
    L = am.lib()
    src = ctypes.c_void_p()
    assert L.am_plan_source_create(b"t", (ctypes.c_void_p * 1)(), (ctypes.c_char_p * 1)(), 0, ctypes.byref(src)) == 0
    try:
        scan = {"op": "scan", "source": "t"}
        for expr in ['(add (col "a") (int 1))', '(cast (col "a") f64)', '(int 3)', '(col "a") extra', 'col a']:
            plan = {"op": "group_by", "keys": [["k", expr]], "aggs": [], "input": scan}
            with pytest.raises(am.ArrowMetalError) as e:
                am.plan_explain(plan, [type("S", (), {"_h": src})()])
            assert str(e.value) == (f"Unsupported Arrow type: plan: the expression {expr} is not supported by this "
                                    "build (only (col \"name\"))"), expr
        for expr in ['(col "a")', '(col a)', ' ( col  "a b" ) ']:
            plan = {"op": "group_by", "keys": [["k", expr]], "aggs": [["count", "n", ""]], "input": scan}
            text = am.plan_explain(plan, [type("S", (), {"_h": src})()])
            assert "HASH_AGGREGATE keys [k]" in text and "SOURCE t" in text
    finally:
        L.am_plan_source_release(src)


def test_import_chunks_refuses_without_a_gpu():
    """am_import_chunks: NULL arguments return 2; a type it does not take returns 3 with the
    reference's text and moves nothing (ArrowMetalC.swift:52-67); am_import_chunks_supported."""
    L = am.lib()
    sch = am._ArrowSchema()
    sch.format = b"u"
    out = ctypes.c_void_p()
    assert L.am_import_chunks(None, None, 0, ctypes.byref(out)) == 2
    assert L.am_import_chunks(ctypes.addressof(sch), None, 2, ctypes.byref(out)) == 2
    assert L.am_import_chunks(ctypes.addressof(sch), None, 0, ctypes.byref(out)) == 3
    assert L.am_last_error().decode() == ("Unsupported Arrow type: am_import_chunks does not take u arrays; "
                                          "import their concatenation")
    assert L.am_import_chunks_supported(ctypes.addressof(sch)) == 0
    assert L.am_import_chunks_supported(None) == 0
    for fmt in (b"l", b"g", b"i", b"C", b"tsu:UTC", b"tdD"):
        sch.format = fmt
        assert L.am_import_chunks_supported(ctypes.addressof(sch)) == 1, fmt

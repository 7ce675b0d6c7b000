"""GPU parity of the plan engine's subset (am_plan_source_create / am_plan_run / am_plan_explain) against
the oracle's plan interpreter, through the C ABI.

Cases follow what the reference's engine computes (Sources/ArrowMetal/Engine/Executor.swift:46-160,
243-350; PhysicalPlan.swift:118-126; PlanJSON.swift's error texts): ORDER BY of whole tables (the
one-key sort that moves every column with the MSD partition, and the lexsort + take form), sort under a
limit as a top-k, limit with offsets, group_by with every aggregate, keep-first unique, and the errors.
"""
import json
import zlib

import numpy as np
import pytest

import arrowmetal_amd as am
import oracle_py as O

pytestmark = pytest.mark.gpu


def assert_tables_equal(got, want, label=""):
    assert list(got) == list(want), (label, list(got), list(want))
    for name in want:
        wv, wva = want[name]
        gv, gva = got[name].to_numpy()
        assert gv.dtype == wv.dtype, (label, name, gv.dtype, wv.dtype)
        assert len(gv) == len(wv), (label, name)
        ok = np.ones(len(wv), bool) if wva is None else wva
        gok = np.ones(len(gv), bool) if gva is None else gva
        np.testing.assert_array_equal(gok, ok, err_msg=f"{label} {name} validity")
        w = wv.view(f"u{wv.dtype.itemsize}")
        g = gv.view(f"u{gv.dtype.itemsize}")
        np.testing.assert_array_equal(g[ok], w[ok], err_msg=f"{label} {name} values")


def run_both(plan, tables):
    sources = [am.PlanSource(name, {c: am.MetalArray.from_numpy(v, va) for c, (v, va) in t.items()})
               for name, t in tables.items()]
    got = am.plan_run(plan, sources)
    want = O.plan_run(plan, tables)
    return got, want


def profiled_names(fn):
    am.prof_reset()
    am.prof_enable(True)
    try:
        out = fn()
    finally:
        am.prof_enable(False)
    names = set(am.prof_report())
    am.prof_reset()
    return out, names


def order_by_table(rng, n, key_kind="i64", key_nulls=0.02):
    if key_kind == "i64":
        key = (O.splitmix64(n, n) >> np.uint64(1)).view(np.int64) - (1 << 62)
    elif key_kind == "f64":
        key = ((O.splitmix64(n + 1, n) >> np.uint64(11)).astype(np.float64) / 2.0**53 - 0.5) * 1e6
        sel = rng.random(n)
        key[sel < 0.001] = np.nan
        key[(sel >= 0.001) & (sel < 0.002)] = -0.0
        key[(sel >= 0.002) & (sel < 0.003)] = 0.0
        pay = sel > 0.9995
        key.view(np.uint64)[pay] = np.uint64(0xFFF8000000000000) | (O.splitmix64(9, int(pay.sum())) & np.uint64(0xFFFF))
    elif key_kind == "u64":
        key = O.splitmix64(n + 2, n)
    else:   # 3,000 distinct full-range keys: equal-key runs, copies and the LSD queue
        key = O.splitmix64(n + 3, 3000).view(np.int64)[rng.integers(0, 3000, n)]
    kv = rng.random(n) >= key_nulls
    t = {
        "key": (key, kv),
        "pay": (rng.standard_normal(n) * 1e6, rng.random(n) >= 0.05),
        "i32": (rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32), None),
        "u8": (rng.integers(0, 256, n).astype(np.uint8), rng.random(n) >= 0.5),
        "f32": (rng.standard_normal(n).astype(np.float32), None),
        "i16": (rng.integers(-30000, 30000, n).astype(np.int16), rng.random(n) >= 0.01),
    }
    return t


@pytest.mark.parametrize("key_kind,desc,nulls,forder", [
    ("i64", True, "last", "ieee"),
    ("i64", False, "first", "ieee"),
    ("f64", True, "last", "total"),
    ("f64", False, "first", "ieee"),
    ("f64", True, "first", "nan_largest"),
    ("u64", False, "last", "ieee"),
    ("runs", True, "last", "ieee"),
])
def test_order_by_moves_every_column(key_kind, desc, nulls, forder):
    """One 64-bit key over 300k rows: the columns move with the sort's top-digit partition (the
    order_by kernels run), and every column equals the oracle's lexsort + take."""
    rng = np.random.default_rng(zlib.crc32(f"{key_kind}{desc}{nulls}{forder}".encode()))
    n = 300_007
    t = order_by_table(rng, n, key_kind)
    plan = {"op": "sort", "by": [{"column": "key", "descending": desc, "nulls": nulls, "float_order": forder}],
            "input": {"op": "scan", "source": "t"}}
    (got, want), names = profiled_names(lambda: run_both(plan, {"t": t}))
    assert "order_by_gather" in names and "order_by_partition" in names, sorted(names)
    assert_tables_equal(got, want, f"{key_kind} desc={desc} nulls={nulls} {forder}")


def test_order_by_without_nulls_and_all_null_payload():
    rng = np.random.default_rng(3)
    n = 1 << 18
    key = (O.splitmix64(77, n) >> np.uint64(1)).view(np.int64)
    t = {"key": (key, None), "p": (np.arange(n, dtype=np.float64), np.zeros(n, bool)),
         "q": (rng.integers(0, 9, n).astype(np.int64), None)}
    plan = {"op": "sort", "by": [["key", True]], "input": {"op": "scan", "source": "t"}}
    (got, want), names = profiled_names(lambda: run_both(plan, {"t": t}))
    assert "order_by_gather" in names
    assert_tables_equal(got, want, "no key nulls")


@pytest.mark.parametrize("n", [0, 1, 1000, 100_003])
def test_small_and_multi_key_sorts_take_lexsort(n):
    rng = np.random.default_rng(n + 11)
    t = {"a": (rng.integers(0, 5, n).astype(np.int32), rng.random(n) > 0.1),
         "b": (rng.standard_normal(n), rng.random(n) > 0.1),
         "c": (rng.integers(-9, 9, n).astype(np.int64), None)}
    for by in ([["b", True]], [["a", False], ["b", True, {"nulls": "first", "float_order": "total"}]],
               [{"column": "c"}, {"column": "a", "descending": True}]):
        plan = {"op": "sort", "nulls": "last", "by": by, "input": {"op": "scan", "source": "t"}}
        got, want = run_both(plan, {"t": t})
        assert_tables_equal(got, want, json.dumps(by))


def test_multi_key_sort_of_a_large_table():
    rng = np.random.default_rng(12)
    n = 400_001
    t = {"g": (rng.integers(0, 50, n).astype(np.int64), None), "k": order_by_table(rng, n)["key"],
         "v": (rng.standard_normal(n), None)}
    plan = {"op": "sort", "by": [["g", False], ["k", True]], "input": {"op": "scan", "source": "t"}}
    got, want = run_both(plan, {"t": t})
    assert_tables_equal(got, want, "two keys")


@pytest.mark.parametrize("count,offset", [(10, None), (1024, 0), (1025, 0), (5, 3), (0, 0), (10**12, 2), (-4, 0),
                                          (7, -3), (3, 10**9)])
def test_limit_and_top_k(count, offset):
    rng = np.random.default_rng(13)
    n = 400_000
    t = {"k": (rng.standard_normal(n), rng.random(n) > 0.05), "v": (np.arange(n, dtype=np.int64), None)}
    for nulls in ("last", "first"):
        inner = {"op": "sort", "by": [["k", True, {"nulls": nulls}]], "input": {"op": "scan", "source": "t"}}
        plan = {"op": "limit", "count": count, "input": inner}
        if offset is not None:
            plan["offset"] = offset
        (got, want), names = profiled_names(lambda: run_both(plan, {"t": t}))
        if offset in (None, 0) and 0 < count <= 1024 and nulls == "last":
            # a top-k selection: no full sort of the 400k rows
            assert "topk_select" in names and "sort_msd_first" not in names, sorted(names)
        assert_tables_equal(got, want, f"limit {count} offset {offset} nulls {nulls}")
    # A limit straight over a scan is a slice.
    plan = {"op": "limit", "count": count, "input": {"op": "scan", "source": "t"}}
    if offset is not None:
        plan["offset"] = offset
    got, want = run_both(plan, {"t": t})
    assert_tables_equal(got, want, "slice")


@pytest.mark.parametrize("n,span", [(0, 10), (1, 10), (10_000, 50), (300_000, 100_000)])
def test_group_by_every_aggregate(n, span):
    rng = np.random.default_rng(n + span)
    t = {"k": (rng.integers(-span, span, n).astype(np.int64), rng.random(n) > 0.05),
         "k2": (rng.integers(0, 3, n).astype(np.int32), None),
         "v": (rng.integers(-10**12, 10**12, n).astype(np.int64), rng.random(n) > 0.1),
         "f": (rng.standard_normal(n) * 1e3, rng.random(n) > 0.1),
         "u": (rng.integers(0, 2**32, n, dtype=np.int64).astype(np.uint32), None),
         "i": (rng.integers(-1000, 1000, n).astype(np.int32), None)}
    aggs = [["sum", "s_v", "(col \"v\")"], ["sum", "s_f", "(col \"f\")"], ["sum", "s_u", "(col u)"],
            ["sum", "s_i", "(col \"i\")"], ["count", "n", ""], ["count", "c_v", "(col \"v\")"],
            ["mean", "m_v", "(col \"v\")"], ["mean", "m_f", "(col \"f\")"], ["min", "lo_f", "(col \"f\")"],
            ["max", "hi_v", "(col \"v\")"], ["min", "lo_u", "(col \"u\")"], ["max", "hi_i", "(col \"i\")"]]
    for keys in ([["key", "(col \"k\")"]], [["a", "(col \"k2\")"], ["b", "(col \"k\")"]]):
        plan = {"op": "group_by", "keys": keys, "aggs": aggs, "input": {"op": "scan", "source": "t"}}
        got, want = run_both(plan, {"t": t})
        assert_tables_equal(got, want, f"group_by {keys}")


def test_group_by_then_sort_then_limit():
    """The reference header's example shape: sum by key, biggest first, top 10."""
    rng = np.random.default_rng(21)
    n = 200_000
    t = {"region": (rng.integers(0, 500, n).astype(np.int64), None),
         "amount": (rng.integers(0, 10_000, n).astype(np.int64), None)}
    plan = {"op": "limit", "count": 10, "input":
            {"op": "sort", "by": [["total", True]], "input":
             {"op": "group_by", "keys": [["region", "(col \"region\")"]],
              "aggs": [["sum", "total", "(col \"amount\")"]], "input": {"op": "scan", "source": "sales"}}}}
    got, want = run_both(plan, {"sales": t})
    assert_tables_equal(got, want, "example")


@pytest.mark.parametrize("n", [0, 7, 100_000])
def test_unique_keeps_first_rows(n):
    rng = np.random.default_rng(n + 5)
    t = {"a": (rng.integers(0, 1000, n).astype(np.int64), rng.random(n) > 0.05),
         "b": (rng.integers(0, 30, n).astype(np.int64), None),
         "v": (rng.standard_normal(n), None)}
    for plan in ({"op": "unique", "subset": ["a", "b"], "input": {"op": "scan", "source": "t"}},
                 {"op": "unique", "input": {"op": "scan", "source": "t", "columns": ["b", "a"]}}):
        got, want = run_both(plan, {"t": t})
        assert_tables_equal(got, want, json.dumps(plan))


def test_plan_errors_and_explain():
    t = {"k": (np.arange(10, dtype=np.int64), None), "v": (np.arange(10, dtype=np.float64), None)}
    src = am.PlanSource("t", {c: am.MetalArray.from_numpy(v, va) for c, (v, va) in t.items()})
    scan = {"op": "scan", "source": "t"}

    def err(plan):
        with pytest.raises(am.ArrowMetalError) as e:
            am.plan_run(plan, [src])
        return str(e.value)

    assert err({"op": "sort", "by": [["nope", True]], "input": scan}) == "Invalid ArrowArray: no column named nope"
    assert err({"op": "scan", "source": "t", "columns": ["zz"]}) == "Invalid ArrowArray: no column named zz"
    assert err({"op": "unique", "subset": ["q"], "input": scan}) == "Invalid ArrowArray: unique: no column named q"
    assert err({"op": "scan", "source": "u"}) == 'Invalid ArrowArray: plan: unknown source "u"'
    assert err({"op": "group_by", "keys": [["k", "(col \"k\")"]], "aggs": [["median", "m", "(col \"v\")"]],
                "input": scan}) == "Invalid ArrowArray: plan: aggs entries are [op, name, sexpr?]"
    assert err({"op": "group_by", "keys": [["k", "(col \"k\")"]], "aggs": [["sum", "s"]], "input": scan}) == \
        "Invalid ArrowArray: sum needs an expression"
    assert err({"op": "filter", "predicate": "(gt (col \"v\") (int 1))", "input": scan}).startswith(
        "Unsupported Arrow type: plan: operator \"filter\"")
    assert err({"op": "sort", "by": [["k", True, {"nulls": "middle"}]], "input": scan}) == \
        "Invalid ArrowArray: plan: sort \"nulls\" must be \"first\" or \"last\""
    assert err("{not json") == "Invalid ArrowArray: plan: not a JSON object"
    with pytest.raises(am.ArrowMetalError, match="Array length mismatch: 10 vs 3"):
        am.PlanSource("bad", {"a": np.arange(10, dtype=np.int64), "b": np.arange(3, dtype=np.int64)})
    text = am.plan_explain({"op": "limit", "count": 3, "input": {"op": "sort", "by": [["v", True]], "input": scan}}, [src])
    assert "TOP_K 3" in text and "SOURCE t" in text

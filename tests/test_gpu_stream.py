"""GPU parity of the streaming group-by (am_stream_from_c_stream + am_stream_group_by) against the oracle
over the concatenated batches, through the C ABI, with enough batches and distinct keys that the
partial states are folded mid-stream. Integer aggregates are bit-exact; float sums are compared within
a relative 1e-12 (the reference adds its batches' float sums in float64 as well)."""
import numpy as np
import pytest

import arrowmetal_amd as am
import oracle_py as O

pytestmark = pytest.mark.gpu
pa = pytest.importorskip("pyarrow")


def reader(cols, valids, sizes):
    batches, at = [], 0
    for s in sizes:
        arrays = [pa.array(c[at:at + s], mask=None if v is None else ~v[at:at + s]) for c, v in zip(cols.values(), valids)]
        batches.append(pa.record_batch(arrays, names=list(cols)))
        at += s
    schema = pa.schema([(k, pa.from_numpy_dtype(v.dtype)) for k, v in cols.items()])
    return pa.RecordBatchReader.from_batches(schema, batches)


def bits(a, dtype):
    dtype = np.dtype(dtype)
    return a.view(dtype) if dtype.itemsize == 8 else (a & np.uint64((1 << 8 * dtype.itemsize) - 1)).astype(
        f"u{dtype.itemsize}").view(dtype)


def check_exact(got, want_bits, want_ok, dtype, label):
    gv, gva = got.to_numpy()
    gva = np.ones(len(gv), bool) if gva is None else gva
    np.testing.assert_array_equal(gva, want_ok, err_msg=label)
    w = bits(want_bits, dtype)
    np.testing.assert_array_equal(gv[want_ok].view(f"u{gv.dtype.itemsize}"), w[want_ok].view(f"u{w.dtype.itemsize}"),
                                  err_msg=label)


def test_stream_group_by_many_batches_with_folds():
    rng = np.random.default_rng(77)
    sizes = [int(x) for x in rng.integers(40_000, 120_000, 64)]
    n = sum(sizes)
    k = rng.integers(0, 2_000_000, n).astype(np.int64) * 7 - 5_000_000
    kv = rng.random(n) > 0.01
    v = rng.integers(-10**12, 10**12, n).astype(np.int64)
    vv = rng.random(n) > 0.1
    u = rng.integers(0, 2**32, n).astype(np.uint32)
    f = rng.standard_normal(n) * 1e3
    fv = rng.random(n) > 0.05
    cols = {"k": k, "v": v, "u": u, "f": f}
    valids = [kv, vv, None, fv]
    aggs = [("sum", "v", "s_v"), ("count", "v", "c_v"), ("count", None, "rows"), ("min", "v", "lo_v"),
            ("max", "v", "hi_v"), ("mean", "v", "m_v"), ("sum", "u", "s_u"), ("min", "f", "lo_f"),
            ("max", "f", "hi_f"), ("sum", "f", "s_f"), ("mean", "f", "m_f")]
    got, stats = am.stream_group_by(reader(cols, valids, sizes), ["k"], aggs)
    assert stats["batches"] == 64 and stats["rows"] == n
    assert stats["merge_s"] > 0.0
    assert list(got) == ["k"] + [a[2] for a in aggs]
    ids, K, rep = O.group_keys([k], [kv])
    assert stats["rows_out"] == K
    gk, gkv = got["k"].to_numpy()
    wk, wkv = O.take(k, rep.astype(np.uint32), kv)
    np.testing.assert_array_equal(gkv if gkv is not None else np.ones(K, bool), wkv)
    np.testing.assert_array_equal(gk[wkv], wk[wkv])
    for name, col, valid, op, dt in (("s_v", v, vv, 0, np.int64), ("c_v", v, vv, 2, np.int64),
                                     ("rows", None, None, 1, np.int64), ("lo_v", v, vv, 4, np.int64),
                                     ("hi_v", v, vv, 5, np.int64), ("m_v", v, vv, 3, np.float64),
                                     ("s_u", u, None, 0, np.uint64), ("lo_f", f, fv, 4, np.float64),
                                     ("hi_f", f, fv, 5, np.float64)):
        wb, wok = O.group_agg(ids, K, col, op, valid)
        check_exact(got[name], wb, wok, dt, name)
    for name, op in (("s_f", 0), ("m_f", 3)):
        wb, wok = O.group_agg(ids, K, f, op, fv)
        gv, gva = got[name].to_numpy()
        gva = np.ones(K, bool) if gva is None else gva
        np.testing.assert_array_equal(gva, wok)
        np.testing.assert_allclose(gv[wok], wb.view(np.float64)[wok], rtol=1e-12, atol=1e-9)


def test_stream_group_by_dense_and_multi_key():
    rng = np.random.default_rng(5)
    sizes = [1000, 0, 25_000, 7, 60_000]
    n = sum(sizes)
    d = rng.integers(-3, 1003, n).astype(np.int32)
    g = rng.integers(0, 4, n).astype(np.int16)
    v = rng.integers(-1000, 1000, n).astype(np.int64)
    cols = {"d": d, "g": g, "v": v}
    got, _ = am.stream_group_by(reader(cols, [None, None, None], sizes), ["d"],
                                [("sum", "v", "s"), ("count", None, "n"), ("mean", "v", "m")], dense_key_count=1000)
    inr = (d >= 0) & (d < 1000)
    present = np.unique(d[inr])
    gk, _ = got["d"].to_numpy()
    assert gk.dtype == np.int32 and gk.tolist() == present.tolist()
    s, _ = got["s"].to_numpy()
    cnt, _ = got["n"].to_numpy()
    want_s = np.bincount(d[inr], weights=None, minlength=1000)
    np.testing.assert_array_equal(cnt, want_s[present])
    sums = np.zeros(1000, np.int64)
    np.add.at(sums, d[inr], v[inr])
    np.testing.assert_array_equal(s, sums[present])
    m, _ = got["m"].to_numpy()
    np.testing.assert_array_equal(m, sums[present].astype(np.float64) / want_s[present])
    # Two key columns: lexicographic ascending order, every group once.
    got2, _ = am.stream_group_by(reader(cols, [None, None, None], sizes), ["g", "d"], [("count", None, "n")])
    ids, K, rep = O.group_keys([g, d])
    g2, _ = got2["g"].to_numpy()
    d2, _ = got2["d"].to_numpy()
    np.testing.assert_array_equal(g2, g[rep])
    np.testing.assert_array_equal(d2, d[rep])


def test_stream_group_by_edges_and_refusals():
    cols = {"k": np.arange(10, dtype=np.int64), "v": np.arange(10, dtype=np.float64)}
    got, stats = am.stream_group_by(reader(cols, [None, None], []), ["k"], [("sum", "v", "s"), ("count", None, "n")])
    assert stats["batches"] == 0 and len(got["k"]) == 0 and got["s"].format == "g" and got["n"].format == "l"
    with pytest.raises(am.ArrowMetalError, match="no column named zz"):
        am.stream_group_by(reader(cols, [None, None], [10]), ["zz"], [("count", None, "n")])
    with pytest.raises(am.ArrowMetalError, match="streaming aggregate op 5 is not supported by this build"):
        am.stream_group_by(reader(cols, [None, None], [10]), ["k"], [("variance", "v", "x")])
    with pytest.raises(am.ArrowMetalError, match="a column of NULL is only valid for count"):
        am.stream_group_by(reader(cols, [None, None], [10]), ["k"], [("sum", None, "x")])

"""GPU parity of am_query's fused dense group-by (the DuckDB extension's small-span form,
duckdb-extension/src/arrowmetal_rewrite.cpp:1245-1257, 1319-1350) against the oracle's dense group-by
(or_group_by_dense) on the same keys and transformed values, through the C ABI."""
import numpy as np
import pytest

import arrowmetal_amd as am
import oracle_py as O

pytestmark = pytest.mark.gpu

DENSE_AGG = {"sum": 0, "count_rows": 1, "min": 2, "max": 3, "mean": 4, "count": 5}


def want(keys, kvalid, K, bias, agg, values, valid, dtype):
    shifted = keys.astype(np.int64) + bias
    bits, ok = O.group_by_dense(shifted, K, values, DENSE_AGG[agg], kvalid, valid)
    dtype = np.dtype(dtype)
    out = bits.view(dtype) if dtype.itemsize == 8 else (bits & np.uint64((1 << 8 * dtype.itemsize) - 1)).astype(
        f"u{dtype.itemsize}").view(dtype)
    return out, ok


def check(got, w, label):
    gv, gva = got.to_numpy()
    wv, wva = w
    gva = np.ones(len(gv), bool) if gva is None else gva
    np.testing.assert_array_equal(gva, wva, err_msg=label)
    np.testing.assert_array_equal(gv[wva].view(f"u{gv.dtype.itemsize}"), wv[wva].view(f"u{wv.dtype.itemsize}"),
                                  err_msg=label)


@pytest.mark.parametrize("n,span", [(1, 4), (1000, 37), (2_000_003, 3000)])
def test_duckdb_dense_group_by_query(n, span):
    rng = np.random.default_rng(n)
    key_min = -12345
    c0 = rng.integers(key_min - 3, key_min + span + 3, n).astype(np.int32)
    k_valid = rng.random(n) > 0.02
    c1 = rng.integers(-2**62, 2**62, n)
    v1 = rng.random(n) > 0.1
    c2 = rng.integers(-2**31, 2**31, n).astype(np.int32)
    c3 = (rng.standard_normal(n) * 100).astype(np.float32)
    c4 = rng.integers(0, 2**32, n).astype(np.uint32)
    cols = {"c0": am.MetalArray.from_numpy(c0, k_valid), "c1": am.MetalArray.from_numpy(c1, v1),
            "c2": am.MetalArray.from_numpy(c2), "c3": am.MetalArray.from_numpy(c3), "c4": am.MetalArray.from_numpy(c4)}
    key = f'(sub (cast (col "c0") i64) (i64 {key_min}))'
    text = (f'(query (group_by {span} "key" {key}) (aggregate (count "n0") (count "n1" (col "c1")) '
            f'(sum "n2" (shr (col "c1") (i64 32))) (sum "n3" (bit_and (col "c1") (i64 4294967295))) '
            f'(min "n4" (col "c2")) (max "n5" (col "c3")) (sum "n6" (col "c2")) (mean "n7" (col "c2")) '
            f'(sum "n8" (col "c4")) (sum "n9" (col "c3")) (max "n10" (col "c4"))))')
    got = am.query(cols, text)
    assert list(got) == ["key"] + [f"n{i}" for i in range(11)]
    kv, _ = got["key"].to_numpy()
    assert kv.dtype == np.int32 and kv.tolist() == list(range(span))
    b = -key_min
    check(got["n0"], want(c0, k_valid, span, b, "count_rows", None, None, np.int64), "count rows")
    check(got["n1"], want(c0, k_valid, span, b, "count", c1, v1, np.int64), "count values")
    check(got["n2"], want(c0, k_valid, span, b, "sum", c1 >> 32, v1, np.int64), "sum hi")
    check(got["n3"], want(c0, k_valid, span, b, "sum", c1 & 0xFFFFFFFF, v1, np.int64), "sum lo")
    check(got["n4"], want(c0, k_valid, span, b, "min", c2, None, np.int32), "min i32")
    check(got["n5"], want(c0, k_valid, span, b, "max", c3, None, np.float32), "max f32")
    check(got["n6"], want(c0, k_valid, span, b, "sum", c2, None, np.int64), "sum i32")
    check(got["n7"], want(c0, k_valid, span, b, "mean", c2, None, np.float64), "mean i32")
    check(got["n8"], want(c0, k_valid, span, b, "sum", c4, None, np.uint64), "sum u32")
    # The float32 sum is am_group_by's (its exact sum rounded once to float32, test_gpu_boundary.py pins it
    # against exact rational sums), over the keys moved into [0, span).
    shifted = am.MetalArray.from_numpy((c0.astype(np.int64) + b), k_valid)
    f32w, f32v = shifted.group_by(span).sum(cols["c3"]).to_numpy()
    check(got["n9"], (f32w, np.ones(span, bool) if f32v is None else f32v), "sum f32")
    check(got["n10"], want(c0, k_valid, span, b, "max", c4, None, np.uint32), "max u32")
    assert got["n8"].format == "L" and got["n2"].format == "l"
    # The split sums add back up to the full 64-bit sums (the DuckDB extension's recombination).
    hi, hv = got["n2"].to_numpy()
    lo, _ = got["n3"].to_numpy()
    full, fv = want(c0, k_valid, span, b, "sum", c1, v1, np.int64)
    ok = fv if hv is None else hv & fv
    recombined = (hi[ok].astype(object) * 2**32 + lo[ok].astype(object))
    assert [int(x) % 2**64 for x in recombined] == [int(x) % 2**64 for x in full[ok]]


def test_query_refusals():
    cols = {"k": am.MetalArray.from_numpy(np.arange(10, dtype=np.int32)),
            "f": am.MetalArray.from_numpy(np.arange(10, dtype=np.float64)),
            "l": am.MetalArray.from_numpy(np.arange(10, dtype=np.int64))}

    def err(text):
        with pytest.raises(am.ArrowMetalError) as e:
            am.query(cols, text)
        return str(e.value)

    assert err('(query (group_by 10 "k" (col "k")) (aggregate (sum "s" (col "f"))))') == \
        'Unsupported Arrow type: group_by sum "s" over float64; cast the expression to float32'
    assert "needs a 32-bit or narrower integer or a float32 expression" in \
        err('(query (group_by 10 "k" (col "k")) (aggregate (min "m" (col "l"))))')
    assert err('(query (group_by 10 "k" (col "zz")) (aggregate (count "n")))') == "Invalid ArrowArray: no column named zz"
    assert err('(query (group_by 10 "k" (col "f")) (aggregate (count "n")))').startswith(
        "Unsupported Arrow type: group_by key must be an integer expression")

"""GPU parity of the hash group-by (am_group_by_keys / am_group_agg_ex / am_group_by) against the CPU
oracle: bit-exact group order, ids, representative keys, integer aggregates, float64 min / max, and
correctly rounded float64 sums and means (0 ulp). Cases follow the reference's suites
(Tests/ArrowMetalTests/GroupByTests.swift, GroupByKeysTests.swift, GroupSumExactTests.swift,
HashTableTests.swift)."""
import json
import os

import numpy as np
import pytest

import arrowmetal_amd as am
import oracle_py as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def vectors():
    with open(os.path.join(GOLDEN, "reference_vectors.json")) as fh:
        return json.load(fh)


def runs(pairs):
    out = []
    for v, n in pairs:
        out += [v] * n
    return out


def agg_bits(arr):
    """(bits uint64, valid bool) of an aggregate result, widened like the oracle's."""
    vals, valid = arr.to_numpy()
    if valid is None:
        valid = np.ones(len(vals), dtype=bool)
    if vals.dtype.kind == "i":
        bits = vals.astype(np.int64).view(np.uint64)
    elif vals.dtype == np.float32:
        bits = vals.view(np.uint32).astype(np.uint64)
    elif vals.dtype.kind == "f":
        bits = vals.view(np.uint64)
    else:
        bits = vals.astype(np.uint64)
    return bits, valid


def assert_agg(got, want, label):
    gb, gv = agg_bits(got)
    wb, wv = want
    np.testing.assert_array_equal(gv, wv, err_msg=f"{label}: validity")
    if not np.array_equal(gb[wv], wb[wv]):
        k = int(np.nonzero(gb[wv] != wb[wv])[0][0])
        raise AssertionError(f"{label}: group {k}: got {gb[wv][k]:#x} want {wb[wv][k]:#x}")


def check_grouping(key_cols, key_valids, values=None, vvalid=None, label=""):
    mcols = [am.MetalArray.from_numpy(k, v) for k, v in zip(key_cols, key_valids)]
    g = am.group_by(mcols)
    ids_w, K, rep = O.group_keys(key_cols, key_valids)
    assert g.group_count == K, label
    ids, iv = g.ids().to_numpy()
    assert iv is None
    np.testing.assert_array_equal(ids, ids_w, err_msg=f"{label}: ids")
    for c, (kc, kv) in enumerate(zip(key_cols, key_valids)):
        got, gv = g.key_arrays()[c].to_numpy()
        want_valid = np.ones(K, dtype=bool) if kv is None else kv[rep]
        if gv is None:
            gv = np.ones(K, dtype=bool)
        np.testing.assert_array_equal(gv, want_valid, err_msg=f"{label}: key validity")
        np.testing.assert_array_equal(got[want_valid].view(np.uint8), kc[rep][want_valid].view(np.uint8),
                                      err_msg=f"{label}: keys")
    if values is not None:
        mv = am.MetalArray.from_numpy(values, vvalid)
        # Integer columns: the sum runs first, so the count after it is the one taken from the sum's
        # own count of non-null values; float columns count on their own.
        ops = [1, 4, 5, 2] if values.dtype.kind == "f" else [1, 0, 2, 3, 4, 5]
        if values.dtype == np.float64:
            ops += [0, 3]
        for op in ops:
            name = {0: "sum", 1: "count_all", 2: "count", 3: "mean", 4: "min", 5: "max"}[op]
            got = g.count_all() if op == 1 else getattr(g, name)(mv)
            assert_agg(got, O.group_agg(ids_w, K, values, op, vvalid), f"{label} {name}")
    return g


@pytest.mark.parametrize("n,span", [(0, 1), (1, 1), (33, 7), (4097, 1024), (200003, 100000), (200003, 7)])
def test_sparse_negative_int_keys(n, span):
    rng = np.random.default_rng(n + span)
    keys = rng.integers(-span, span + 1, n).astype(np.int64) * 1_000_003
    kv = rng.random(n) >= 0.125
    vals = rng.integers(-1000, 1001, n).astype(np.int64)
    vv = rng.random(n) >= 0.1
    check_grouping([keys], [kv], vals, vv, f"n={n} span={span}")


def test_pyarrow_fixtures():
    fx = np.load(os.path.join(GOLDEN, "groupby_fixtures.npz"))
    for n in (1000, 20000, 30000):
        keys, kv = fx[f"n{n}_keys"], fx[f"n{n}_kvalid"]
        vals, vv = fx[f"n{n}_vals"], fx[f"n{n}_vvalid"]
        g = am.group_by([am.MetalArray.from_numpy(keys, kv)])
        assert g.group_count == len(fx[f"n{n}_gkeys"])
        mv = am.MetalArray.from_numpy(vals, vv)
        for name in ("sum", "min", "max"):
            got, gv = g._agg(name, mv).to_numpy()
            gv = np.ones(len(got), bool) if gv is None else gv
            np.testing.assert_array_equal(gv, fx[f"n{n}_{name}_valid"])
            np.testing.assert_array_equal(got[gv], fx[f"n{n}_{name}"][gv])
        np.testing.assert_array_equal(g.count(mv).to_numpy()[0], fx[f"n{n}_count"])
        np.testing.assert_array_equal(g.count_all().to_numpy()[0], fx[f"n{n}_count_all"])


def test_reference_vectors():
    v = vectors()
    f = v["group_float_keys"]
    keys = np.array(f["keys"], dtype=np.uint64).view(np.float64)
    g = am.group_by([am.MetalArray.from_numpy(keys, np.array(f["keys_valid"], bool))])
    assert g.group_count == f["group_count"]
    assert g.sum(np.array(f["values"], dtype=np.int64)).to_numpy()[0].tolist() == f["group_sums"]
    d = v["group_dense_fast_path"]
    k = am.MetalArray.from_numpy(np.array(d["keys"], np.int32), np.array(d["keys_valid"], bool))
    vals = am.MetalArray.from_numpy(np.array(d["values"], np.int64))
    s, sv = k.group_by(d["key_count"]).sum(vals).to_numpy()
    assert [int(x) if ok else None for x, ok in zip(s, sv)] == d["dense_sums"]
    g = am.group_by([k])
    assert sorted(g.sum(vals).to_numpy()[0].tolist()) == d["mapped_sums_sorted"]
    a = v["group_all_null"]
    g = am.group_by([am.MetalArray.from_numpy(np.zeros(3, np.int64), np.array(a["keys_valid"], bool))])
    assert g.group_count == 1
    assert g.sum(np.array(a["values"], np.int64)).to_numpy()[0].tolist() == a["sums"]
    w = v["group_sum_wrap"]
    keys = am.MetalArray.from_numpy(np.array(runs(w["keys"]), np.int32))
    for name, dt in (("values", np.int64), ("neg_values", np.int64), ("u64_values", np.uint64)):
        vals = np.array([x % 2**64 for x in runs(w[name])], dtype=np.uint64).view(dt)
        got = keys.group_by(2).sum(am.MetalArray.from_numpy(vals)).to_numpy()[0]
        want = w["sums" if name == "values" else name.replace("values", "sums")]
        assert got.tolist() == want, name
    m = v["group_minmax_dense"]
    k = am.MetalArray.from_numpy(np.array(m["keys"], np.int32), np.array(m["keys_valid"], bool))
    i32 = am.MetalArray.from_numpy(np.array(m["i32"], np.int32), np.array(m["i32_valid"], bool))
    gb = k.group_by(m["key_count"])
    for name, key in (("min", "min"), ("max", "max")):
        x, xv = getattr(gb, name)(i32).to_numpy()
        assert [int(a) if ok else None for a, ok in zip(x, xv)] == m[key]
    assert gb.count_values(i32).to_numpy()[0].tolist() == m["count_values"]
    assert gb.count().to_numpy()[0].tolist() == m["count_rows"]
    u8 = am.MetalArray.from_numpy(np.array(m["u8"], np.uint8))
    x, xv = gb.min(u8).to_numpy()
    assert [int(a) if ok else None for a, ok in zip(x, xv)] == m["u8_min"]
    with pytest.raises(am.ArrowMetalError, match="32-bit or narrower"):
        gb.min(am.MetalArray.from_numpy(np.arange(7, dtype=np.int64)))


def test_float_keys_and_values():
    rng = np.random.default_rng(12)
    n = 100003
    keys = rng.integers(-50, 50, n).astype(np.float64) / 4
    keys[::17] = -0.0
    keys[::29] = np.nan
    keys.view(np.uint64)[::31] = np.uint64(0x7FF0000000000001)
    kv = rng.random(n) > 0.02
    pick = rng.integers(0, 20, n)
    vals = rng.uniform(-1e6, 1e6, n)
    vals[pick == 0] = np.nan
    vals[pick == 1] = -0.0
    vals[pick == 2] = 0.0
    vals[pick == 3] = np.inf
    vv = rng.random(n) > 0.07
    check_grouping([keys], [kv], vals, vv, "f64 keys")


def test_multi_column_and_unique():
    rng = np.random.default_rng(8)
    n = 300007
    a = rng.integers(0, 1000, n).astype(np.int64)
    b = rng.integers(0, 100, n).astype(np.int32)
    av = rng.random(n) > 0.01
    vals = rng.integers(-100, 101, n).astype(np.int64)
    check_grouping([a, b], [av, None], vals, None, "two columns")
    ma, mb = am.MetalArray.from_numpy(a, av), am.MetalArray.from_numpy(b)
    got = am.unique_rows([ma, mb]).to_numpy()[0]
    np.testing.assert_array_equal(got, O.unique_rows([a, b], [av, None]))


def test_exact_float_sum_edges():
    groups = vectors()["group_sum_edges"]["groups"]
    keys, vals = [], []
    for k, g in enumerate(groups):
        keys += [k] * len(g) + [k]
        vals += g + [0]
    keys = np.array(keys, dtype=np.int64)
    vals = np.array(vals, dtype=np.uint64).view(np.float64)
    vv = np.ones(len(vals), dtype=bool)
    pos = 0
    for g in groups:                      # a null value in every group
        pos += len(g)
        vv[pos] = False
        pos += 1
    check_grouping([keys], [None], vals, vv, "edge groups")
    # The same groups spread over many more groups and reversed (device-table path).
    k2 = np.concatenate([keys * 700 + r for r in range(3)])[::-1].copy()
    v2 = np.concatenate([vals] * 3)[::-1].copy()
    vv2 = np.concatenate([vv] * 3)[::-1].copy()
    check_grouping([k2], [None], v2, vv2, "edge groups, device table")
    nanv = vectors()["group_sum_nan_payload"]
    g = am.group_by([am.MetalArray.from_numpy(np.array(nanv["keys"], np.int64))])
    got = g.sum(np.array(nanv["values"], dtype=np.uint64).view(np.float64)).to_numpy()[0].view(np.uint64)
    assert got.tolist() == nanv["sums"]


@pytest.mark.parametrize("K,spread", [(1, False), (7, True), (1024, False), (1025, True), (5000, False),
                                      (100000, True)])
def test_exact_float_sum_random(K, spread):
    rng = np.random.default_rng(K)
    n = 400000
    keys = rng.integers(0, K, n).astype(np.int64)
    kv = rng.random(n) > 0.02
    e = rng.integers(-300, 301, n) if spread else rng.integers(-3, 21, n)
    vals = rng.uniform(-1, 1, n) * np.exp2(e.astype(np.float64))
    vv = rng.random(n) > 0.1
    check_grouping([keys], [kv], vals, vv, f"K={K} spread={spread}")


def test_float_specials_many_groups():
    """Float64 sum, mean, min and max over 100k groups (most outside the LDS cache, in the global
    tables) with NaN, +inf, -inf, -0.0 and nulls mixed into ordinary values: the special-value flags
    and the lowest NaN row of each group, and groups left with only zeros or only specials."""
    rng = np.random.default_rng(0x5BEC)
    n, K = 600_000, 100_000
    keys = rng.integers(0, K, n).astype(np.int64)
    vals = rng.uniform(-1e3, 1e3, n) * np.exp2(rng.integers(-40, 41, n).astype(np.float64))
    pick = rng.random(n)
    vals[pick < 0.004] = np.nan
    vals[(pick >= 0.004) & (pick < 0.008)] = np.inf
    vals[(pick >= 0.008) & (pick < 0.012)] = -np.inf
    vals[(pick >= 0.012) & (pick < 0.03)] = -0.0
    vals.view(np.uint64)[(pick >= 0.03) & (pick < 0.032)] = np.uint64(0xFFF8000000000123)   # a NaN payload
    vv = rng.random(n) > 0.05
    g = am.group_by([am.MetalArray.from_numpy(keys)])
    ids_w, K2, _ = O.group_keys([keys])
    mv = am.MetalArray.from_numpy(vals, vv)
    for op, name in ((0, "sum"), (3, "mean"), (4, "min"), (5, "max")):
        print(f"float specials many groups {name}", flush=True)
        assert_agg(getattr(g, name)(mv), O.group_agg(ids_w, K2, vals, op, vv), f"specials {name}")


def test_cancellation_heavy():
    rng = np.random.default_rng(42)
    for K in (3, 2000):
        m = 200000
        k = rng.integers(0, K, m)
        big = rng.uniform(1e15, 1e17, m)
        small = rng.uniform(-1, 1, m)
        keys = np.stack([k, k, k], 1).reshape(-1).astype(np.int64)
        vals = np.stack([big, -big, small], 1).reshape(-1)
        check_grouping([keys], [None], vals, None, f"cancel K={K}")


def test_zipf_keys_match_oracle():
    """The 100k-key Zipf(1.1) shape of the group-by workload, at 2M rows."""
    n, K = 2_000_000, 100_000
    rng = np.random.default_rng(0x21F)
    w = 1.0 / np.arange(1, K + 1) ** 1.1
    cdf = np.cumsum(w) / w.sum()
    ranks = np.searchsorted(cdf, rng.random(n))
    table = O.splitmix64(0x5EED, K).view(np.int64)
    keys = table[ranks]
    ivals = rng.integers(-10**9, 10**9 + 1, n).astype(np.int64)
    fvals = rng.uniform(-1e6, 1e6, n)
    g = check_grouping([keys], [None], ivals, None, "zipf int")
    ids_w, K2, _ = O.group_keys([keys])
    mv = am.MetalArray.from_numpy(fvals)
    for op, name in ((0, "sum"), (4, "min"), (5, "max"), (3, "mean")):
        assert_agg(getattr(g, name)(mv), O.group_agg(ids_w, K2, fvals, op), f"zipf f64 {name}")


def test_uniform_200_keys_match_oracle():
    """The 200-uniform-key shape of the group-by workload (the LDS-private table path), at 4M rows:
    keys, ids, int64 sum / count and float64 min / max / correctly rounded sum, all bit for bit."""
    n, K = 4_000_000, 200
    rng = np.random.default_rng(0x200)
    table = O.splitmix64(0x200, K).view(np.int64)
    keys = table[rng.integers(0, K, n)]
    ivals = rng.integers(-10**9, 10**9 + 1, n).astype(np.int64)
    fvals = rng.uniform(-1e6, 1e6, n)
    g = check_grouping([keys], [None], ivals, None, "uniform200 int")
    assert g.group_count == K
    ids_w, K2, _ = O.group_keys([keys])
    mv = am.MetalArray.from_numpy(fvals)
    for op, name in ((0, "sum"), (4, "min"), (5, "max"), (3, "mean")):
        assert_agg(getattr(g, name)(mv), O.group_agg(ids_w, K2, fvals, op), f"uniform200 f64 {name}")


def test_dense_group_by_out_of_range_and_nulls():
    rng = np.random.default_rng(3)
    n, K = 300000, 1025
    keys = rng.integers(-5, K + 5, n).astype(np.int32)
    kv = rng.random(n) > 0.05
    vals = rng.integers(-10**6, 10**6, n).astype(np.int64)
    vv = rng.random(n) > 0.1
    mk = am.MetalArray.from_numpy(keys, kv)
    mv = am.MetalArray.from_numpy(vals, vv)
    gb = mk.group_by(K)
    for agg, name in ((0, "sum"), (1, "count"), (4, "mean"), (5, "count_values")):
        got = gb.count() if agg == 1 else getattr(gb, name)(mv)
        assert_agg(got, O.group_by_dense(keys, K, vals if agg != 1 else None, agg, kv, vv), f"dense {name}")
    for it in (np.int64, np.uint32):
        keys2 = np.abs(keys).astype(it)
        gb2 = am.MetalArray.from_numpy(keys2, kv).group_by(K)
        assert_agg(gb2.sum(mv), O.group_by_dense(keys2, K, vals, 0, kv, vv), f"dense sum {it}")


@pytest.mark.parametrize("dtype", [np.int64, np.int32, np.float64])
def test_min_max_struct(dtype):
    """op 6 hash_min_max: a struct<min, max> whose fields equal ops 4 and 5 bit for bit."""
    pa = pytest.importorskip("pyarrow")
    rng = np.random.default_rng(66)
    n, span = 50_001, 900
    keys = rng.integers(0, span, n).astype(np.int64)
    if dtype == np.float64:
        vals = rng.normal(0, 1e3, n)
        vals[rng.random(n) < 0.01] = -0.0
        vals[rng.random(n) < 0.01] = np.nan
    else:
        vals = rng.integers(-10**6, 10**6, n).astype(dtype)
    vv = rng.random(n) >= 0.05
    vv[keys == 7] = False   # one group with no value: null min and max
    g = am.group_by([am.MetalArray.from_numpy(keys)])
    ids_w, K, _ = O.group_keys([keys], [None])
    st = g.min_max(am.MetalArray.from_numpy(vals, vv))
    assert st.format == "+s" and len(st) == K and st.null_count == 0
    arr = st.to_arrow()
    assert isinstance(arr, pa.StructArray) and [arr.type.field(i).name for i in range(2)] == ["min", "max"]
    for f, op in (("min", 4), ("max", 5)):
        col = arr.field(f)
        want_bits, want_valid = O.group_agg(ids_w, K, vals, op, vv)
        got_valid = np.asarray(col.is_valid())
        np.testing.assert_array_equal(got_valid, want_valid, err_msg=f)
        got = col.fill_null(0).to_numpy(zero_copy_only=False)
        got_bits = (got.view(np.uint64) if dtype == np.float64 else got.astype(np.int64).view(np.uint64))
        np.testing.assert_array_equal(got_bits[want_valid], want_bits[want_valid], err_msg=f)


@pytest.mark.parametrize("n,span,null_frac", [(0, 5, 0.0), (1, 5, 1.0), (300, 20, 0.1), (20001, 5000, 0.02),
                                              (300_000, 100_000, 0.01), (4097, 3, 0.0)])
@pytest.mark.parametrize("order", [0, 1])
def test_unique_ex(n, span, null_frac, order):
    """am_unique_ex: first appearance (the null once, at the first null row) or ascending (nulls
    dropped), against the oracle's rows."""
    rng = np.random.default_rng(n + span + order)
    vals = (rng.integers(-span, span + 1, n) * 7919).astype(np.int64)
    valid = rng.random(n) >= null_frac
    got = am.MetalArray.from_numpy(vals, valid).unique("sorted" if order == 1 else "first_appearance")
    rows = O.unique_values(vals, valid, order)
    gv, gvalid = got.to_numpy()
    assert len(gv) == len(rows)
    want_valid = valid[rows]
    if gvalid is None:
        gvalid = np.ones(len(gv), dtype=bool)
    np.testing.assert_array_equal(gvalid, want_valid)
    np.testing.assert_array_equal(gv[want_valid], vals[rows][want_valid])


def test_unique_ex_float_and_narrow():
    rng = np.random.default_rng(5)
    for vals in (rng.integers(-3, 4, 10_000).astype(np.int8),
                 rng.integers(0, 50, 10_000).astype(np.uint16),
                 np.round(rng.normal(0, 4, 10_000)).astype(np.float64)):
        valid = rng.random(len(vals)) >= 0.03
        for order in (0, 1):
            got = am.MetalArray.from_numpy(vals, valid).unique("sorted" if order else "first_appearance")
            rows = O.unique_values(vals, valid, order)
            gv, gvalid = got.to_numpy()
            if gvalid is None:
                gvalid = np.ones(len(gv), dtype=bool)
            np.testing.assert_array_equal(gvalid, valid[rows])
            np.testing.assert_array_equal(gv[gvalid].view(np.uint8), vals[rows][gvalid].view(np.uint8))


@pytest.mark.parametrize("n,span", [(3_000_001, 4_000_000), (1_000_000, 3), (700_000, 65_536)])
def test_table_sizing_from_distinct_estimate(n, span):
    """Inputs of 2^16 rows and more size the hash table from a HyperLogLog estimate of the distinct
    keys: nearly-all-distinct, tiny and mid-size key sets must all group exactly."""
    print(f"sizing n={n} span={span}", flush=True)
    rng = np.random.default_rng(n ^ span)
    keys = (rng.integers(0, span, n).astype(np.int64) * 7_777_777_777) ^ 0x5555
    kv = rng.random(n) >= 0.01
    vals = rng.integers(-10**9, 10**9, n).astype(np.int64)
    check_grouping([keys], [kv], vals, None, f"n={n} span={span}")


@pytest.mark.parametrize("with_nulls", [False, True])
def test_count_reuse_in_any_op_order(with_nulls):
    """count_all after an integer sum of a column without nulls, and count after count_all, reuse the
    table already built; with nulls in the value column they must not. Every order of sum / count_all /
    count gives the oracle's answers, with null keys present."""
    rng = np.random.default_rng(17 + with_nulls)
    n = 300_007
    keys = rng.integers(0, 5_000, n).astype(np.int64)
    kv = rng.random(n) >= 0.02
    vals = rng.integers(-10**12, 10**12, n).astype(np.int64)
    vv = (rng.random(n) >= 0.1) if with_nulls else None
    ids_w, K, _ = O.group_keys([keys], [kv])
    mk, mv = am.MetalArray.from_numpy(keys, kv), am.MetalArray.from_numpy(vals, vv)
    for order in ([0, 1, 2], [1, 2, 0], [2, 1, 0], [1, 0, 2]):
        g = am.group_by([mk])
        for op in order:
            name = {0: "sum", 1: "count_all", 2: "count"}[op]
            got = g.count_all() if op == 1 else getattr(g, name)(mv)
            assert_agg(got, O.group_agg(ids_w, K, vals, op, vv), f"order {order} {name} nulls={with_nulls}")


@pytest.mark.parametrize("case", ["small_span", "negative", "nulls", "all_null", "u64_top", "int8", "two_columns",
                                  "unique_rows", "gaps", "bits_edge", "bits_over"])
def test_range_path_equals_hash_path(case, monkeypatch):
    """Integer keys over a small span take the range path (occupancy, scan, rank; GroupByKeys.swift:268-337);
    the same keys through the hash path (ARROWMETAL_GROUPBY_RANGE=0) give the same ids, keys, counts and
    representative rows, and both match the oracle."""
    rng = np.random.default_rng(len(case))
    n = 300_001
    valid = None
    if case == "small_span":
        cols = [rng.integers(0, 5000, n).astype(np.int64)]
    elif case == "negative":
        cols = [rng.integers(-70_000, -10_000, n).astype(np.int32)]
    elif case == "nulls":
        cols = [rng.integers(100, 300, n).astype(np.uint32)]
        valid = [rng.random(n) > 0.3]
    elif case == "all_null":
        cols = [rng.integers(0, 9, n).astype(np.int64)]
        valid = [np.zeros(n, bool)]
    elif case == "gaps":   # a third of the span present, with nulls: ranks are looked up
        cols = [rng.integers(0, 5000, n).astype(np.int64) * 3 - 7000]
        valid = [rng.random(n) > 0.1]
    elif case in ("bits_edge", "bits_over"):   # the span at, and one past, the LDS bitmap's limit
        span = (1 << 19) + (case == "bits_over")
        k = rng.integers(0, span, n).astype(np.int64)
        k[:2] = [0, span - 1]
        cols = [k + 123]
    elif case == "u64_top":
        cols = [(np.uint64(2**64 - 1000) + rng.integers(0, 1000, n).astype(np.uint64))]
    elif case == "int8":
        cols = [rng.integers(-128, 128, n).astype(np.int8)]
    else:
        cols = [rng.integers(0, 50, n).astype(np.int64), rng.integers(-3, 3, n).astype(np.int16)]
        valid = [rng.random(n) > 0.05, None]
    arrays = [am.MetalArray.from_numpy(c, None if valid is None else valid[i]) for i, c in enumerate(cols)]
    want_ids, K, want_rep = O.group_keys(cols, valid)
    results = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("ARROWMETAL_GROUPBY_RANGE", mode)
        if case == "unique_rows":
            results[mode] = am.unique_rows(arrays).to_numpy()[0]
            continue
        am.prof_reset()
        am.prof_enable(True)
        g = am.group_by(arrays)
        ids = g.ids().to_numpy()[0]
        am.prof_enable(False)
        names = set(am.prof_report())
        assert ("gb_range_scan" in names) == (mode == "1"), names
        assert g.group_count == K
        np.testing.assert_array_equal(ids, want_ids)
        keys = [k.to_numpy() for k in g.key_arrays()]
        cnt, _ = g.count_all().to_numpy()
        results[mode] = (ids, keys, cnt)
    if case == "unique_rows":
        np.testing.assert_array_equal(results["1"], results["0"])
        np.testing.assert_array_equal(results["1"], O.unique_rows(cols, valid))
        return
    for (a, av), (b, bv) in zip(results["1"][1], results["0"][1]):
        ok = np.ones(len(a), bool) if av is None else av
        np.testing.assert_array_equal(ok, np.ones(len(b), bool) if bv is None else bv)
        np.testing.assert_array_equal(a[ok], b[ok])
    np.testing.assert_array_equal(results["1"][2], results["0"][2])


@pytest.mark.parametrize("case", ["int64_nulls", "int32", "uint16", "float64_minmax", "dense_keys", "max_buckets"])
def test_partitioned_aggregates_many_groups(case):
    """More groups than one workgroup's LDS holds: the aggregates run through the partitioned passes
    (rows partitioned by group-id bucket, each bucket reduced in LDS, no global atomics). Sums, counts,
    means, minima and maxima match the oracle bit for bit, with null keys and null values, narrow value
    widths, and the dense-key form up to the largest bucket count (2^21 ids)."""
    rng = np.random.default_rng(sum(map(ord, case)))
    n = 1_500_007
    span = 2**21 if case == "max_buckets" else 120_000
    keys = rng.integers(0, span, n).astype(np.int64)
    kv = rng.random(n) >= 0.03
    vv = rng.random(n) >= 0.08
    if case == "int32":
        vals = rng.integers(-2**31, 2**31, n).astype(np.int32)
    elif case == "uint16":
        vals = rng.integers(0, 2**16, n).astype(np.uint16)
    elif case == "float64_minmax":
        vals = rng.normal(size=n)
        vals[rng.random(n) < 0.01] = np.nan
    else:
        vals = rng.integers(-10**15, 10**15, n).astype(np.int64)
    mv = am.MetalArray.from_numpy(vals, vv)
    am.prof_reset()
    am.prof_enable(True)
    if case in ("dense_keys", "max_buckets"):
        K = span   # max_buckets: 2^21 dense ids, 512 buckets of 4096
        k32 = keys.astype(np.int32)
        gb = am.MetalArray.from_numpy(k32, kv).group_by(K)
        for agg, name in ((0, "sum"), (1, "count"), (4, "mean"), (5, "count_values")):
            got = gb.count() if agg == 1 else getattr(gb, name)(mv)
            assert_agg(got, O.group_by_dense(k32, K, vals if agg != 1 else None, agg, kv, vv), f"dense {name}")
    else:
        g = am.group_by([am.MetalArray.from_numpy(keys, kv)])
        ids_w, K, _ = O.group_keys([keys], [kv])
        assert g.group_count == K
        ops = (4, 5, 1, 2) if case == "float64_minmax" else (0, 1, 2, 3, 4, 5)
        for op in ops:
            name = {0: "sum", 1: "count_all", 2: "count", 3: "mean", 4: "min", 5: "max"}[op]
            got = g.count_all() if op == 1 else getattr(g, name)(mv)
            assert_agg(got, O.group_agg(ids_w, K, vals, op, vv), f"{case} {name}")
    am.prof_enable(False)
    assert "gb_part_scatter" in am.prof_report(), "the partitioned passes did not run"


def test_handle_used_from_another_thread():
    """am_group_by_keys returns while its last kernels may still run on the calling thread's stream;
    aggregates, ids and keys asked for from another thread (its own stream) are ordered after them,
    and the handle can be released there. Results equal the oracle's."""
    import threading
    n, K = 3_000_000, 50_000
    rng = np.random.default_rng(0x7AD)
    table = O.splitmix64(0x7AD, K).view(np.int64)
    keys = table[rng.integers(0, K, n)]
    vals = rng.integers(-10**9, 10**9 + 1, n).astype(np.int64)
    mk, mv = am.MetalArray.from_numpy(keys), am.MetalArray.from_numpy(vals)
    ids_w, K2, _ = O.group_keys([keys])
    for rep in range(3):
        g = am.group_by([mk])   # no wait for the mapping here
        out = {}

        def worker():
            try:
                out["sum"] = g.sum(mv)
                out["ids"] = g.ids().to_numpy()[0]
                out["keys"] = g.key_arrays()[0].to_numpy()[0]
                out["count"] = g.count_all()
            except Exception as e:   # reported below, in the test's thread
                out["error"] = e

        t = threading.Thread(target=worker)
        t.start()
        t.join()
        assert "error" not in out, out.get("error")
        assert g.group_count == K2
        np.testing.assert_array_equal(out["ids"], ids_w)
        assert_agg(out["sum"], O.group_agg(ids_w, K2, vals, 0), f"cross-thread sum, rep {rep}")
        assert_agg(out["count"], O.group_agg(ids_w, K2, vals, 1), f"cross-thread count_all, rep {rep}")
        want_keys = np.sort(np.unique(keys))
        np.testing.assert_array_equal(out["keys"], want_keys)
        del g

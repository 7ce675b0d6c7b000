"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker of the HIP library.

Each oracle function is checked against (a) the reference's own literal golden vectors
(tests/golden/reference_vectors.json), (b) fixtures written by pyarrow / numpy / math.fsum
(tests/golden/*.npz, see tests/golden/make_fixtures.py), and (c) closed forms and brute force written
independently of the oracle's C (Python sorted() with tuple keys, exact rational arithmetic with
fractions.Fraction, dictionaries of first occurrences).
"""
import json
import math
import os
import struct
from fractions import Fraction

import numpy as np
import pytest

import oracle_py as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_vectors():
    with open(os.path.join(GOLDEN, "reference_vectors.json")) as f:
        return json.load(f)


def fb(b):
    return struct.unpack("<d", struct.pack("<Q", int(b)))[0]


def bits(x):
    return struct.unpack("<Q", struct.pack("<d", float(x)))[0]


def runs(pairs):
    out = []
    for v, n in pairs:
        out += [v] * n
    return out


# ---------------------------------------------------------------- argsort


def test_argsort_reference_specials():
    v = load_vectors()["sort_specials"]
    vals = np.array(v["values"], dtype=np.uint64).view(np.float64)
    valid = np.array(v["valid"], dtype=bool)
    assert O.argsort(vals, valid).tolist() == v["ascending"]
    assert O.argsort(vals, valid, descending=True).tolist() == v["descending"]


def test_argsort_reference_total_order():
    v = load_vectors()["sort_total_specials"]
    vals = np.array(v["values"], dtype=np.uint64).view(np.float64)
    valid = np.array(v["valid"], dtype=bool)
    assert O.argsort(vals, valid, descending=True).tolist() == v["desc_ieee"]
    assert O.argsort(vals, valid, descending=True, float_order=1).tolist() == v["desc_total"]
    asc = O.argsort(vals, valid, float_order=1)
    got = [int(vals.view(np.uint64)[i]) for i in asc[:9]]
    assert got == v["sorted_total_asc"]
    desc_first = O.argsort(vals, valid, descending=True, null_placement=1, float_order=1)
    assert [bool(valid[i]) for i in desc_first[:2]] == [False, False]
    got = [int(vals.view(np.uint64)[i]) for i in desc_first[2:]]
    assert got == v["sorted_total_desc_nulls_first"]


@pytest.fixture(scope="module")
def sort_fx():
    return np.load(os.path.join(GOLDEN, "sort_fixtures.npz"))


@pytest.mark.parametrize("n", [0, 1, 2, 33, 257, 4097, 20001])
def test_argsort_matches_pyarrow_and_total_order(sort_fx, n):
    b = sort_fx[f"f64_{n}_bits"]
    valid = sort_fx[f"f64_{n}_valid"]
    f = b.view(np.float64)
    for desc in (0, 1):
        for pl, place in ((0, "at_end"), (1, "at_start")):
            got = O.argsort(f, valid, descending=desc, null_placement=pl, float_order=0)
            np.testing.assert_array_equal(got, sort_fx[f"f64_{n}_ieee_{desc}_{place}"])
            got = O.argsort(f, valid, descending=desc, null_placement=pl, float_order=1)
            np.testing.assert_array_equal(got, sort_fx[f"f64_{n}_total_{desc}_{place}"])
            ints = sort_fx[f"i64_{n}_vals"]
            got = O.argsort(ints, valid, descending=desc, null_placement=pl)
            np.testing.assert_array_equal(got, sort_fx[f"i64_{n}_{desc}_{place}"])


def brute_argsort(vals, valid, desc, place, order):
    """Stable argsort from Python comparisons (no bit tricks for ieee / nan_largest)."""
    n = len(vals)
    rows = [i for i in range(n) if valid is None or valid[i]]
    nulls = [i for i in range(n) if valid is not None and not valid[i]]
    if vals.dtype.kind == "f" and order == 0:
        nans = [i for i in rows if math.isnan(vals[i])]
        rows = [i for i in rows if not math.isnan(vals[i])]
    else:
        nans = []
    if vals.dtype.kind == "f" and order == 2:
        keyf = lambda i: (1, 0.0) if math.isnan(vals[i]) else (0, float(vals[i]) + 0.0)
    elif vals.dtype.kind == "f" and order == 1:
        keyf = lambda i: (lambda b: b ^ 0xFFFFFFFFFFFFFFFF if b >> 63 else b | (1 << 63))(int(vals.view(np.uint64)[i]))
    elif vals.dtype.kind == "f":
        keyf = lambda i: float(vals[i]) + 0.0
    else:
        keyf = lambda i: int(vals[i])
    # Stable: Python's sort is stable; a descending sort reverses the key only, not the ties.
    srt = sorted(rows, key=keyf, reverse=False)
    if desc:
        # Group equal keys, reverse the groups, keep each group's input order.
        groups = []
        for i in srt:
            if groups and keyf(groups[-1][0]) == keyf(i):
                groups[-1].append(i)
            else:
                groups.append([i])
        srt = [i for g in reversed(groups) for i in g]
    return nulls + nans + srt if place == 1 else srt + nans + nulls


@pytest.mark.parametrize("seed", range(6))
def test_argsort_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(0, 300))
    f = rng.integers(-8, 8, n).astype(np.float64) / 2
    sp = np.array([np.nan, -0.0, 0.0, np.inf, -np.inf])
    m = rng.random(n) < 0.2
    f[m] = sp[rng.integers(0, len(sp), int(m.sum()))]
    valid = rng.random(n) >= 0.15
    i64 = rng.integers(-5, 5, n).astype(np.int64)
    u32 = rng.integers(0, 7, n).astype(np.uint32)
    for desc in (0, 1):
        for place in (0, 1):
            for order in (0, 1, 2):
                assert O.argsort(f, valid, desc, place, order).tolist() == brute_argsort(f, valid, desc, place, order)
            assert O.argsort(i64, valid, desc, place).tolist() == brute_argsort(i64, valid, desc, place, 1)
            assert O.argsort(u32, None, desc, place).tolist() == brute_argsort(u32, None, desc, place, 1)


def test_splitmix64_known_values():
    # SplitMix64 from seed 0: the published first outputs of the generator (Vigna).
    got = O.splitmix64(0, 3).tolist()
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


# ---------------------------------------------------------------- take


def test_take_semantics():
    vals = np.array([10, 20, 30, 40], dtype=np.int64)
    valid = np.array([1, 0, 1, 1], dtype=bool)
    idx = np.array([3, 1, 0, 2, 2], dtype=np.int32)
    ivalid = np.array([1, 1, 0, 1, 1], dtype=bool)
    out, ov = O.take(vals, idx, valid, ivalid)
    assert ov.tolist() == [True, False, False, True, True]
    assert out[ov].tolist() == [40, 30, 30]
    with pytest.raises(IndexError, match="array length 4"):
        O.take(vals, np.array([4], dtype=np.int64))
    with pytest.raises(IndexError):
        O.take(vals, np.array([-1], dtype=np.int64))


# ---------------------------------------------------------------- group keys and aggregates


@pytest.fixture(scope="module")
def gb_fx():
    return np.load(os.path.join(GOLDEN, "groupby_fixtures.npz"))


@pytest.mark.parametrize("n", [1000, 20000, 30000])
def test_group_by_matches_pyarrow(gb_fx, n):
    keys, kvalid = gb_fx[f"n{n}_keys"], gb_fx[f"n{n}_kvalid"]
    vals, vvalid = gb_fx[f"n{n}_vals"], gb_fx[f"n{n}_vvalid"]
    ids, K, rep = O.group_keys([keys], [kvalid])
    assert K == len(gb_fx[f"n{n}_gkeys"])
    gk = gb_fx[f"n{n}_gkeys"]
    gkv = gb_fx[f"n{n}_gkeys_valid"]
    # Representative rows carry the group's key (null group last).
    np.testing.assert_array_equal(kvalid[rep], gkv)
    np.testing.assert_array_equal(keys[rep][gkv], gk[gkv])
    for op, name in ((0, "sum"), (4, "min"), (5, "max")):
        out, ov = O.group_agg(ids, K, vals, op, vvalid)
        np.testing.assert_array_equal(ov, gb_fx[f"n{n}_{name}_valid"])
        np.testing.assert_array_equal(out.view(np.int64)[ov], gb_fx[f"n{n}_{name}"][ov])
    out, _ = O.group_agg(ids, K, vals, 2, vvalid)
    np.testing.assert_array_equal(out.view(np.int64), gb_fx[f"n{n}_count"])
    out, _ = O.group_agg(ids, K, None, 1)
    np.testing.assert_array_equal(out.view(np.int64), gb_fx[f"n{n}_count_all"])
    # Representatives are the lowest row of each group.
    first = {}
    for i in range(len(keys)):
        first.setdefault(int(ids[i]), i)
    assert [first[g] for g in range(K)] == rep.tolist()


def test_group_sum_wrap_vectors():
    v = load_vectors()["group_sum_wrap"]
    keys = np.array(runs(v["keys"]), dtype=np.int32)
    for name, dt in (("values", np.int64), ("neg_values", np.int64), ("u64_values", np.uint64)):
        vals = np.array([x % 2**64 for x in runs(v[name])], dtype=np.uint64).view(dt)
        out, ok = O.group_by_dense(keys, 2, vals, 0)
        want = v["sums" if name == "values" else name.replace("values", "sums")]
        assert ok.all()
        if dt == np.uint64:
            assert out.tolist() == want
        else:
            assert out.view(np.int64).tolist() == want


def test_group_dense_minmax_vectors():
    v = load_vectors()["group_minmax_dense"]
    keys = np.array(v["keys"], dtype=np.int32)
    kv = np.array(v["keys_valid"], dtype=bool)
    i32 = np.array(v["i32"], dtype=np.int32)
    iv = np.array(v["i32_valid"], dtype=bool)
    K = v["key_count"]

    def as_list(out, ok, dt):
        return [int(np.array([o], dtype=np.uint64).view(np.int64)[0]) if ok_ else None for o, ok_ in zip(out, ok)]

    out, ok = O.group_by_dense(keys, K, i32, 2, kv, iv)
    assert as_list(out, ok, np.int64) == v["min"]
    out, ok = O.group_by_dense(keys, K, i32, 3, kv, iv)
    assert as_list(out, ok, np.int64) == v["max"]
    out, _ = O.group_by_dense(keys, K, i32, 5, kv, iv)
    assert out.tolist() == v["count_values"]
    out, _ = O.group_by_dense(keys, K, None, 1, kv)
    assert out.tolist() == v["count_rows"]
    u8 = np.array(v["u8"], dtype=np.uint8)
    out, ok = O.group_by_dense(keys, K, u8, 2, kv)
    assert [int(o) if k else None for o, k in zip(out, ok)] == v["u8_min"]
    out, ok = O.group_by_dense(keys, K, u8, 3, kv)
    assert [int(o) if k else None for o, k in zip(out, ok)] == v["u8_max"]


def test_group_float_keys_vector():
    v = load_vectors()["group_float_keys"]
    keys = np.array(v["keys"], dtype=np.uint64).view(np.float64)
    kv = np.array(v["keys_valid"], dtype=bool)
    ids, K, rep = O.group_keys([keys], [kv])
    assert K == v["group_count"]
    out, _ = O.group_agg(ids, K, np.array(v["values"], dtype=np.int64), 0)
    assert out.view(np.int64).tolist() == v["group_sums"]


def test_group_dense_and_mapped_vectors():
    v = load_vectors()["group_dense_fast_path"]
    keys = np.array(v["keys"], dtype=np.int32)
    kv = np.array(v["keys_valid"], dtype=bool)
    vals = np.array(v["values"], dtype=np.int64)
    out, ok = O.group_by_dense(keys, v["key_count"], vals, 0, kv)
    assert [int(o) if k else None for o, k in zip(out.view(np.int64), ok)] == v["dense_sums"]
    ids, K, _ = O.group_keys([keys], [kv])
    out, _ = O.group_agg(ids, K, vals, 0)
    assert sorted(out.view(np.int64).tolist()) == v["mapped_sums_sorted"]
    v = load_vectors()["group_all_null"]
    ids, K, _ = O.group_keys([np.zeros(3, dtype=np.int64)], [np.array(v["keys_valid"], dtype=bool)])
    assert K == v["group_count"]
    out, _ = O.group_agg(ids, K, np.array(v["values"], dtype=np.int64), 0)
    assert out.view(np.int64).tolist() == v["sums"]


def test_group_keys_empty():
    ids, K, rep = O.group_keys([np.zeros(0, dtype=np.int64)])
    assert K == 0 and len(ids) == 0


def test_group_keys_multi_column_lexicographic():
    rng = np.random.default_rng(5)
    n = 3000
    a = rng.integers(-3, 4, n).astype(np.int64)
    b = rng.integers(0, 5, n).astype(np.int32)
    av = rng.random(n) > 0.1
    bv = rng.random(n) > 0.1
    ids, K, rep = O.group_keys([a, b], [av, bv])
    # Brute force: tuples with None after every value, per column.
    def key(i):
        return ((0, int(a[i])) if av[i] else (1, 0), (0, int(b[i])) if bv[i] else (1, 0))
    distinct = sorted({key(i) for i in range(n)})
    assert K == len(distinct)
    pos = {k: g for g, k in enumerate(distinct)}
    assert [pos[key(i)] for i in range(n)] == ids.tolist()


# ---------------------------------------------------------------- exact float64 sums


def exact_sum_expected(xs_bits, mean=False):
    """The reference's rule from first principles: IEEE specials, then exact rational arithmetic
    rounded once by CPython's correctly rounded int / int division."""
    xs = [fb(b) for b in xs_bits]
    if not xs:
        return None
    for b, x in zip(xs_bits, xs):
        if math.isnan(x):
            return int(b) | (1 << 51)
    pinf, ninf = math.inf in xs, -math.inf in xs
    if pinf and ninf:
        return 0x7FF8000000000000
    if pinf:
        return bits(math.inf)
    if ninf:
        return bits(-math.inf)
    if all(x == 0 for x in xs):
        return bits(-0.0) if all(math.copysign(1, x) < 0 for x in xs) else bits(0.0)
    F = sum((Fraction(x) for x in xs), Fraction(0))
    if mean:
        F /= len(xs)
    if F == 0:
        return bits(0.0)
    try:
        return bits(float(F))
    except OverflowError:
        return bits(math.inf if F > 0 else -math.inf)


def test_exact_sum_edge_groups():
    groups = load_vectors()["group_sum_edges"]["groups"]
    keys, vals = [], []
    for k, g in enumerate(groups):
        keys += [k] * len(g)
        vals += g
    keys = np.array(keys, dtype=np.int32)
    vals = np.array(vals, dtype=np.uint64).view(np.float64)
    K = len(groups)
    s, ok = O.group_agg(keys, K, vals, 0)
    m, _ = O.group_agg(keys, K, vals, 3)
    assert ok.all()
    for k, g in enumerate(groups):
        assert int(s[k]) == exact_sum_expected(g), (k, [fb(x) for x in g])
        assert int(m[k]) == exact_sum_expected(g, mean=True), (k, [fb(x) for x in g])
    # The comments of GroupSumExactTests.swift:180-190 fix a few of them outright.
    tiny = 5e-324
    maxd = 1.7976931348623157e308
    def sum_of(g):
        return fb(s[groups.index([bits(x) for x in g])])
    assert sum_of([1e300, 1.0, -1e300]) == 1.0
    assert sum_of([2.0**53, 1.0]) == 2.0**53                  # ties to even
    assert sum_of([2.0**53, 1.0, 2.0**-80]) == 2.0**53 + 2   # just above the tie
    assert sum_of([maxd, 2.0**970]) == math.inf              # half an ulp above maxd
    assert sum_of([maxd, 2.0**969]) == maxd
    assert sum_of([2.0**-1022 - tiny, tiny]) == 2.0**-1022
    assert math.copysign(1, sum_of([-1.0, 1.0, -0.0])) == 1.0


def test_exact_sum_nan_payload_vector():
    v = load_vectors()["group_sum_nan_payload"]
    keys = np.array(v["keys"], dtype=np.int32)
    vals = np.array(v["values"], dtype=np.uint64).view(np.float64)
    s, _ = O.group_agg(keys, 3, vals, 0)
    assert s.tolist() == v["sums"]


def test_exact_sum_matches_fsum(gb_fx):
    keys = gb_fx["fsum_keys"].astype(np.int32)
    vals = gb_fx["fsum_vals"]
    K = len(gb_fx["fsum_sums"])
    s, ok = O.group_agg(keys, K, vals, 0)
    assert ok.all()
    np.testing.assert_array_equal(s, gb_fx["fsum_sums"].view(np.uint64))


def test_exact_sum_random_vs_fraction():
    rng = np.random.default_rng(11)
    n, K = 4000, 37
    keys = rng.integers(0, K, n).astype(np.int32)
    vals = rng.uniform(-1, 1, n) * np.exp2(rng.integers(-1070, 1020, n).astype(np.float64))
    valid = rng.random(n) > 0.1
    s, ok = O.group_agg(keys, K, vals, 0, valid)
    m, _ = O.group_agg(keys, K, vals, 3, valid)
    for k in range(K):
        g = [int(b) for b in vals.view(np.uint64)[(keys == k) & valid]]
        assert int(s[k]) == exact_sum_expected(g)
        assert int(m[k]) == exact_sum_expected(g, mean=True)


def test_float_minmax_brute_force():
    rng = np.random.default_rng(2)
    n, K = 5000, 23
    keys = rng.integers(0, K, n).astype(np.int32)
    pick = rng.integers(0, 20, n)
    vals = rng.uniform(-1e6, 1e6, n)
    vals[pick == 0] = np.nan
    vals[pick == 1] = -0.0
    vals[pick == 2] = 0.0
    vals[pick == 3] = np.inf
    keys[keys == 1] = 2     # group 1 empty
    vals[keys == 3] = np.nan  # group 3 only NaN
    mn, okn = O.group_agg(keys, K, vals, 4)
    mx, okx = O.group_agg(keys, K, vals, 5)
    order = lambda x: (x, math.copysign(1, x))      # -0.0 below +0.0
    for k in range(K):
        xs = [float(x) for x in vals[keys == k] if not math.isnan(x)]
        if not xs:
            assert not okn[k] and not okx[k]
            continue
        assert int(mn[k]) == bits(min(xs, key=order))
        assert int(mx[k]) == bits(max(xs, key=order))


def test_integer_mean():
    keys = np.array([0, 0, 0, 1], dtype=np.int32)
    vals = np.array([1, 2, 2, -7], dtype=np.int64)
    m, ok = O.group_agg(keys, 2, vals, 3)
    assert m.view(np.float64).tolist() == [5 / 3, -7.0]


def test_integer_mean_rounds_sum_and_count_first():
    """The reference's integer mean is d_div(d_from_long(sum), d_from_ulong(count)), i.e.
    Double(sum) / Double(count) with each step rounded (GroupBySource.swift:28-40,
    GroupBy.swift:108-121). Past 2^53 that differs from rounding the exact quotient once: a sum of
    2^53 + 3 over 3 values gives (2^53 + 4) / 3 = 3002399751580332.0, not 3002399751580331.5. A
    mistaken exact-quotient oracle fails here."""
    from fractions import Fraction
    keys = np.array([0, 0, 0, 1, 1], dtype=np.int32)
    vals = np.array([2**53 + 1, 1, 1, -(2**62), -(2**62) + 1], dtype=np.int64)
    m, ok = O.group_agg(keys, 2, vals, 3)
    got = m.view(np.float64).tolist()
    assert got[0] == 3002399751580332.0 and float(Fraction(2**53 + 3, 3)) == 3002399751580331.5
    assert got[1] == float(np.float64(-(2**63) + 1) / 2.0) == -(2.0**62)


# ---------------------------------------------------------------- unique


def test_unique_rows_keep_first():
    rng = np.random.default_rng(9)
    n = 5000
    a = rng.integers(0, 30, n).astype(np.int64)
    b = rng.integers(0, 4, n).astype(np.int64)
    bv = rng.random(n) > 0.2
    got = O.unique_rows([a, b], [None, bv])
    seen, want = set(), []
    for i in range(n):
        k = (int(a[i]), int(b[i]) if bv[i] else None)
        if k not in seen:
            seen.add(k)
            want.append(i)
    assert got.tolist() == want


def test_planted_mistake_is_caught(sort_fx):
    """A plausible oracle bug (ties broken in reverse row order) must fail the fixture check."""
    n = 4097
    f = sort_fx[f"f64_{n}_bits"].view(np.float64)
    valid = sort_fx[f"f64_{n}_valid"]
    good = O.argsort(f, valid)
    want = sort_fx[f"f64_{n}_ieee_0_at_end"]
    np.testing.assert_array_equal(good, want)
    # Reverse the order inside every run of equal keys: an unstable sort.
    keys = np.where(np.isnan(f), np.inf, f + 0.0)
    bad = good.copy()
    i = 0
    while i < len(bad):
        j = i
        while j + 1 < len(bad) and valid[bad[j + 1]] == valid[bad[i]] and keys[bad[j + 1]] == keys[bad[i]]:
            j += 1
        bad[i:j + 1] = bad[i:j + 1][::-1]
        i = j + 1
    assert not np.array_equal(bad, want)


@pytest.mark.parametrize("n,span,null_frac", [(0, 5, 0.0), (1, 5, 1.0), (300, 20, 0.1), (20001, 5000, 0.02),
                                              (4097, 3, 0.0)])
def test_unique_values_matches_arrow_and_numpy(n, span, null_frac):
    """First appearance with the null kept once is Arrow's own `unique`; sorted is numpy's `unique` of
    the non-null values."""
    pa = pytest.importorskip("pyarrow")
    pc = pytest.importorskip("pyarrow.compute")
    rng = np.random.default_rng(n + span)
    vals = (rng.integers(-span, span + 1, n) * 7919).astype(np.int64)
    valid = rng.random(n) >= null_frac
    rows0 = O.unique_values(vals, valid, 0)
    want0 = pc.unique(pa.array(vals, mask=~valid)).to_pylist()
    got0 = [int(vals[r]) if valid[r] else None for r in rows0]
    assert got0 == want0
    rows1 = O.unique_values(vals, valid, 1)
    assert vals[rows1].tolist() == np.unique(vals[valid]).tolist()
    assert all(valid[rows1])


# ---------------------------------------------------------------- lexsort and top-k


def _dense_ranks(col, valid, desc, place, forder):
    """Rank of every row in its column's own stable order, equal keys sharing a rank. Equality comes
    from the (pinned) single-key oracle itself: in the sorted order, a row ties the one before it
    exactly when sorting the pair the other way round keeps it first (nothing strictly smaller)."""
    order = O.argsort(col, valid, desc, place, forder)
    rank = np.empty(len(col), dtype=np.int64)
    r = 0
    for i, row in enumerate(order):
        if i:
            prev = order[i - 1]
            pair = np.array([col[row], col[prev]], dtype=col.dtype)
            pv = None if valid is None else np.array([valid[row], valid[prev]])
            if O.argsort(pair, pv, desc, place, forder)[0] == 1:
                r += 1
        rank[row] = r
    return rank


def _lex_columns(rng, n):
    a = rng.integers(-3, 4, n).astype(np.int64)
    b = rng.choice(np.array([-0.0, 0.0, 1.5, -2.0, np.nan, np.inf, -np.inf], dtype=np.float64), n)
    b[rng.random(n) < 0.05] = np.frombuffer(np.uint64(0xFFF8000000000123).tobytes(), dtype=np.float64)[0]
    c = rng.integers(0, 5, n).astype(np.uint8)
    d = rng.integers(-2, 3, n).astype(np.int32)
    valids = [rng.random(n) > 0.15, rng.random(n) > 0.2, None, rng.random(n) > 0.1]
    return [a, b, c, d], valids


@pytest.mark.parametrize("seed", range(6))
def test_lexsort_brute_force_per_key_options(seed):
    """The definition of lexicographic order: rows ordered by (rank in key 0, rank in key 1, ...,
    row), each rank from that key's own pinned single-key order with its own options."""
    rng = np.random.default_rng(seed)
    n = 400
    cols, valids = _lex_columns(rng, n)
    k = 2 + seed % 3
    pick = rng.permutation(4)[:k]
    cols = [cols[i] for i in pick]
    valids = [valids[i] for i in pick]
    desc = rng.integers(0, 2, k).tolist()
    place = rng.integers(0, 2, k).tolist()
    forder = rng.integers(0, 3, k).tolist()
    ranks = [_dense_ranks(c, v, d, p, f) for c, v, d, p, f in zip(cols, valids, desc, place, forder)]
    want = sorted(range(n), key=lambda i: tuple(r[i] for r in ranks) + (i,))
    got = O.lexsort(cols, valids, desc, place, forder)
    np.testing.assert_array_equal(got, np.array(want, dtype=np.uint32))


@pytest.mark.parametrize("placement", ["at_end", "at_start"])
def test_lexsort_matches_arrow_multi_key_sort(placement):
    """Arrow C++ sort_indices over several keys (one null_placement for every key, ieee floats) is
    the reference's am_lexsort_ex order (MultiSort.swift:9-13)."""
    pa = pytest.importorskip("pyarrow")
    pc = pytest.importorskip("pyarrow.compute")
    rng = np.random.default_rng(11)
    n = 3000
    cols, valids = _lex_columns(rng, n)
    desc = [0, 1, 1, 0]
    t = pa.table({f"c{i}": pa.array(c, mask=None if v is None else ~v) for i, (c, v) in enumerate(zip(cols, valids))})
    keys = [(f"c{i}", "descending" if d else "ascending") for i, d in enumerate(desc)]
    want = pc.sort_indices(t, sort_keys=keys, null_placement=placement).to_numpy()
    p = 1 if placement == "at_start" else 0
    got = O.lexsort(cols, valids, desc, [p] * 4, [0] * 4)
    np.testing.assert_array_equal(got.astype(np.uint64), want)


def test_lexsort_planted_mistake_is_caught():
    """Sorting the keys most significant first (the wrong way round) fails the brute-force order."""
    rng = np.random.default_rng(3)
    n = 300
    a = rng.integers(0, 3, n).astype(np.int64)
    b = rng.integers(0, 3, n).astype(np.int64)
    ra, rb = _dense_ranks(a, None, 0, 0, 0), _dense_ranks(b, None, 0, 0, 0)
    want = np.array(sorted(range(n), key=lambda i: (ra[i], rb[i], i)), dtype=np.uint32)
    wrong = O.lexsort([b, a])   # keys swapped: b most significant
    assert not np.array_equal(wrong, want)
    np.testing.assert_array_equal(O.lexsort([a, b]), want)


# ---------------------------------------------------------------- the plan engine's subset


def _plan_table(rng, n):
    k = rng.integers(-20, 20, n).astype(np.int64)
    kv = rng.random(n) > 0.1
    g = rng.integers(0, 3, n).astype(np.int32)
    v = rng.integers(-1000, 1000, n).astype(np.int64)
    vv = rng.random(n) > 0.2
    f = rng.standard_normal(n)
    return {"k": (k, kv), "g": (g, None), "v": (v, vv), "f": (f, None)}


def _arrow_table(pa, t):
    return pa.table({c: pa.array(v, mask=None if va is None else ~va) for c, (v, va) in t.items()})


@pytest.mark.parametrize("nulls", ["last", "first"])
def test_plan_sort_matches_arrow_sort_by(nulls):
    """A plan sort is the stable multi-key sort of every column (Kernels/MultiSort.swift:201-213);
    Arrow's Table.sort_by gives the same rows (ieee order, one null placement)."""
    pa = pytest.importorskip("pyarrow")
    rng = np.random.default_rng(5)
    t = _plan_table(rng, 2000)
    plan = {"op": "sort", "nulls": nulls, "by": [["g", False], ["k", True]], "input": {"op": "scan", "source": "t"}}
    got = O.plan_run(plan, {"t": t})
    want = _arrow_table(pa, t).sort_by([("g", "ascending"), ("k", "descending")],
                                       null_placement="at_end" if nulls == "last" else "at_start")
    for c in t:
        a = want.column(c)
        ok = ~np.asarray(a.is_null())
        np.testing.assert_array_equal(got[c][0][ok], a.to_numpy(zero_copy_only=False)[ok])
        gv = got[c][1] if got[c][1] is not None else np.ones(len(ok), bool)
        np.testing.assert_array_equal(gv, ok)


def test_plan_group_by_matches_arrow():
    """group_by: keys ascending with the null group last, then sum / count / count_all / mean / min /
    max of integers, against Arrow's hash aggregates sorted by key."""
    pa = pytest.importorskip("pyarrow")
    rng = np.random.default_rng(6)
    t = _plan_table(rng, 3000)
    plan = {"op": "group_by", "keys": [["key", "(col \"k\")"]],
            "aggs": [["sum", "s", "(col \"v\")"], ["count", "c", "(col \"v\")"], ["count", "n", ""],
                     ["mean", "m", "(col \"v\")"], ["min", "lo", "(col \"v\")"], ["max", "hi", "(col \"v\")"]],
            "input": {"op": "scan", "source": "t"}}
    got = O.plan_run(plan, {"t": t})
    at = _arrow_table(pa, t)
    agg = at.group_by("k", use_threads=False).aggregate(
        [("v", "sum"), ("v", "count"), ([], "count_all"), ("v", "mean"), ("v", "min"), ("v", "max")])
    agg = agg.sort_by([("k", "ascending")], null_placement="at_end")
    kv = got["key"]
    want_k = agg.column("k")
    np.testing.assert_array_equal(kv[1], ~np.asarray(want_k.is_null()))
    np.testing.assert_array_equal(kv[0][kv[1]], want_k.drop_null().to_numpy())
    for ours, theirs in (("s", "v_sum"), ("c", "v_count"), ("n", "count_all"), ("m", "v_mean"), ("lo", "v_min"),
                         ("hi", "v_max")):
        a = agg.column(theirs).to_numpy(zero_copy_only=False)
        np.testing.assert_array_equal(got[ours][0], a.astype(got[ours][0].dtype), err_msg=ours)


def test_plan_unique_limit_and_top_k_by_brute_force():
    rng = np.random.default_rng(7)
    t = _plan_table(rng, 500)
    # unique over a subset keeps each combination's first row, in input order.
    got = O.plan_run({"op": "unique", "subset": ["g", "k"], "input": {"op": "scan", "source": "t"}}, {"t": t})
    seen, keep = set(), []
    k, kv = t["k"]
    for i in range(500):
        key = (int(t["g"][0][i]), int(k[i]) if kv[i] else None)
        if key not in seen:
            seen.add(key)
            keep.append(i)
    np.testing.assert_array_equal(got["f"][0], t["f"][0][keep])
    # limit with an offset is a slice; a sort under a limit is the sort's first rows.
    got = O.plan_run({"op": "limit", "count": 7, "offset": 3, "input": {"op": "scan", "source": "t"}}, {"t": t})
    np.testing.assert_array_equal(got["v"][0], t["v"][0][3:10])
    plan = {"op": "limit", "count": 5, "input": {"op": "sort", "by": [["f", True]], "input": {"op": "scan", "source": "t"}}}
    got = O.plan_run(plan, {"t": t})
    np.testing.assert_array_equal(got["f"][0], np.sort(t["f"][0])[::-1][:5])

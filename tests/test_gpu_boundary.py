"""The drop-in boundary's texts and edge rules, checked on the GPU through the C ABI:

- am_last_error carries the reference's ArrowMetalError.description texts
  (Sources/ArrowMetal/MetalContext.swift:19-31): "Invalid ArrowArray: ...", "Unsupported Arrow type: ...",
  "Array length mismatch: a vs b";
- the uint32 row limit: lowered as the reference's tests lower Dispatch.maxRowIndex
  (Tests/ArrowMetalTests/IndexTypeTests.swift:137-160), every index-returning call refuses one row past
  it with the golden message (tests/golden/reference_vectors.json "row_limit_message");
- the dense am_group_by value rules (Sources/ArrowMetalC/ArrowMetalC.swift:416-456): a float32 sum is a
  float64 column, float64 sums and float means are refused;
- the integer mean is Double(sum) / Double(count) (GroupBySource.swift:28-40), pinned by a group whose
  sum passes 2^53, where that rule and the exactly rounded mean differ;
- float32 values under am_group_agg_ex sum / mean: the exact sum (and mean) rounded once, as float64.
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

import arrowmetal_amd as am
import oracle_py as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def err(fn):
    with pytest.raises(am.ArrowMetalError) as e:
        fn()
    return str(e.value)


def test_error_texts_carry_the_reference_prefixes():
    a = am.MetalArray.from_numpy(np.arange(10, dtype=np.float64))
    assert err(lambda: a.argsort(False, "at_end", 7)) == "Invalid ArrowArray: unknown float_order 7"
    assert err(lambda: a.take(np.array([3, 10], dtype=np.int64))) == \
        "Invalid ArrowArray: take: index out of range (array length 10)"
    assert err(lambda: a.take(np.array([1], dtype=np.int16))) == \
        "Unsupported Arrow type: take indices must be int32, int64 or uint32"
    assert err(lambda: a.slice(4, 7)) == "Invalid ArrowArray: slice out of range"
    k = am.MetalArray.from_numpy(np.arange(10, dtype=np.int64) % 3)
    g = am.group_by([k])
    h = am._P()
    rc = am.lib().am_group_by_keys_result(g._h, 1, h)
    assert rc == 1 and am.lib().am_last_error().decode() == "Invalid ArrowArray: key column 1 of 1"
    assert err(lambda: g.sum(np.arange(9, dtype=np.int64))) == "Array length mismatch: 10 vs 9"
    assert err(lambda: g.sum(None)) == "Invalid ArrowArray: am_group_agg_ex op 0 needs a value column"
    assert err(lambda: am.group_by([k, am.MetalArray.from_numpy(np.arange(9, dtype=np.int64))])) == \
        "Array length mismatch: 10 vs 9"
    out = am._P()
    assert am.lib().am_group_agg_ex(g._h, k._h, 99, 0.0, out) == 1
    assert am.lib().am_last_error().decode() == "Invalid ArrowArray: unknown am_group_agg_ex op 99"


def test_row_limit_refusal_matches_the_reference_message():
    template = json.load(open(os.path.join(GOLDEN, "reference_vectors.json")))["row_limit_message"]["template"]
    prev = am.set_max_row_index(99)
    try:
        fits = am.MetalArray.from_numpy((99 - np.arange(100)).astype(np.int32))
        assert fits.argsort().to_numpy()[0][0] == 99
        big = am.MetalArray.from_numpy((np.arange(101) % 7).astype(np.int32))
        bigf = am.MetalArray.from_numpy((np.arange(101) % 5).astype(np.float64))
        want = {op: "Invalid ArrowArray: " + template.format(op=op, last=100).replace("4294967295", "99")
                for op in ("argsort", "row indices", "unique")}
        assert err(lambda: big.argsort()) == want["argsort"]
        assert err(lambda: bigf.argsort(True, "at_start", "total")) == want["argsort"]
        assert err(lambda: am.unique_rows([big])) == want["row indices"]
        assert err(lambda: big.unique("first_appearance")) == want["unique"]
    finally:
        am.set_max_row_index(prev)
    assert am.set_max_row_index(None) == 2**32 - 1
    big = am.MetalArray.from_numpy((np.arange(101) % 7).astype(np.int32))
    np.testing.assert_array_equal(big.argsort().to_numpy()[0], O.argsort(np.arange(101) % 7))


def rn_binary32(x: Fraction) -> np.float32:
    """Round-to-nearest-even of an exact rational to binary32 (normal range), from the definition."""
    if x == 0:
        return np.float32(0.0)
    s = -1 if x < 0 else 1
    a = abs(x)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    scaled = a / Fraction(2) ** (e - 23)          # in [2^23, 2^24)
    q, r = divmod(scaled.numerator, scaled.denominator)
    twice = 2 * r
    if twice > scaled.denominator or (twice == scaled.denominator and q & 1):
        q += 1
    return np.float32(s * q * 2.0 ** (e - 23))


def test_rn_binary32_matches_single_roundings():
    """Pins rn_binary32: a float64 value rounds to float32 once in numpy, so both must agree."""
    rng = np.random.default_rng(8)
    d = rng.standard_normal(20000) * 10.0 ** rng.integers(-30, 30, 20000)
    d = np.concatenate([d, np.float64(np.float32(d[:500])) * (1 + 2.0 ** -24)])   # ties and near-ties
    got = np.array([rn_binary32(Fraction(float(x))) for x in d], dtype=np.float32)
    np.testing.assert_array_equal(got.view(np.uint32), d.astype(np.float32).view(np.uint32))


def test_dense_group_by_float_value_rules():
    rng = np.random.default_rng(21)
    n, K = 200_003, 97
    keys = rng.integers(-3, K + 3, n).astype(np.int32)
    mk = am.MetalArray.from_numpy(keys)
    gb = mk.group_by(K)
    # Integer-valued float32: every float32 running total is exact, so any order gives the exact sum.
    f = rng.integers(-1000, 1000, n).astype(np.float32)
    got, gv = gb.sum(am.MetalArray.from_numpy(f)).to_numpy()
    assert got.dtype == np.float64 and gv is None
    inr = (keys >= 0) & (keys < K)
    want = np.bincount(keys[inr], weights=f[inr].astype(np.float64), minlength=K)
    np.testing.assert_array_equal(got, want)
    # Random float32 values: the exact sum per key, rounded once to float32.
    f2 = (rng.standard_normal(n) * 100).astype(np.float32)
    got2, _ = gb.sum(am.MetalArray.from_numpy(f2)).to_numpy()
    exact = [Fraction(0)] * K
    for k, v in zip(keys[inr].tolist(), f2[inr].astype(np.float64).tolist()):
        exact[k] += Fraction(v)
    want2 = np.array([rn_binary32(x) for x in exact], dtype=np.float32).astype(np.float64)
    np.testing.assert_array_equal(got2.view(np.uint64), want2.view(np.uint64))
    # The pairings the reference refuses.
    d = am.MetalArray.from_numpy(rng.standard_normal(n))
    assert err(lambda: gb.sum(d)) == "Unsupported Arrow type: group-by agg 0 on g"
    assert err(lambda: gb.mean(d)) == "Unsupported Arrow type: group-by agg 4 on g"
    assert err(lambda: gb.mean(am.MetalArray.from_numpy(f))) == "Unsupported Arrow type: group-by agg 4 on f"
    assert err(lambda: gb.min(am.MetalArray.from_numpy(np.arange(n, dtype=np.int64)))) == \
        "Unsupported Arrow type: group-by min/max needs a 32-bit or narrower type"
    assert err(lambda: gb.sum(None)) == "Invalid ArrowArray: values required for this aggregation"
    # min / max of float32 still answer (-0.0 below +0.0, NaN skipped).
    fm = np.array([-0.0, 0.0, np.nan, 1.5] * 10, dtype=np.float32)
    km = am.MetalArray.from_numpy((np.arange(40) % 2).astype(np.int32))
    # Group 0 holds -0.0 and NaN, group 1 holds +0.0 and 1.5.
    lo, _ = km.group_by(2).min(am.MetalArray.from_numpy(fm)).to_numpy()
    assert lo.view(np.uint32).tolist() == [0x80000000, 0x00000000]
    hi, _ = km.group_by(2).max(am.MetalArray.from_numpy(fm)).to_numpy()
    assert hi.view(np.uint32).tolist() == [0x80000000, 0x3FC00000]


def test_integer_mean_is_double_sum_over_double_count():
    """sum = 2^53 + 3 over 3 values: Double(sum) rounds to 2^53 + 4, so the reference's mean is
    (2^53 + 4) / 3 = 3002399751580332.0 exactly, while the exactly rounded mean would be
    3002399751580331.5. Dense and hash forms both give the reference's value."""
    vals = np.array([2**53 + 1, 1, 1, 7], dtype=np.int64)
    keys = np.array([0, 0, 0, 1], dtype=np.int32)
    want = float(np.float64(2**53 + 3) / 3.0)
    assert want == 3002399751580332.0
    assert Fraction(2**53 + 3, 3) != Fraction(want)
    dense, _ = am.MetalArray.from_numpy(keys).group_by(2).mean(am.MetalArray.from_numpy(vals)).to_numpy()
    assert dense.tolist() == [want, 7.0]
    g = am.group_by([am.MetalArray.from_numpy(keys)])
    hashed, _ = g.mean(vals).to_numpy()
    assert hashed.tolist() == [want, 7.0]
    ow, _ = O.group_agg(np.array([0, 0, 0, 1], dtype=np.int32), 2, vals, 3)
    assert ow.view(np.float64).tolist() == [want, 7.0]
    # Many groups above 2^53 against the oracle.
    rng = np.random.default_rng(5)
    n, K = 100_000, 1000
    k = rng.integers(0, K, n).astype(np.int64)
    v = rng.integers(2**50, 2**52, n).astype(np.int64) * rng.choice([-1, 1], n)
    g2 = am.group_by([am.MetalArray.from_numpy(k)])
    ids, K2, _ = O.group_keys([k])
    got, _ = g2.mean(v).to_numpy()
    wb, _ = O.group_agg(ids, K2, v, 3)
    np.testing.assert_array_equal(got.view(np.uint64), wb)


def test_float32_hash_sum_and_mean_are_exact_then_rounded():
    rng = np.random.default_rng(17)
    n, K = 50_000, 37
    keys = rng.integers(0, K, n).astype(np.int64)
    f = (rng.standard_normal(n) * 10.0 ** rng.integers(-6, 6, n)).astype(np.float32)
    f[::1001] = np.float32(3.0e38)          # large values: float64 totals stay finite
    g = am.group_by([am.MetalArray.from_numpy(keys)])
    s, sv = g.sum(f).to_numpy()
    m, _ = g.mean(f).to_numpy()
    assert s.dtype == np.float64 and sv is None
    exact = [Fraction(0)] * K
    cnt = np.bincount(keys, minlength=K)
    for k, v in zip(keys.tolist(), f.astype(np.float64).tolist()):
        exact[k] += Fraction(v)
    np.testing.assert_array_equal(s, np.array([float(x) for x in exact]))
    np.testing.assert_array_equal(m, np.array([float(x / int(c)) for x, c in zip(exact, cnt)]))
    # Specials: NaN (first in row order, quieted and widened), infinities, signed zeros, nulls.
    fs = np.array([1.0, np.inf, -np.inf, 2.0, -0.0, -0.0, 0.0, 5.0], dtype=np.float32)
    ks = np.array([0, 1, 1, 2, 3, 3, 4, 5], dtype=np.int64)
    nan_pay = np.array([0x7F800001], dtype=np.uint32).view(np.float32)[0]
    fs = np.append(fs, nan_pay)
    ks = np.append(ks, 0)
    valid = np.ones(len(fs), dtype=bool)
    valid[7] = False
    gs = am.group_by([am.MetalArray.from_numpy(ks)])
    got, gv = gs.sum(am.MetalArray.from_numpy(fs, valid)).to_numpy()
    assert gv.tolist() == [True, True, True, True, True, False]
    bits = got.view(np.uint64)
    assert bits[0] == 0x7FF8000020000000        # quieted payload of the first NaN, widened
    assert bits[1] == 0x7FF8000000000000        # +inf with -inf
    assert got[2] == 2.0
    assert bits[3] == 0x8000000000000000        # only -0.0
    assert bits[4] == 0

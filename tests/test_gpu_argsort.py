"""GPU parity of the radix argsort (am_argsort / am_argsort_ex / am_argsort_ex2) and am_take against the
CPU oracle, bit-exact, on the sizes and key shapes of the reference's own suites
(Tests/ArrowMetalTests/RadixSortScaleTests.swift:28-201, SortOptionsTests.swift, SortTests.swift)."""
import json
import os

import numpy as np
import pytest

import arrowmetal_amd as am
import oracle_py as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# RadixSortScaleTests.swift:28-29: lane, chunk, block and pass-skip boundaries.
SIZES = [0, 1, 2, 31, 32, 33, 63, 64, 65, 255, 256, 257, 511, 512, 513, 4095, 4096, 4097, 65536,
         (1 << 18) - 1, 1 << 18, (1 << 18) + 1]


def gpu_argsort(vals, valid=None, desc=False, place=0, order=0):
    a = am.MetalArray.from_numpy(vals, valid)
    r = a.argsort(bool(desc), am.NULL_PLACEMENTS[place], am.FLOAT_ORDERS[order])
    assert r.format == "I"
    out, v = r.to_numpy()
    assert v is None
    return out


def check(vals, valid=None, desc=False, place=0, order=0, label=""):
    got = gpu_argsort(vals, valid, desc, place, order)
    want = O.argsort(vals, valid, desc, place, order)
    if not np.array_equal(got, want):
        bad = int(np.nonzero(got != want)[0][0])
        raise AssertionError(f"{label} n={len(vals)} desc={desc} place={place} order={order}: "
                             f"first difference at {bad}: got {got[bad]} want {want[bad]}")


@pytest.mark.parametrize("n", SIZES)
def test_int64_full_range_sizes(n):
    vals = O.splitmix64(0xA11CE501 + n, n).view(np.int64)
    for desc in (False, True):
        check(vals, None, desc, label="i64")


@pytest.mark.parametrize("n", [1, 33, 4097, 70001, (1 << 18) + 1])
def test_float64_orders_and_placements(n):
    rng = np.random.default_rng(n)
    f = rng.uniform(-1e3, 1e3, n).round(0)
    sp = np.array([np.nan, -np.nan, np.inf, -np.inf, -0.0, 0.0, 5e-324, -5e-324], dtype=np.float64)
    m = rng.random(n) < 0.1
    f[m] = sp[rng.integers(0, len(sp), int(m.sum()))]
    b = f.view(np.uint64)
    b[rng.random(n) < 0.01] = np.uint64(0x7FF0000000000001)     # signalling NaN payload
    valid = rng.random(n) >= 0.05
    for desc in (0, 1):
        for place in (0, 1):
            for order in (0, 1, 2):
                check(f, valid, desc, place, order, "f64")


def test_reference_vectors():
    with open(os.path.join(GOLDEN, "reference_vectors.json")) as fh:
        v = json.load(fh)
    s = v["sort_specials"]
    vals = np.array(s["values"], dtype=np.uint64).view(np.float64)
    valid = np.array(s["valid"], dtype=bool)
    assert gpu_argsort(vals, valid).tolist() == s["ascending"]
    assert gpu_argsort(vals, valid, desc=True).tolist() == s["descending"]
    t = v["sort_total_specials"]
    vals = np.array(t["values"], dtype=np.uint64).view(np.float64)
    valid = np.array(t["valid"], dtype=bool)
    assert gpu_argsort(vals, valid, desc=True).tolist() == t["desc_ieee"]
    assert gpu_argsort(vals, valid, desc=True, order=1).tolist() == t["desc_total"]


def test_pyarrow_fixtures():
    fx = np.load(os.path.join(GOLDEN, "sort_fixtures.npz"))
    for n in [0, 1, 2, 33, 257, 4097, 20001]:
        f = fx[f"f64_{n}_bits"].view(np.float64)
        valid = fx[f"f64_{n}_valid"]
        ints = fx[f"i64_{n}_vals"]
        for desc in (0, 1):
            for pl, place in ((0, "at_end"), (1, "at_start")):
                np.testing.assert_array_equal(gpu_argsort(f, valid, desc, pl, 0), fx[f"f64_{n}_ieee_{desc}_{place}"])
                np.testing.assert_array_equal(gpu_argsort(f, valid, desc, pl, 1), fx[f"f64_{n}_total_{desc}_{place}"])
                np.testing.assert_array_equal(gpu_argsort(ints, valid, desc, pl), fx[f"i64_{n}_{desc}_{place}"])


@pytest.mark.parametrize("n", [4097, (1 << 18) + 1, 1 << 20])
def test_key_shapes(n):
    rng = np.random.default_rng(7)
    # All keys equal (every pass skipped; the single pass is a stable compaction).
    check(np.full(n, 42, dtype=np.int64), None, label="all-equal")
    check(np.full(n, 42, dtype=np.int64), rng.random(n) > 0.5, True, 1, label="all-equal-nulls")
    # Constant high digits, then constant low digits (pass skipping both ways).
    check(rng.integers(0, 256, n).astype(np.int64), None, label="low-byte-only")
    check((rng.integers(0, 256, n).astype(np.int64) << 40) | 0x55, None, True, label="one-mid-byte")
    # Monotone input, both directions.
    check(np.arange(n, dtype=np.int64), None, label="ascending-input")
    check(np.arange(n, 0, -1, dtype=np.int64), None, label="descending-input")
    # Nulls at 1/1000, 1/4, 1/2, and all null.
    vals = rng.integers(-1000, 1000, n).astype(np.int64)
    for frac in (0.001, 0.25, 0.5, 1.0):
        valid = rng.random(n) >= frac
        for place in (0, 1):
            check(vals, valid, False, place, label=f"nulls {frac}")
            check(vals, valid, True, place, label=f"nulls {frac}")


@pytest.mark.parametrize("dtype", [np.int32, np.uint32, np.float32, np.int8, np.int16, np.uint8, np.uint16, np.uint64])
def test_other_key_widths(dtype):
    rng = np.random.default_rng(3)
    n = 70001
    if np.dtype(dtype).kind == "f":
        vals = rng.uniform(-100, 100, n).astype(dtype)
        vals[::37] = np.nan
        vals[::41] = -0.0
    else:
        info = np.iinfo(dtype)
        vals = rng.integers(info.min, info.max, n, dtype=dtype, endpoint=True)
    valid = rng.random(n) > 0.1
    for desc in (0, 1):
        for place in (0, 1):
            for order in ((0, 1, 2) if np.dtype(dtype).kind == "f" else (0,)):
                check(vals, valid, desc, place, order, str(np.dtype(dtype)))


def test_offset_slice_import():
    rng = np.random.default_rng(1)
    vals = rng.integers(-50, 50, 10000).astype(np.int64)
    valid = rng.random(10000) > 0.2
    for off in (0, 3, 64, 1001):
        a = am.MetalArray.from_numpy(vals, valid, offset=off)
        got, v = a.argsort(True, "at_start").to_numpy()
        want = O.argsort(vals[off:], valid[off:], True, 1)
        np.testing.assert_array_equal(got, want)


def test_repeat_determinism():
    vals = O.splitmix64(99, 300001).view(np.int64) >> 40      # many duplicate keys
    first = gpu_argsort(vals)
    for _ in range(3):
        np.testing.assert_array_equal(gpu_argsort(vals), first)


def test_bad_float_order_is_an_error():
    a = am.MetalArray.from_numpy(np.arange(4, dtype=np.float64))
    with pytest.raises(am.ArrowMetalError, match="unknown float_order 7"):
        a.argsort(False, "at_end", 7)


def test_batch_matches_unbatched():
    vals = O.splitmix64(5, 100000).view(np.int64)
    a = am.MetalArray.from_numpy(vals)
    want = a.argsort().to_numpy()[0]
    with am.batch():
        r1 = a.argsort()
        r2 = a.argsort(True)
    np.testing.assert_array_equal(r1.to_numpy()[0], want)
    np.testing.assert_array_equal(r2.to_numpy()[0], O.argsort(vals, None, True))


# ---------------------------------------------------------------- take


def test_take_matches_oracle():
    rng = np.random.default_rng(4)
    n = 100003
    vals = rng.uniform(-1e6, 1e6, n)
    valid = rng.random(n) > 0.1
    a = am.MetalArray.from_numpy(vals, valid)
    for it in (np.int32, np.int64, np.uint32):
        idx = rng.integers(0, n, 77777).astype(it)
        ivalid = rng.random(77777) > 0.05
        got, gv = a.take(am.MetalArray.from_numpy(idx, ivalid)).to_numpy()
        want, wv = O.take(vals, idx, valid, ivalid)
        np.testing.assert_array_equal(gv, wv)
        np.testing.assert_array_equal(got.view(np.uint64)[wv], want.view(np.uint64)[wv])
    got, gv = a.take(np.arange(n, dtype=np.uint32)[::-1].copy()).to_numpy()
    np.testing.assert_array_equal(got.view(np.uint64)[gv], vals[::-1].view(np.uint64)[gv])


def test_take_out_of_range():
    a = am.MetalArray.from_numpy(np.arange(10, dtype=np.int64))
    with pytest.raises(am.ArrowMetalError, match=r"take: index out of range \(array length 10\)"):
        a.take(np.array([1, 10], dtype=np.int64))
    with pytest.raises(am.ArrowMetalError, match="out of range"):
        a.take(np.array([-1], dtype=np.int32))


def test_order_by_gather_config_shape():
    """ORDER BY: argsort descending nulls last, then gather a float64 payload (20 B/row)."""
    n = 1_000_003
    keys = (O.splitmix64(0xB0B, n) >> np.uint64(1)).view(np.int64) - (1 << 62)
    valid = (O.splitmix64(0xC0C, n) >> np.uint64(11)).astype(np.float64) / 2.0**53 < 0.98
    payload = (O.splitmix64(0xD0D, n) >> np.uint64(11)).astype(np.float64) / 2.0**53 * 2e6 - 1e6
    k = am.MetalArray.from_numpy(keys, valid)
    p = am.MetalArray.from_numpy(payload)
    idx = k.argsort(True, "at_end")
    got = p.take(idx).to_numpy()[0]
    want_idx = O.argsort(keys, valid, True, 0)
    np.testing.assert_array_equal(idx.to_numpy()[0], want_idx)
    np.testing.assert_array_equal(got, payload[want_idx])


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("width", [8, 4])
def test_large_tiles_first_pass_and_nulls(order, width):
    """9M rows, NaN and null bins included, against the oracle. 64-bit keys take the hybrid sort;
    32-bit keys keep the LSD passes, whose scatter passes take 20 keys per thread above 8.4M rows and
    whose first pass takes 16: both tile shapes and their look-back chains."""
    n = 9_000_001
    print(f"large sort n={n} order={order} width={width}", flush=True)
    bits = O.splitmix64(0x1A2B + order, n)
    valid = (O.splitmix64(0x3C4D, n) >> np.uint64(11)).astype(np.float64) / 2.0**53 < 0.97
    if order == 0:
        vals = (bits >> np.uint64(1)).view(np.int64) - (1 << 62)
        if width == 4:
            vals = (vals >> 32).astype(np.int32)
    else:
        vals = ((bits >> np.uint64(11)).astype(np.float64) / 2.0**53 - 0.5) * 1e6
        vals[(bits & np.uint64(1023)) == 7] = np.nan
        vals[(bits & np.uint64(1023)) == 9] = -0.0
        if width == 4:
            vals = vals.astype(np.float32)
    for desc, place in ((True, 0), (False, 1)):
        check(vals, valid, desc, place, order, label="large")


def sort_with_kernels(vals, valid=None, desc=False, place=0, order=0):
    """The GPU argsort with kernel timing on: (indices, names of the kernels it launched)."""
    am.prof_reset()
    am.prof_enable(True)
    try:
        got = gpu_argsort(vals, valid, desc, place, order)
    finally:
        am.prof_enable(False)
    names = set(am.prof_report())
    am.prof_reset()
    return got, names


def msd_case(case, n):
    bits = O.splitmix64(0x7E57 + len(case), n)
    if case == "uniform":
        return bits.view(np.int64), None, "hybrid"
    if case == "low_digits_constant":   # only the top two digits vary: the local sort has no pass
        return ((bits & np.uint64(0xFFFF000000000000)) | np.uint64(0x1234)).view(np.int64), None, "hybrid"
    if case == "correlated_top":        # second digit equals the top one: every top bucket is carried a level
        top = bits >> np.uint64(56)
        k = (top << np.uint64(56)) | (top << np.uint64(48)) | (bits & np.uint64((1 << 48) - 1))
        return k.view(np.int64), None, "hybrid"
    if case == "duplicates":            # one key on 0.5% of the rows: a bucket that goes down every level
        k = bits.copy()
        k[(bits % np.uint64(200)) == 3] = np.uint64(0x0123456789ABCDEF)
        return k.view(np.int64), None, "hybrid"
    if case == "few_distinct":          # 3,000 distinct keys: buckets of equal keys, carried, then copied
        return O.splitmix64(0xD15, 3000)[(bits >> np.uint64(5)) % np.uint64(3000)].view(np.int64), None, "hybrid"
    if case == "narrow_floats":         # float64 in [1, 2) and [-2, -1): two top digits, the levels go deeper
        v = 1.0 + (bits >> np.uint64(12)).astype(np.float64) / 2.0**52
        v[(bits & np.uint64(1)) == 1] *= -1.0
        return v, None, "hybrid"
    if case == "repeats":               # 1/16 of the rows on 2,000 keys (~140 each): runs the bucket LSD queue
        k = bits.copy()
        pick = (bits >> np.uint64(7)) % np.uint64(16) == 0
        k[pick] = O.splitmix64(0x5EED, 2000)[(bits[pick] >> np.uint64(20)) % np.uint64(2000)]
        return k.view(np.int64), None, "hybrid"
    if case == "wide_buckets":          # top 16 bits from 1,500 values: ~3,000-row buckets, 16 keys per thread
        tops = np.unique(O.splitmix64(0x70B5, 1600) >> np.uint64(48))[:1500] << np.uint64(48)
        k = tops[(bits >> np.uint64(3)) % np.uint64(len(tops))] | (bits & np.uint64(0x0000FFFFFFFFFFFF))
        return k.view(np.int64), None, "hybrid"
    if case == "half_range":            # the ORDER BY keys: [-2^62, 2^62), 128 top digits, ~1,500-row buckets
        return ((bits >> np.uint64(1)).view(np.int64) - (1 << 62)), None, "hybrid"
    if case == "float_bits":            # random bit patterns: every exponent, NaNs, infinities, -0.0
        v = bits.view(np.float64).copy()
        # Equal keys must fit one LDS bucket (4096 rows): about 450 rows of -0.0, and a quarter of
        # the 1-in-2048 NaN patterns kept (nan_largest orders every NaN as one key).
        v[(bits % np.uint64(9973)) == 5] = -0.0
        drop = np.isnan(v) & ((bits >> np.uint64(3)) % np.uint64(4) != 0)
        v[drop] = (bits[drop] >> np.uint64(11)).astype(np.float64)
        valid = (bits >> np.uint64(20)) % np.uint64(33) != 0
        return v, valid, "hybrid"
    if case == "skewed":                # top digit nearly constant: the estimate sends it to the LSD passes
        return (bits >> np.uint64(10)).view(np.int64), None, "lsd"
    raise ValueError(case)


@pytest.mark.parametrize("case", ["uniform", "low_digits_constant", "correlated_top", "duplicates",
                                  "few_distinct", "narrow_floats", "repeats", "wide_buckets", "half_range",
                                  "float_bits", "skewed"])
def test_msd_hybrid_and_fallback(case):
    """64-bit keys above 4M rows whose top digit varies take the hybrid sort: a stable partition by the
    top digit, then level by level a segmented partition of every bucket above 4096 rows by the next
    digit (buckets whose rows share the digit carried without moving), then an LDS sort of each
    bucket by its remaining digits (LSD passes in LDS for buckets with many equal keys, a copy for
    buckets of equal keys). A constant top digit keeps the LSD passes. Each case checks the path it
    took and the order against the oracle."""
    n = 4_500_000
    vals, valid, path = msd_case(case, n)
    combos = [(False, 0, 0), (True, 1, 0)] + ([(False, 1, 1), (True, 0, 2)] if vals.dtype == np.float64 else [])
    for desc, place, order in combos:
        print(f"msd case={case} n={n} desc={desc} place={place} order={order}", flush=True)
        got, names = sort_with_kernels(vals, valid, desc, place, order)
        want = O.argsort(vals, valid, desc, place, order)
        if not np.array_equal(got, want):
            bad = int(np.nonzero(got != want)[0][0])
            raise AssertionError(f"{case} desc={desc} place={place} order={order}: first difference at {bad}: "
                                 f"got {got[bad]} want {want[bad]}")
        bucket_sort = "sort_msd_local" in names or "sort_msd_group" in names
        took = ("hybrid" if bucket_sort else "fallback" if "sort_msd_first" in names else "lsd")
        assert took == path, f"{case}: took the {took} path, kernels {sorted(names)}"


def test_msd_list_overflow_falls_back(monkeypatch):
    """A hybrid sort whose bucket-sort list runs out of room (forced here by capping the list at 64
    entries) drops the bucket sort and finishes with the LSD passes, whose order is the same."""
    monkeypatch.setenv("ARROWMETAL_MSD_LIST_CAP", "64")
    vals, valid, _ = msd_case("uniform", 4_500_000)
    for desc, place, order in ((False, 0, 0), (True, 1, 0)):
        print(f"msd list overflow desc={desc} place={place}", flush=True)
        got, names = sort_with_kernels(vals, valid, desc, place, order)
        want = O.argsort(vals, valid, desc, place, order)
        assert np.array_equal(got, want), f"desc={desc}: first difference at {int(np.nonzero(got != want)[0][0])}"
        assert "sort_msd_first" in names and "sort_msd_local" not in names and "sort_msd_group" not in names, \
            sorted(names)


@pytest.mark.parametrize("n", [1 << 18, 300_007, 1_000_000, (1 << 22) - 1])
@pytest.mark.parametrize("case", ["uniform", "float_bits", "half_range", "duplicates"])
def test_msd_hybrid_small_inputs(n, case):
    """From 2^18 rows, 64-bit keys take the hybrid sort with 4K-row units (16K from 2^22 rows): units
    that end inside a tile, top-digit buckets that go straight to the bucket sort, and the levels
    when a bucket is larger."""
    vals, valid, _ = msd_case(case, n)
    for desc, place, order in ((False, 0, 0), (True, 1, 1)):
        print(f"small msd case={case} n={n} desc={desc} place={place} order={order}", flush=True)
        got, names = sort_with_kernels(vals, valid, desc, place, order)
        want = O.argsort(vals, valid, desc, place, order)
        if not np.array_equal(got, want):
            bad = int(np.nonzero(got != want)[0][0])
            raise AssertionError(f"{case} n={n}: first difference at {bad}: got {got[bad]} want {want[bad]}")
        assert "sort_msd_local" in names or "sort_msd_group" in names, f"{case} n={n}: kernels {sorted(names)}"

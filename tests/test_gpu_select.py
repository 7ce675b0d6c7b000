"""GPU parity of am_sort_ex2 (sorted values), am_top_k_ex (radix select) and am_lexsort_ex2 against the
oracle, through the C ABI. Cases follow the reference's suites: top-k agrees with argsort element for
element (Tests/ArrowMetalTests/SortTests.swift:80-83), sorted values keep -0.0 and NaN payloads
(Kernels/Sort.swift:67-93), lexsort per-key options (SortOptionsTests.swift:242-284)."""
import numpy as np
import pytest

import arrowmetal_amd as am
import oracle_py as O

pytestmark = pytest.mark.gpu

ORDERS = ("ieee", "total", "nan_largest")
PLACES = ("at_end", "at_start")


def float_column(rng, n, nulls=0.1):
    x = rng.standard_normal(n) * 10.0 ** rng.integers(-3, 4, n)
    sel = rng.random(n)
    x[sel < 0.03] = np.nan
    x[(sel >= 0.03) & (sel < 0.05)] = -0.0
    x[(sel >= 0.05) & (sel < 0.06)] = 0.0
    bits = x.view(np.uint64)
    pay = sel > 0.995
    bits[pay] = np.uint64(0xFFF0000000000000) | (O.splitmix64(3, int(pay.sum())) & np.uint64(0xFFFFFFFFFFFFF)) | np.uint64(1)
    valid = rng.random(n) >= nulls
    return x, valid


@pytest.mark.parametrize("n", [0, 1, 37, 4097, 300_001])
def test_sorted_values_keep_bits(n):
    rng = np.random.default_rng(n)
    x, valid = float_column(rng, n)
    a = am.MetalArray.from_numpy(x, valid)
    for desc in (False, True):
        for p, place in enumerate(PLACES):
            for f, fo in enumerate(ORDERS):
                got, gv = a.sort(desc, place, fo).to_numpy()
                order = O.argsort(x, valid, desc, p, f)
                want_valid = valid[order]
                if gv is None:
                    gv = np.ones(n, dtype=bool)
                np.testing.assert_array_equal(gv, want_valid)
                np.testing.assert_array_equal(got.view(np.uint64)[gv], x[order].view(np.uint64)[want_valid])
    ints = rng.integers(-2**63, 2**63 - 1, n, dtype=np.int64)
    got, _ = am.MetalArray.from_numpy(ints).sort(True).to_numpy()
    np.testing.assert_array_equal(got, ints[O.argsort(ints, None, True)])


@pytest.mark.parametrize("n,ks", [(1, [1, 2]), (1000, [1, 7, 999, 1000, 5000]),
                                  (300_007, [1, 10, 1000, 65_536, 299_000]),
                                  (2_000_003, [1, 100, 100_000])])
def test_top_k_is_the_argsort_prefix(n, ks):
    rng = np.random.default_rng(n)
    ints = O.splitmix64(n, n).view(np.int64)
    ivalid = rng.random(n) > 0.02
    x, fvalid = float_column(rng, n)
    narrow = rng.integers(0, 7, n).astype(np.uint8)          # few distinct values: long runs of ties
    cols = [(ints, ivalid), (ints, None), (x, fvalid), (narrow, None), (narrow.astype(np.int32) - 3, ivalid)]
    for vals, valid in cols:
        a = am.MetalArray.from_numpy(vals, valid)
        for k in ks:
            for largest in (True, False):
                for p, place in enumerate(PLACES):
                    forders = ORDERS if vals.dtype.kind == "f" else ("ieee",)
                    for fo in forders:
                        got, _ = a.top_k(k, largest, place, fo).to_numpy()
                        want = O.top_k(vals, k, valid, largest, p, ORDERS.index(fo))
                        assert np.array_equal(got, want), (vals.dtype, valid is None, k, largest, place, fo)


def test_top_k_takes_the_select_path():
    """A top-k that stays among the values selects: the 1M-row argsort's MSD partition never runs (the
    1,000 chosen values are sorted on their own)."""
    n = 1_000_000
    vals = O.splitmix64(77, n).view(np.int64)
    a = am.MetalArray.from_numpy(vals)
    am.prof_reset()
    am.prof_enable(True)
    got, _ = a.top_k(1000, True).to_numpy()
    am.prof_enable(False)
    names = set(am.prof_report())
    assert "topk_select" in names and "sort_msd_first" not in names
    np.testing.assert_array_equal(got, O.argsort(vals, None, True)[:1000])


def test_top_k_errors_and_edges():
    a = am.MetalArray.from_numpy(np.arange(10, dtype=np.int64))
    with pytest.raises(am.ArrowMetalError, match="Invalid ArrowArray: top_k needs k >= 0"):
        a.top_k(-1)
    assert len(a.top_k(0)) == 0
    np.testing.assert_array_equal(a.top_k(100).to_numpy()[0], np.arange(9, -1, -1))
    e = am.MetalArray.from_numpy(np.zeros(0, dtype=np.float64))
    assert len(e.top_k(5)) == 0
    prev = am.set_max_row_index(99)
    try:
        big = am.MetalArray.from_numpy(np.arange(101, dtype=np.int32))
        with pytest.raises(am.ArrowMetalError, match=r"top_k: row 100 does not fit the UInt32 index type"):
            big.top_k(3)
        with pytest.raises(am.ArrowMetalError, match=r"lexsort: row 100 does not fit the UInt32 index type"):
            am.lexsort([big, big])
        assert len(am.MetalArray.from_numpy(np.arange(100, dtype=np.int32)).top_k(1)) == 1
    finally:
        am.set_max_row_index(None)


@pytest.mark.parametrize("n", [1, 1000, 200_003])
def test_lexsort_matches_oracle(n):
    rng = np.random.default_rng(n + 5)
    a = rng.integers(-3, 4, n).astype(np.int64)
    b, bvalid = float_column(rng, n)
    c = rng.integers(0, 5, n).astype(np.uint8)
    d = O.splitmix64(n, n).view(np.int64)
    avalid = rng.random(n) > 0.1
    cols = [(a, avalid), (b, bvalid), (c, None), (d, None)]
    mcols = [am.MetalArray.from_numpy(v, m) for v, m in cols]
    for trial in range(6):
        k = 2 + trial % 3
        pick = rng.permutation(4)[:k]
        desc = rng.integers(0, 2, k).tolist()
        place = rng.integers(0, 2, k).tolist()
        forder = rng.integers(0, 3, k).tolist()
        got, _ = am.lexsort([mcols[i] for i in pick], desc, place, forder).to_numpy()
        want = O.lexsort([cols[i][0] for i in pick], [cols[i][1] for i in pick], desc, place, forder)
        np.testing.assert_array_equal(got, want, err_msg=f"pick {pick} desc {desc} place {place} forder {forder}")
    # Defaults (NULL option arrays) and the one-key case.
    got, _ = am.lexsort(mcols[:2]).to_numpy()
    np.testing.assert_array_equal(got, O.lexsort([a, b], [avalid, bvalid]))
    got, _ = am.lexsort(mcols[3:]).to_numpy()
    np.testing.assert_array_equal(got, O.argsort(d))


def test_lexsort_errors():
    a = am.MetalArray.from_numpy(np.arange(10, dtype=np.int64))
    b = am.MetalArray.from_numpy(np.arange(9, dtype=np.int64))
    with pytest.raises(am.ArrowMetalError, match="Array length mismatch: 10 vs 9"):
        am.lexsort([a, b])
    with pytest.raises(am.ArrowMetalError, match="Invalid ArrowArray: unknown float_order 5"):
        am.lexsort([a, a], float_order=[0, 5])
    out = am._P()
    assert am.lib().am_lexsort_ex2(None, None, None, None, 1, out) == 2

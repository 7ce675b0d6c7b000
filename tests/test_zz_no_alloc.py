"""am_group_by_noalloc (the no-allocation dense group-by) against am_group_by, through the C ABI: two
calls in a row on one GPU, the second with fewer rows and fewer groups, every aggregate it takes, and
its refusals. Outputs go into device buffers the caller owns (arrays imported from NumPy)."""
import ctypes

import numpy as np
import pytest

import arrowmetal_amd as am

pytestmark = pytest.mark.gpu


def noalloc(keys, K, agg, values, out_dtype):
    """Runs am_group_by_noalloc into fresh caller buffers; returns (values, valid, null_count)."""
    out = am.MetalArray.from_numpy(np.zeros(K, dtype=out_dtype))
    bits = am.MetalArray.from_numpy(np.zeros((K + 63) // 64, dtype=np.uint64))
    nulls = am.MetalArray.from_numpy(np.full(1, -7, dtype=np.int64))
    ptr = lambda a: ctypes.c_void_p(a.device_buffers()[0])
    rc = am.lib().am_group_by_noalloc(keys._h, K, agg, values._h if values is not None else None, ptr(out),
                                      ptr(bits), ptr(nulls))
    if rc != 0:
        raise am.ArrowMetalError((am.lib().am_last_error() or b"").decode())
    v, _ = out.to_numpy()
    b, _ = bits.to_numpy()
    valid = np.unpackbits(b.view(np.uint8), bitorder="little")[:K].astype(bool)
    return v, valid, int(nulls.to_numpy()[0][0])


def reference(keys, K, agg, values):
    got, gv = keys.group_by(K)._agg({0: "sum", 1: "count", 2: "min", 3: "max", 4: "mean", 5: "count_values"}[agg],
                                    values).to_numpy()
    return got, (np.ones(K, bool) if gv is None else gv)


CASES = [(0, np.int64), (0, np.int32), (0, np.uint64), (1, None), (2, np.int32), (3, np.float32), (3, np.uint16),
         (4, np.int64), (4, np.uint32), (5, np.float64)]


def make(rng, n, K, vdt):
    keys = rng.integers(-2, K + 2, n).astype(np.int32)
    kvalid = rng.random(n) > 0.03
    if vdt is None:
        return keys, kvalid, None, None
    if np.dtype(vdt).kind == "f":
        v = (rng.standard_normal(n) * 100).astype(vdt)
    else:
        info = np.iinfo(vdt)
        v = rng.integers(max(info.min, -10**15), min(info.max, 10**15), n, dtype=np.int64).astype(vdt)
    return keys, kvalid, v, rng.random(n) > 0.1


@pytest.mark.parametrize("agg,vdt", CASES)
def test_noalloc_matches_group_by_on_two_calls(agg, vdt):
    rng = np.random.default_rng(agg * 31 + (0 if vdt is None else np.dtype(vdt).itemsize))
    for n, K in ((2_000_003, 65_536), (100_001, 3_000)):   # the second call is smaller
        keys, kvalid, v, vvalid = make(rng, n, K, vdt)
        ka = am.MetalArray.from_numpy(keys, kvalid)
        va = None if v is None else am.MetalArray.from_numpy(v, vvalid)
        want, wvalid = reference(ka, K, agg, va)
        got, gvalid, nulls = noalloc(ka, K, agg, va, want.dtype)
        np.testing.assert_array_equal(gvalid, wvalid)
        assert nulls == int((~wvalid).sum())
        np.testing.assert_array_equal(got[wvalid].view(f"u{got.dtype.itemsize}"),
                                      want[wvalid].view(f"u{want.dtype.itemsize}"))


def test_noalloc_refusals():
    keys = am.MetalArray.from_numpy(np.arange(100, dtype=np.int32) % 7)
    f32 = am.MetalArray.from_numpy(np.ones(100, dtype=np.float32))
    i64 = am.MetalArray.from_numpy(np.ones(100, dtype=np.int64))
    with pytest.raises(am.ArrowMetalError, match="key_count 65537 exceeds the 65536 groups of its fixed workspace"):
        noalloc(keys, 65_537, 0, i64, np.int64)
    with pytest.raises(am.ArrowMetalError, match="does not take the float32 sum"):
        noalloc(keys, 7, 0, f32, np.float64)
    with pytest.raises(am.ArrowMetalError, match="min/max needs a 32-bit or narrower type"):
        noalloc(keys, 7, 2, i64, np.int64)
    L = am.lib()
    assert L.am_group_by_noalloc(None, 7, 0, None, None, None, None) == 2

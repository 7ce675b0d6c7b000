"""GPU parity of am_import_chunks (reference include/arrowmetal.h:22-34, ChunkedImport.swift): chunks with
offsets, every validity form (a bitmap with its null count, a bitmap with null_count -1, none), empty
chunks and unaligned bit positions, against the concatenation; a pyarrow ChunkedArray; the refusals."""
import numpy as np
import pytest

import arrowmetal_amd as am

pytestmark = pytest.mark.gpu


def concat(chunks):
    vals, valid = [], []
    for v, va, off, ln in chunks:
        ln = len(v) - off if ln is None else ln
        vals.append(v[off:off + ln])
        if va is None:
            valid.append(np.ones(ln, bool))
        elif va.dtype == bool:
            valid.append(va[off:off + ln])
        else:
            valid.append(np.unpackbits(va, bitorder="little")[off:off + ln].astype(bool))
    return np.concatenate(vals), np.concatenate(valid)


@pytest.mark.parametrize("dtype", [np.int64, np.float64, np.int32, np.uint16, np.int8])
def test_chunks_match_their_concatenation(dtype):
    rng = np.random.default_rng(np.dtype(dtype).itemsize)
    chunks = []
    for i, ln in enumerate([1000, 0, 7, 13, 64, 100_003, 5, 3000]):
        off = int(rng.integers(0, 9))
        v = rng.integers(-100, 100, off + ln).astype(dtype)
        mode = i % 3
        if mode == 0:
            va = rng.random(off + ln) > 0.2
        elif mode == 1:
            va = np.packbits(rng.random(off + ln) > 0.3, bitorder="little")   # null_count -1
        else:
            va = None
        chunks.append((v, va, off, ln))
    got = am.MetalArray.from_chunks(chunks)
    gv, gva = got.to_numpy()
    wv, wva = concat(chunks)
    assert len(gv) == len(wv) and got.null_count == int((~wva).sum())
    np.testing.assert_array_equal(gva if gva is not None else np.ones(len(wv), bool), wva)
    np.testing.assert_array_equal(gv[wva], wv[wva])
    # The imported column sorts like the concatenation.
    np.testing.assert_array_equal(got.argsort(True).to_numpy()[0],
                                  am.MetalArray.from_numpy(wv, wva).argsort(True).to_numpy()[0])


def test_chunks_without_nulls_and_edges():
    a = np.arange(10, dtype=np.int64)
    got = am.MetalArray.from_chunks([(a, None, 0, None), (a * 2, None, 3, 4)])
    v, va = got.to_numpy()
    assert va is None and v.tolist() == list(range(10)) + [6, 8, 10, 12]
    one = am.MetalArray.from_chunks([(a, np.arange(10) % 2 == 0, 1, 5)])   # n == 1 is am_import
    v, va = one.to_numpy()
    assert v.tolist() == [1, 2, 3, 4, 5] and va.tolist() == [False, True, False, True, False]
    empty = am.MetalArray.from_chunks([])
    assert len(empty) == 0


def test_pyarrow_chunked_array_and_refusals():
    pa = pytest.importorskip("pyarrow")
    ca = pa.chunked_array([pa.array([1.5, None, 3.0]), pa.array([], pa.float64()), pa.array([None, -0.0]).slice(1)])
    got = am.MetalArray.from_chunks(ca)
    v, va = got.to_numpy()
    assert va.tolist() == [True, False, True, True]
    assert v[[0, 2, 3]].view(np.uint64).tolist() == np.array([1.5, 3.0, -0.0]).view(np.uint64).tolist()
    assert got.to_arrow().to_pylist() == ca.combine_chunks().to_pylist()
    s = pa.chunked_array([pa.array(["a", "b"]), pa.array(["c"])])
    assert am.MetalArray.from_chunks(s) is None   # 3: not taken, nothing moved
    assert am.lib().am_last_error().decode() == \
        "Unsupported Arrow type: am_import_chunks does not take u arrays; import their concatenation"

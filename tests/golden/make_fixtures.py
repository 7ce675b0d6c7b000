"""Writes the golden fixtures under tests/golden (run in the build container; the files are committed).

Two kinds of data:

* reference_vectors.json -- the literal known-answer vectors of the reference's own tests,
  transcribed as data (inputs and expected outputs only), each entry naming the test it comes from.
  Floats are stored as their 64-bit patterns so NaN payloads and the sign of zero survive JSON.
* *.npz -- inputs and expected outputs produced here by third-party routines the reference's own
  tests check against: pyarrow (Arrow C++) `array_sort_indices` for the `ieee` float order
  (python/tests/test_sort_options.py), numpy's stable argsort of the totalOrder key for `total`
  (test_sort_options.py:59-78), pyarrow `Table.group_by` for integer aggregates
  (python/tests/test_group_by_2_24.py:52-60), and `math.fsum` for float64 sums
  (python/tests/test_group_sum_exact.py:38-61).

Usage: python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import json
import math
import os
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def bits(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def fbits(b: int) -> float:
    return struct.unpack("<d", struct.pack("<Q", b))[0]


NAN = float("nan")
INF = float("inf")


def reference_vectors():
    tiny = 5e-324
    maxd = 1.7976931348623157e308
    v = {}
    # Tests/ArrowMetalTests/SortTests.swift:70-92 -- [1, NaN, inf, -inf, -0, 0, null, NaN]
    v["sort_specials"] = {
        "source": "Tests/ArrowMetalTests/SortTests.swift:70-92",
        "values": [bits(x) for x in [1.0, NAN, INF, -INF, -0.0, 0.0, 0.0, NAN]],
        "valid": [1, 1, 1, 1, 1, 1, 0, 1],
        "ascending": [3, 4, 5, 0, 2, 1, 7, 6],
        "descending": [2, 0, 4, 5, 3, 1, 7, 6],
    }
    # Tests/ArrowMetalTests/SortOptionsTests.swift:209-222
    v["sort_total_specials"] = {
        "source": "Tests/ArrowMetalTests/SortOptionsTests.swift:209-222",
        "values": [bits(x) for x in [1.0, NAN, -0.0, 0.0, -NAN, 0.0, -INF, INF, -0.0, -1.0, 0.0]],
        "valid": [1, 1, 1, 0, 1, 1, 1, 1, 1, 1, 0],
        "desc_ieee": [7, 0, 2, 5, 8, 9, 6, 1, 4, 3, 10],
        "desc_total": [1, 7, 0, 5, 2, 8, 9, 6, 4, 3, 10],
        # sorted(floatOrder: .total) ascending and descending + nulls first, as bit patterns
        "sorted_total_asc": [bits(x) for x in [-NAN, -INF, -1.0, -0.0, -0.0, 0.0, 1.0, INF, NAN]],
        "sorted_total_desc_nulls_first": [bits(x) for x in [NAN, INF, 1.0, 0.0, -0.0, -0.0, -1.0, -INF, -NAN]],
    }
    # Tests/ArrowMetalTests/GroupByTests.swift:42-54 -- carry and wrap
    v["group_sum_wrap"] = {
        "source": "Tests/ArrowMetalTests/GroupByTests.swift:42-54",
        # run-length: [value, repeat] pairs
        "keys": [[0, 5000], [1, 2]],
        "values": [[0xFFFFFFFF, 5000], [2**63 - 1, 1], [1, 1]],
        "sums": [5000 * 0xFFFFFFFF, -2**63],
        "neg_values": [[-3, 5000], [-5, 1], [2, 1]],
        "neg_sums": [-15000, -3],
        "u64_values": [[(2**64 - 1) // 2, 5000], [7, 1], [8, 1]],
        "u64_sums": [((2**64 - 1) // 2 * 5000) % 2**64, 15],
    }
    # Tests/ArrowMetalTests/GroupByTests.swift:80-101 -- int min / max / counts over dense keys
    v["group_minmax_dense"] = {
        "source": "Tests/ArrowMetalTests/GroupByTests.swift:80-101",
        "keys": [0, 1, 0, 2, 1, 0, 0], "keys_valid": [1, 1, 1, 1, 1, 0, 1],
        "i32": [5, -7, 0, 9, 3, 100, -2], "i32_valid": [1, 1, 0, 1, 1, 1, 1],
        "key_count": 4,
        "min": [-2, -7, 9, None], "max": [5, 3, 9, None],
        "count_values": [2, 2, 1, 0], "count_rows": [3, 2, 1, 0],
        "u8": [200, 1, 255, 0, 2, 9, 7],
        "u8_min": [7, 1, 0, None], "u8_max": [255, 2, 0, None],
    }
    # Tests/ArrowMetalTests/GroupByKeysTests.swift:85-102 -- float keys: -0 == 0, NaN one group
    v["group_float_keys"] = {
        "source": "Tests/ArrowMetalTests/GroupByKeysTests.swift:85-102",
        "keys": [bits(x) for x in [0.0, -0.0, NAN, NAN, 1.5, 0.0, 1.5, -0.0, INF, -INF]],
        "keys_valid": [1, 1, 1, 1, 1, 0, 1, 1, 1, 1],
        "values": [1, 2, 4, 8, 16, 32, 64, 128, 256, 512],
        # groups ascending, null last: -inf, 0, 1.5, inf, NaN, null
        "group_sums": [512, 131, 80, 256, 12, 32],
        "group_count": 6,
    }
    # Tests/ArrowMetalTests/GroupByKeysTests.swift:190-212 -- dense ids, empty and all-null
    v["group_dense_fast_path"] = {
        "source": "Tests/ArrowMetalTests/GroupByKeysTests.swift:190-201",
        "keys": [0, 1, 0, 2, 1, 0, 0], "keys_valid": [1, 1, 1, 1, 1, 0, 1],
        "values": [1, 2, 3, 4, 5, 6, 7], "key_count": 4,
        "dense_sums": [11, 7, 4, None],
        "mapped_sums_sorted": [4, 6, 7, 11],
    }
    v["group_all_null"] = {
        "source": "Tests/ArrowMetalTests/GroupByKeysTests.swift:203-212",
        "keys_valid": [0, 0, 0], "values": [1, 2, 3], "group_count": 1, "sums": [6],
    }
    # Tests/ArrowMetalTests/GroupSumExactTests.swift:169-222 -- edge groups (inputs only; the
    # expected sums follow from exact rational arithmetic, see tests/test_oracle.py)
    groups = [
        [1e300, 1, -1e300], [2.0**53, 1, -2.0**53], [1, 2.0**-60, -1], [1e16, 1, 1, 1, -1e16],
        [2.0**60, 2.0**-60, -2.0**60, 2.0**-61], [0.1, 0.2, 0.3, -0.6],
        [-0.0], [-0.0, -0.0], [0.0, -0.0], [-0.0, 0.0], [0.0], [1, -1], [-1, 1, -0.0],
        [tiny], [tiny, tiny, -tiny], [tiny, -tiny], [2.0**-1022, -tiny], [-tiny, -tiny],
        [2.0**-1022 - tiny, tiny], [maxd, maxd], [maxd, maxd, -maxd], [-maxd, -maxd],
        [2.0**1023, 2.0**1023, -2.0**1023, -2.0**1023, 1], [maxd, 2.0**970], [maxd, 2.0**969],
        [NAN], [1, NAN, 2], [INF, 1], [-INF, -1], [INF, -INF], [INF, NAN], [INF, INF],
        [fbits(0xFFF8000000000000)], [3, fbits(0x7FF0000000000001)],
        [1, 2, 3, 4, 5], [2.0**52 + 1, 0.5], [2.0**53, 1], [2.0**53, 1, 2.0**-80],
    ]
    v["group_sum_edges"] = {
        "source": "Tests/ArrowMetalTests/GroupSumExactTests.swift:169-222",
        "groups": [[bits(float(x)) for x in g] for g in groups],
    }
    # GroupSumExactTests.swift:215-221 -- the NaN a group returns
    v["group_sum_nan_payload"] = {
        "source": "Tests/ArrowMetalTests/GroupSumExactTests.swift:215-221",
        "keys": [0, 0, 1, 1, 2],
        "values": [0x7FF0000000000001, 0xFFF8000000000005, bits(INF), bits(-INF), bits(NAN)],
        "sums": [0x7FF8000000000001, 0x7FF8000000000000, 0x7FF8000000000000],
    }
    # Tests/ArrowMetalTests/IndexTypeTests.swift:140 (the uint32 row limit message)
    v["row_limit_message"] = {
        "source": "Sources/ArrowMetal/Kernels/Dispatch.swift:198-203",
        "template": "{op}: row {last} does not fit the UInt32 index type (row numbers go up to 4294967295)",
    }
    return v


def special_doubles():
    # Tests/ArrowMetalTests/SortOptionsTests.swift:68-74 -- values the float orders disagree on
    return np.array([bits(x) for x in [NAN, -NAN]] + [0x7FF0000000000001, 0xFFF8000000000042, 0x7FF8000000000007,
                    0xFFF0000000000003] + [bits(x) for x in [0.0, -0.0, INF, -INF, 5e-324, -5e-324,
                                                             2.2250738585072014e-308 / 4,
                                                             -2.2250738585072014e-308 / 8,
                                                             1.7976931348623157e308, -1.7976931348623157e308]],
                    dtype=np.uint64)


def sort_fixtures(rng):
    import pyarrow as pa
    import pyarrow.compute as pc
    out = {}
    sp = special_doubles()
    for n in [0, 1, 2, 33, 257, 4097, 20001]:
        vals = rng.uniform(-50, 50, n).round(1)               # duplicates: stability matters
        bitsv = vals.view(np.uint64).copy()
        pick = rng.random(n) < 0.15
        bitsv[pick] = sp[rng.integers(0, len(sp), int(pick.sum()))]
        valid = rng.random(n) >= 0.1
        f = bitsv.view(np.float64)
        arr = pa.array(f, type=pa.float64(), mask=~valid)
        out[f"f64_{n}_bits"] = bitsv
        out[f"f64_{n}_valid"] = valid
        for desc in (0, 1):
            for place in ("at_end", "at_start"):
                order = "descending" if desc else "ascending"
                idx = pc.array_sort_indices(arr, order=order, null_placement=place).to_numpy()
                out[f"f64_{n}_ieee_{desc}_{place}"] = idx.astype(np.uint32)
                # totalOrder through numpy's stable argsort of the sign-flip key
                key = np.where(bitsv >> np.uint64(63), ~bitsv, bitsv | np.uint64(1 << 63))
                if desc:
                    key = ~key
                rows = np.nonzero(valid)[0]
                ordered = rows[np.argsort(key[rows], kind="stable")]
                nulls = np.nonzero(~valid)[0]
                tot = np.concatenate([nulls, ordered] if place == "at_start" else [ordered, nulls])
                out[f"f64_{n}_total_{desc}_{place}"] = tot.astype(np.uint32)
        ints = rng.integers(-60, 60, n).astype(np.int64)
        ints[rng.random(n) < 0.01] = np.iinfo(np.int64).min
        ints[rng.random(n) < 0.01] = np.iinfo(np.int64).max
        iarr = pa.array(ints, mask=~valid)
        out[f"i64_{n}_vals"] = ints
        for desc in (0, 1):
            for place in ("at_end", "at_start"):
                order = "descending" if desc else "ascending"
                idx = pc.array_sort_indices(iarr, order=order, null_placement=place).to_numpy()
                out[f"i64_{n}_{desc}_{place}"] = idx.astype(np.uint32)
    np.savez_compressed(os.path.join(HERE, "sort_fixtures.npz"), **out)


def groupby_fixtures(rng):
    import pyarrow as pa
    out = {}
    for n, span in [(1000, 7), (20000, 1000), (30000, 20000)]:
        keys = (rng.integers(-span, span, n) * 1_000_003).astype(np.int64)
        kvalid = rng.random(n) >= 0.05
        vals = rng.integers(-10**9, 10**9, n).astype(np.int64)
        vvalid = rng.random(n) >= 0.1
        t = pa.table({"k": pa.array(keys, mask=~kvalid), "v": pa.array(vals, mask=~vvalid)})
        g = t.group_by("k").aggregate([("v", "sum"), ("v", "count"), ("v", "min"), ("v", "max"),
                                       ("k", "count", pa.compute.CountOptions(mode="all"))])
        # Ascending key, null last: the reference's group order.
        g = g.sort_by([("k", "ascending")])
        kk = g.column("k")
        null_first = kk.null_count > 0 and not kk.is_valid()[0].as_py()
        if null_first:   # pyarrow sorts nulls last by default; keep the check explicit
            raise RuntimeError("unexpected null placement")
        out[f"n{n}_keys"] = keys
        out[f"n{n}_kvalid"] = kvalid
        out[f"n{n}_vals"] = vals
        out[f"n{n}_vvalid"] = vvalid
        gk = g.column("k")
        out[f"n{n}_gkeys"] = gk.fill_null(0).to_numpy()
        out[f"n{n}_gkeys_valid"] = gk.is_valid().to_numpy(zero_copy_only=False)
        for col, name in [("v_sum", "sum"), ("v_min", "min"), ("v_max", "max")]:
            c = g.column(col)
            out[f"n{n}_{name}"] = c.fill_null(0).to_numpy()
            out[f"n{n}_{name}_valid"] = c.is_valid().to_numpy(zero_copy_only=False)
        out[f"n{n}_count"] = g.column("v_count").to_numpy()
        out[f"n{n}_count_all"] = g.column("k_count").to_numpy()
    # Float64 sums checked against math.fsum (finite values, no overflow).
    n, K = 30000, 200
    keys = rng.integers(0, K, n).astype(np.int64)
    exps = rng.integers(-300, 300, n)
    vals = rng.uniform(-1, 1, n) * np.exp2(exps.astype(np.float64))
    vals[::97] = -0.0
    out["fsum_keys"] = keys
    out["fsum_vals"] = vals
    sums = np.zeros(K, dtype=np.float64)
    for k in range(K):
        sums[k] = math.fsum(vals[keys == k].tolist())
    out["fsum_sums"] = sums
    np.savez_compressed(os.path.join(HERE, "groupby_fixtures.npz"), **out)


def main():
    rng = np.random.default_rng(0xA11CE501)
    with open(os.path.join(HERE, "reference_vectors.json"), "w") as f:
        json.dump(reference_vectors(), f, indent=None, separators=(",", ":"))
    sort_fixtures(rng)
    groupby_fixtures(rng)
    for name in sorted(os.listdir(HERE)):
        p = os.path.join(HERE, name)
        if name.endswith((".json", ".npz")):
            print(f"{name}: {os.path.getsize(p)} bytes")


if __name__ == "__main__":
    main()

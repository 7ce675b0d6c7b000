"""Benchmark of the argsort / group-by path on one MI355X.

Workload (the metric's ORDER BY configuration): 50,000,000 rows of int64 keys uniform over
[-2^62, 2^62) with 2% nulls (an independent validity stream), a stable descending argsort with nulls
last (am_argsort_ex2), then the gather of a float64 payload uniform in [-1e6, 1e6] (am_take).
One step = argsort + gather over the HBM-resident columns; rows/s = rows / step time.

Also measured and reported beside it (same JSON line):
  * argsort_1M: the plain ascending argsort of 1,000,000 full-range int64 keys;
  * group_by_50M_zipf_100k: am_group_by_keys over 50M int64 keys drawn Zipf(1.1) from 100,000
    distinct values, then int64 sum, count_all and count (am_group_agg_ex ops 0, 1, 2), and the float64
    min / max / sum of a second value column;
  * roofline: the middle scatter pass of the argsort (the dominant kernel), algorithmic bytes per
    launch (24 B per value row: read key + row, write key + row) over its mean launch time, measured
    with HIP events on the library's own stream (am_prof_*);
  * cpu_baseline: the CPU oracle's ORDER BY (an 8-bit LSD radix argsort + gather, one core) on the
    same 50M rows, timed on this host.

Inputs come from SplitMix64 streams (seeded), generated with NumPy here, imported into HBM once
before timing. `--gpus N` with N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import arrowmetal_amd as am  # noqa: E402

METRIC = "rows/s + %HBM roofline for 50M int64 argsort & group-by; p99 served vs plain"
HBM_PEAK_GBS = 8000.0

GOLDEN = 0x9E3779B97F4A7C15
M1 = 0xBF58476D1CE4E5B9
M2 = 0x94D049BB133111EB


def splitmix64(seed: int, n: int, chunk: int = 1 << 22) -> np.ndarray:
    """SplitMix64 outputs 1..n from `seed` (Tests/ArrowMetalTests/RadixSortScaleTests.swift:14-24)."""
    out = np.empty(n, dtype=np.uint64)
    g = np.uint64(GOLDEN)
    with np.errstate(over="ignore"):
        for a in range(0, n, chunk):
            b = min(n, a + chunk)
            z = np.uint64(seed) + (np.arange(a + 1, b + 1, dtype=np.uint64) * g)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(M1)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(M2)
            out[a:b] = z ^ (z >> np.uint64(31))
    return out


def unit(u: np.ndarray) -> np.ndarray:
    """Uniform [0, 1) doubles from the top 53 bits."""
    return (u >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def order_by_inputs(n: int):
    keys = (splitmix64(0xA11CE501, n) >> np.uint64(1)).view(np.int64) - (1 << 62)
    valid = unit(splitmix64(0xA11CE502, n)) < 0.98
    payload = unit(splitmix64(0xA11CE503, n)) * 2e6 - 1e6
    return keys, valid, payload


def zipf_ranks(n: int, support: int, s: float, seed: int) -> np.ndarray:
    w = 1.0 / np.arange(1, support + 1, dtype=np.float64) ** s
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    out = np.empty(n, dtype=np.int32)
    chunk = 1 << 23
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        out[a:b] = np.minimum(np.searchsorted(cdf, unit(splitmix64(seed + a, b - a))), support - 1)
    return out


def group_by_inputs(n: int, support: int = 100_000):
    table = splitmix64(0x6B0DE5, support).view(np.int64)       # distinct int64 key values
    keys = table[zipf_ranks(n, support, 1.1, 0x21F0)]
    ivals = (splitmix64(0x1AA1, n) % np.uint64(2_000_000_001)).astype(np.int64) - 1_000_000_000
    fvals = unit(splitmix64(0x2BB2, n)) * 2e6 - 1e6
    return keys, ivals, fvals


def timed(fn, steps: int, warmup: int):
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(steps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return ts


def kernel_ms(name: str, prof: dict):
    c, ms = prof.get(name, (0, 0.0))
    return (ms / c if c else None), c


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows", type=int, default=50_000_000)
    ap.add_argument("--skip-cpu", action="store_true", help="omit the CPU baseline leg")
    ap.add_argument("--skip-groupby", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU design is described in DESIGN.md, nothing here launches on a second GPU)",
              file=sys.stderr)
        sys.exit(2)
    n = args.rows
    log = lambda *a: print("[bench]", *a, file=sys.stderr, flush=True)
    log("device:", am.device_name())

    # ---------------------------------------------------------------- ORDER BY (headline)
    keys, valid, payload = order_by_inputs(n)
    m = int(valid.sum())
    k = am.MetalArray.from_numpy(keys, valid)
    p = am.MetalArray.from_numpy(payload)
    log(f"order-by inputs resident: {n} rows, {n - m} nulls")
    lib = am.lib()

    def order_by_step():
        idx = k.argsort(True, "at_end", "ieee")
        out = p.take(idx)
        del idx, out

    ts = timed(order_by_step, args.steps, args.warmup)
    ms = 1e3 * float(np.median(ts))
    best = 1e3 * float(np.min(ts))
    log(f"order-by: median {ms:.3f} ms, best {best:.3f} ms")

    # Kernel times with HIP events on the library's stream, in a separate profiled run.
    am.prof_reset()
    am.prof_enable(True)
    for _ in range(3):
        order_by_step()
    am.prof_enable(False)
    prof = am.prof_report()
    log("kernels:", json.dumps(prof))
    scat_ms, scat_n = kernel_ms("sort_scatter", prof)
    bytes_per_launch = 24 * m
    achieved = bytes_per_launch / (scat_ms * 1e-3) / 1e9 if scat_ms else None

    # Parity of the timed step against the CPU oracle (and the CPU baseline timing).
    idx, _ = k.argsort(True, "at_end", "ieee").to_numpy()
    gathered, _ = p.take(am.MetalArray.from_numpy(idx)).to_numpy()
    cpu = None
    parity = None
    if not args.skip_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_py as O   # the checker and the timed CPU baseline; never the measured GPU path
        t0 = time.perf_counter()
        want = O.argsort(keys, valid, True, 0, 0)
        want_payload = payload[want]
        cpu_s = time.perf_counter() - t0
        parity = bool(np.array_equal(idx, want) and np.array_equal(gathered.view(np.uint64),
                                                                   want_payload.view(np.uint64)))
        cpu = {"value": n / cpu_s, "unit": "rows/s", "cores": 1, "kind": "port",
               "sample": f"the full {n}-row ORDER BY (oracle 8-bit LSD radix argsort, descending, nulls "
                         f"last, then the float64 gather), one run, {cpu_s:.2f} s on one core"}
        log(f"cpu oracle: {cpu_s:.2f} s, parity {parity}")
    del idx, gathered

    # ---------------------------------------------------------------- argsort 1M
    k1 = am.MetalArray.from_numpy(splitmix64(0xA11CE501, 1_000_000).view(np.int64))
    ts1 = timed(lambda: k1.argsort(), max(args.steps, 20), args.warmup)
    ms1 = 1e3 * float(np.median(ts1))
    del k1

    # ---------------------------------------------------------------- group-by 50M, Zipf over 100k
    gb_res = None
    if not args.skip_groupby:
        gkeys, ivals, fvals = group_by_inputs(n)
        gk = am.MetalArray.from_numpy(gkeys)
        gi = am.MetalArray.from_numpy(ivals)
        gf = am.MetalArray.from_numpy(fvals)
        del gkeys, ivals, fvals

        def gb_step():
            g = am.group_by([gk])
            a = g.sum(gi)
            b = g.count_all()
            c = g.count(gi)
            del a, b, c, g

        def gb_float_step():
            g = am.group_by([gk])
            a = g.min(gf)
            b = g.max(gf)
            c = g.sum(gf)
            del a, b, c, g

        tg = timed(gb_step, args.steps, args.warmup)
        tf = timed(gb_float_step, max(3, args.steps // 2), 1)
        am.prof_reset()
        am.prof_enable(True)
        gb_step()
        am.prof_enable(False)
        gprof = am.prof_report()
        log("group-by kernels:", json.dumps(gprof))
        gms = 1e3 * float(np.median(tg))
        fms = 1e3 * float(np.median(tf))
        g = am.group_by([gk])
        gb_res = {
            "workload": "group_by_50M_int64_zipf1.1_100k_keys",
            "groups": int(g.group_count),
            "ms_per_step": round(gms, 4),
            "rows_per_s": n / (gms * 1e-3),
            "step": "am_group_by_keys + am_group_agg_ex sum, count_all, count (int64 values)",
            "b_alg_bytes": 16 * n,
            "hbm_frac_of_b_alg": (16 * n) / (gms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "float_step_ms": round(fms, 4),
            "float_step": "am_group_by_keys + float64 min, max, correctly rounded sum",
            "kernels_ms": {kname: round(v[1] / max(v[0], 1), 4) for kname, v in gprof.items()},
        }
        log(f"group-by: {gms:.3f} ms, float step {fms:.3f} ms, groups {g.group_count}")
        del g, gk, gi, gf

    b_sort = 8 * n + (8 * n + 12 * m + 4 * (n - m)) + 6 * 2 * m * 12 + m * 16 + 2 * ((n + 7) // 8)
    b_take = 20 * n
    result = {
        "metric": METRIC,
        "value": n / (ms * 1e-3),
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "best_ms_per_step": round(best, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (SplitMix64 streams; keys uniform over [-2^62, 2^62), 2% nulls, float64 payload)",
        "config": {"workload": "order_by_50M_int64_desc_nulls_last_then_gather_f64", "rows": n,
                   "null_fraction": round(1 - m / n, 5), "descending": True, "null_placement": "at_end",
                   "float_order": "ieee", "digit_bits": 8, "passes": 8},
        "step_hbm_frac_of_b_alg": (b_sort + b_take) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "b_alg_bytes": {"sort": b_sort, "gather": b_take},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": None,
                     "kernel": "sort_scatter (middle radix pass, sort_mid<uint64,16,0>)",
                     "bytes_per_launch": bytes_per_launch, "mean_launch_ms": scat_ms,
                     "launches": scat_n},
        "kernels_ms": {kname: round(v[1] / max(v[0], 1), 4) for kname, v in prof.items()},
        "parity_vs_oracle": parity,
        "cpu_baseline": cpu,
        "argsort_1M": {"ms": round(ms1, 4), "rows_per_s": 1e6 / (ms1 * 1e-3),
                       "workload": "stable ascending argsort of 1,000,000 full-range int64 keys"},
        "group_by": gb_res,
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()

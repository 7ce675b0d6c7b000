"""Benchmark of the argsort / group-by path on one MI355X.

Headline workload (the metric's ORDER BY configuration): 50,000,000 rows of int64 keys uniform over
[-2^62, 2^62) with 2% nulls (an independent validity stream), a stable descending argsort with nulls
last (am_argsort_ex2), then the gather of a float64 payload uniform in [-1e6, 1e6] (am_take).
One step = argsort + gather over the HBM-resident columns; rows/s = rows / step time.

Every leg reports, in the same JSON line:
  * its time and rows/s, and a parity check against the CPU oracle or Arrow C++;
  * `roofline`: the dominant kernel of the leg (the longest mean launch among the kernels with an
    algorithmic byte count), its algorithmic bytes per launch over its mean launch time, measured with
    HIP events on the stream the library launches on (am_prof_*), against the 8 TB/s HBM peak;
    `kernels_roofline` gives the same for every kernel of the leg; `traffic` comes from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary when that kernel is in it;
  * `cpu_baseline`: the reference's CPU path -- Arrow C++ (pyarrow) array_sort_indices + take, and
    the threaded hash group_by -- timed on this host with the threads it used stated, or NumPy's stable
    argsort for the float64 total order (the reference's own recipe for that order).
Legs: the headline ORDER BY; argsort_1M (1,000,000 full-range int64 keys, ascending); the float64
total-order ORDER BY; the Zipf(1.1) group-by over 100,000 keys and over 200 uniform keys; the batched
plan (2,048 lognormal batches); the served request mix at a fixed Poisson rate and saturated.

Inputs come from SplitMix64 streams (seeded), generated with NumPy, imported into HBM once before
timing. `--gpus N` with N > 1 exits: this build runs on one GPU.
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


def order_by_inputs_payload(n: int):
    return unit(splitmix64(0xA11CE503, n)) * 2e6 - 1e6


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


def uniform_200_keys(n: int) -> np.ndarray:
    table200 = splitmix64(0x6B0DE6, 200).view(np.int64)
    return table200[(splitmix64(0x21F1, n) % np.uint64(200)).astype(np.int64)]


def batched_inputs(nb: int, seed: int = 0xBA7C4):
    """Lognormal batch lengths (median 20,000 rows, sigma 1.78, capped at 4,000,000), full-range int64
    sort keys, group keys uniform over 1,000 int64 values, int64 values uniform in [-1e9, 1e9]."""
    z = np.sqrt(-2.0 * np.log(1.0 - unit(splitmix64(seed, nb)))) * np.cos(2 * np.pi * unit(splitmix64(seed + 1, nb)))
    lens = np.minimum(np.exp(np.log(20_000.0) + 1.78 * z), 4_000_000).astype(np.int64)
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    n = int(offs[-1])
    sort_keys = splitmix64(seed + 2, n).view(np.int64)
    key_table = splitmix64(seed + 3, 1000).view(np.int64)
    group_keys = key_table[(splitmix64(seed + 4, n) % np.uint64(1000)).astype(np.int64)]
    values = (splitmix64(seed + 5, n) % np.uint64(2_000_000_001)).astype(np.int64) - 1_000_000_000
    return offs, sort_keys, group_keys, values


# ==================================================================== measurement helpers

def timed(fn, steps: int, warmup: int):
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(steps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return ts


def profiled(fn, reps: int = 1):
    """Kernel times of `reps` calls of fn: {launch name: (launches, total ms)}, HIP events on the
    library's own streams."""
    am.prof_reset()
    am.prof_enable(True)
    for _ in range(reps):
        fn()
    am.prof_enable(False)
    return am.prof_report()


PMC_SUMMARY = os.path.join(ROOT, "profiles", "orderby_groupby_50M_pmc_traffic.txt")

# rocprofv3's kernel names (as written in the committed PMC summary) of the library's launch names.
ROCPROF_NAME = {
    "take": "take_kernel<u64, u32>",
    "sort_upsweep": "sort_upsweep<u64, 8>",
    "sort_msd_first": "msd_part<u64, 8, 16, false>",
    "sort_msd_hist": "lvl_hist<u64>",
    "sort_msd_scatter": "lvl_pass<u64, 20>",
    "sort_msd_local": "hyb_local<u64, 8, false>",
    "sort_first": "sort_first<u64, 8, 16>",
    "sort_scatter": "sort_mid<u64, 20, 0>",
    "sort_scatter_last": "sort_mid<u64, 20, 1>",
    "gb_key_scan": "gb_key_scan<8, 3, false>",
    "gb_build": "gb_build<false, 8, 7, true>",
    "gb_ids": "gb_ids",
    "gb_part_hist": "gb_part_hist<SumCountP, 4>",
    "gb_part_scatter": "gb_part_scatter<SumCountP, 4, 8, false>",
    "gb_sum": "gb_part_reduce<SumCountP, 8, false>",
    "order_by_partition": "ob_partition<8>",
    "order_by_gather": "ob_gather<8>",
}


def committed_traffic(launch_name: str):
    """HBM bytes per launch of a kernel from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (copy-kernel calibrated), or None."""
    rname = ROCPROF_NAME.get(launch_name)
    if not rname:
        return None
    try:
        with open(PMC_SUMMARY) as fh:
            for line in fh:
                if line.startswith("traffic " + rname + " "):
                    return int(line.rsplit(" ", 1)[1])
    except OSError:
        pass
    return None


def sort_alg_bytes(n: int, m: int):
    """Algorithmic HBM bytes per launch of the argsort kernels (n rows, m non-null): every input byte
    read once and every output byte written once, with a description."""
    vb = (n + 7) // 8
    return {
        "sort_upsweep": (8 * n + vb, "digit histograms: read key + validity; 8.125 B/row"),
        "sort_msd_first": (8 * n + vb + 12 * m + 4 * (n - m),
                           "top-digit partition: read key + validity, write key + row (null rows: index); ~20 B/row"),
        "sort_msd_hist": (8 * m, "one level's digit histograms of the large buckets: read key; 8 B/row"),
        "sort_msd_scatter": (24 * m, "one level's segmented partition: read key + row, write key + row; 24 B/row"),
        "sort_msd_local": (16 * m, "LDS bucket sorts: read key + row, write index; 16 B/row"),
        "sort_first": (8 * n + vb + 12 * m + 4 * (n - m), "LSD first pass: read key + validity, write key + row"),
        "sort_scatter": (24 * m, "LSD middle pass: read key + row, write key + row; 24 B/row"),
        "sort_scatter_last": (16 * m, "LSD last pass: read key + row, write index; 16 B/row"),
    }


def orderby_alg_bytes(n: int, m: int):
    d = sort_alg_bytes(n, m)
    d["take"] = (20 * n, "gather: read u32 index + f64 value, write f64; 20 B/row")
    return d


def group_by_alg_bytes(n: int):
    """Algorithmic bytes per launch of the group-by kernels over n rows of 8-byte keys and values."""
    return {
        "gb_key_scan": (8 * n, "key statistics (distinct-key estimate, span): read key; 8 B/row"),
        "gb_build": (12 * n, "hash insert: read key, write slot; 12 B/row"),
        "gb_ids": (8 * n, "slot to ascending id: read slot, write id; 8 B/row"),
        "gb_part_hist": (4 * n, "partitioned aggregate, bucket counts per tile: read id; 4 B/row"),
        "gb_part_scatter": (24 * n, "partitioned aggregate, records to their bucket: read id + value, write id + value; "
                                    "24 B/row"),
        "gb_sum": (12 * n, "int64 sum + count: read id + value (or the partitioned record); 12 B/row"),
        "gb_minmax": (12 * n, "min / max: read id + value; 12 B/row"),
        "fs_scan": (12 * n, "exact float sum, scan pass: read id + value; 12 B/row"),
        "fs_fixed": (12 * n, "exact float sum, fixed-point pass: read id + value; 12 B/row"),
    }


def merge_bucket_sorts(prof: dict) -> dict:
    """The two bucket sorts share the rows of a step: one entry, their summed time per step."""
    prof = dict(prof)
    if "sort_msd_group" in prof:
        c1, t1 = prof.pop("sort_msd_local", (0, 0.0))
        c2, t2 = prof.pop("sort_msd_group")
        prof["sort_msd_local"] = (max(c1, c2), t1 + t2)
    return prof


def roofline_of(prof: dict, alg: dict):
    """(roofline of the dominant kernel, {kernel: roofline row}) for the kernels of `prof` that have
    an algorithmic byte count in `alg`."""
    rows = {}
    for kname, (count, total) in prof.items():
        if kname in alg and count:
            mean_ms = total / count
            rows[kname] = {"mean_ms": round(mean_ms, 4), "launches": count, "alg_bytes": alg[kname][0],
                           "achieved_gbs": round(alg[kname][0] / (mean_ms * 1e-3) / 1e9, 1)}
    if not rows:
        return None, rows
    dom = max(rows, key=lambda k: rows[k]["mean_ms"])
    r = rows[dom]
    achieved = r["alg_bytes"] / (r["mean_ms"] * 1e-3) / 1e9
    traffic = committed_traffic(dom)
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_source": (f"{os.path.relpath(PMC_SUMMARY, ROOT)} ({ROCPROF_NAME[dom]}: FETCH_SIZE + "
                               "WRITE_SIZE, copy-kernel calibrated)") if traffic else None,
            "kernel": f"{dom} ({alg[dom][1]})", "bytes_per_launch": r["alg_bytes"],
            "mean_launch_ms": r["mean_ms"], "launches": r["launches"]}, rows


def sort_passes(prof: dict):
    """The passes an argsort executed, from its launches: the hybrid MSD sort (partition, levels,
    bucket sort) or the LSD onesweep (first pass + scatter passes)."""
    if "sort_msd_first" in prof or "sort_msd_local" in prof:
        return {"path": "msd", "digit_bits": 8, "top_digit_partitions": prof.get("sort_msd_first", (0, 0))[0],
                "levels": prof.get("sort_msd_scatter", (0, 0))[0],
                "bucket_sort_launches": prof.get("sort_msd_local", (0, 0))[0] + prof.get("sort_msd_group", (0, 0))[0]}
    return {"path": "lsd", "digit_bits": 8,
            "scatter_passes": prof.get("sort_first", (0, 0))[0] + prof.get("sort_scatter", (0, 0))[0]
            + prof.get("sort_scatter_last", (0, 0))[0]}


def per_step(prof: dict, steps: int) -> dict:
    return {k: (c // steps, t / steps) for k, (c, t) in prof.items()}


def kernels_ms(prof: dict) -> dict:
    return {k: round(t / max(c, 1), 4) for k, (c, t) in prof.items()}


def host_info():
    info = {"os_cpu_count": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    info["cpu_model"] = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        import pyarrow as pa
        info["pyarrow"] = pa.__version__
        info["pa_cpu_count_default"] = pa.cpu_count()
    except ImportError:
        info["pyarrow"] = None
    return info


def cpu_threads() -> int:
    """Threads for the threaded CPU baselines: the host's CPU share (OMP_NUM_THREADS when set, at most
    the CPUs this process may run on)."""
    share = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
    return max(1, min(share, len(os.sched_getaffinity(0))))


# ==================================================================== CPU baselines (Arrow C++)

def arrow_order_by(keys, valid, payload, descending=True):
    """Arrow C++ array_sort_indices (stable; descending; nulls last) + take, as the reference's CPU
    path. Returns (seconds, uint64 indices)."""
    import pyarrow as pa
    import pyarrow.compute as pc
    ka = pa.array(keys, mask=None if valid is None else ~valid)
    pa_payload = pa.array(payload) if payload is not None else None
    t0 = time.perf_counter()
    idx = pc.array_sort_indices(ka, order="descending" if descending else "ascending", null_placement="at_end")
    if pa_payload is not None:
        out = pc.take(pa_payload, idx)
        del out
    s = time.perf_counter() - t0
    return s, idx.to_numpy(zero_copy_only=False)


def arrow_group_by(keys, vals, float_vals=None):
    """Arrow C++ hash group_by (threaded): int64 sum, count_all, count; or float64 min, max, sum.
    Returns (seconds, table sorted by key)."""
    import pyarrow as pa
    import pyarrow.compute as pc
    if float_vals is None:
        t = pa.table({"k": keys, "v": vals})
        aggs = [("v", "sum"), ("v", "count", pc.CountOptions(mode="all")), ("v", "count")]
    else:
        t = pa.table({"k": keys, "v": float_vals})
        aggs = [("v", "min"), ("v", "max"), ("v", "sum")]
    t0 = time.perf_counter()
    r = t.group_by("k", use_threads=True).aggregate(aggs)
    s = time.perf_counter() - t0
    return s, r.sort_by("k")


def total_order_argsort_numpy(kf, valid, descending=True):
    """The float64 total order as the reference's own test computes it (python/tests/
    test_sort_options.py:59-78): NumPy's stable argsort of the sign-flip key; nulls last."""
    b = kf.view(np.int64)
    key = np.where(b < 0, b ^ 0x7FFFFFFFFFFFFFFF, b)
    if descending:
        key = ~key
    vi = np.nonzero(valid)[0]
    order = vi[np.argsort(key[vi], kind="stable")]
    return np.concatenate([order, np.nonzero(~valid)[0]])


# ==================================================================== legs

def order_by_leg(args, log):
    n = args.rows
    keys, valid, payload = order_by_inputs(n)
    m = int(valid.sum())
    k = am.MetalArray.from_numpy(keys, valid)
    p = am.MetalArray.from_numpy(payload)
    log(f"order-by inputs resident: {n} rows, {n - m} nulls")

    def step():
        idx = k.argsort(True, "at_end", "ieee")
        out = p.take(idx)
        del idx, out

    ts = timed(step, args.steps, args.warmup)
    ms, best = 1e3 * float(np.median(ts)), 1e3 * float(np.min(ts))
    log(f"order-by: median {ms:.3f} ms, best {best:.3f} ms")
    raw = profiled(step, 3)
    prof = merge_bucket_sorts(raw)
    log("kernels:", json.dumps(raw))
    alg = orderby_alg_bytes(n, m)
    roof, rows = roofline_of(prof, alg)
    b_step = sum(alg[kn][0] * prof[kn][0] // 3 for kn in rows)

    idx, _ = k.argsort(True, "at_end", "ieee").to_numpy()
    gathered, _ = p.take(am.MetalArray.from_numpy(idx)).to_numpy()
    res = {"ms": ms, "best": best, "m": m, "roofline": roof, "kernels_roofline": rows,
           "kernels_ms": kernels_ms(prof), "passes": sort_passes(per_step(raw, 3)),
           "b_alg_step_bytes": b_step, "parity_vs_oracle": None, "parity_vs_arrow": None, "cpu_baseline": None}
    if not args.skip_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_py as O   # the checker only; never the measured GPU path
        want = O.argsort(keys, valid, True, 0, 0)
        res["parity_vs_oracle"] = bool(np.array_equal(idx, want) and np.array_equal(
            gathered.view(np.uint64), payload[want].view(np.uint64)))
        del want
        ns = min(n, args.cpu_rows)
        s, aidx = arrow_order_by(keys[:ns], valid[:ns], payload[:ns])
        gi, _ = k.slice(0, ns).argsort(True, "at_end", "ieee").to_numpy()
        res["parity_vs_arrow"] = bool(np.array_equal(gi.astype(np.uint64), aidx))
        res["cpu_baseline"] = {
            "value": ns / s, "unit": "rows/s", "cores": 1, "kind": "reference",
            "sample": (f"Arrow C++ (pyarrow {__import__('pyarrow').__version__}) array_sort_indices (descending, "
                       f"nulls at end) + take of the float64 payload over the first {ns:,} rows of the same "
                       f"columns, one run, {s:.2f} s; Arrow's sort runs on one thread"),
        }
        log(f"order-by parity: oracle {res['parity_vs_oracle']}, arrow {res['parity_vs_arrow']}; "
            f"arrow {ns} rows in {s:.2f} s")
    return res, keys, valid, k, p


def plan_order_by_alg_bytes(n: int, m: int):
    d = sort_alg_bytes(n, m)
    d["order_by_partition"] = (20 * n, "column into the partition's layout: read position + 8 B value, write 8 B; "
                                       "20 B/row")
    d["order_by_gather"] = (20 * n, "column out of the partition: read position + 8 B value (one bucket's "
                                    "range), write 8 B; 20 B/row")
    return d


def plan_order_by_leg(args, log, keys, valid, payload, k, p):
    """ORDER BY through the plan engine (am_plan_run), the route DataFusion / Spark / Polars use: the
    same 50M-row table (key, payload) sorted by key descending, nulls last; both columns come back
    sorted. The one-key sort moves every column with its top-digit partition (no table-wide gather)."""
    n = len(keys)
    m = int(valid.sum())
    src = am.PlanSource("t", {"key": k, "payload": p})
    plan = json.dumps({"op": "sort", "by": [["key", True]], "input": {"op": "scan", "source": "t"}})

    def step():
        out = am.plan_run(plan, [src])
        del out

    ts = timed(step, args.steps, args.warmup)
    ms, best = 1e3 * float(np.median(ts)), 1e3 * float(np.min(ts))
    raw = profiled(step, 3)
    prof = merge_bucket_sorts(raw)
    alg = plan_order_by_alg_bytes(n, m)
    roof, rows = roofline_of(prof, alg)
    b_step = sum(alg[kn][0] * prof[kn][0] // 3 for kn in rows)
    res = {"workload": "plan_order_by_50M_int64_desc_nulls_last_table_of_key_and_f64_payload", "ms": ms,
           "best": best, "rows_per_s": n / (ms * 1e-3), "roofline": roof, "kernels_roofline": rows,
           "kernels_ms": kernels_ms(prof), "b_alg_step_bytes": b_step,
           "step_hbm_frac_of_b_alg": b_step / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "outputs": "both columns sorted (the two-call form returns the index and the payload only)",
           "parity_vs_oracle": None}
    log(f"plan order-by: median {ms:.3f} ms, best {best:.3f} ms; kernels {json.dumps(raw)}")
    if not args.skip_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_py as O   # the checker only
        out = am.plan_run(plan, [src])
        gk, gkv = out["key"].to_numpy()
        gp, _ = out["payload"].to_numpy()
        want = O.argsort(keys, valid, True, 0, 0)
        wv = valid[want]
        res["parity_vs_oracle"] = bool(np.array_equal(gkv if gkv is not None else np.ones(n, bool), wv)
                                       and np.array_equal(gk[wv], keys[want][wv])
                                       and np.array_equal(gp.view(np.uint64), payload[want].view(np.uint64)))
        del want, out
        log(f"plan order-by parity: {res['parity_vs_oracle']}")
    return res


def argsort_1m_leg(args, log):
    n1 = 1_000_000
    keys = splitmix64(0xA11CE501, n1).view(np.int64)
    k1 = am.MetalArray.from_numpy(keys)
    ts = timed(lambda: k1.argsort(), max(args.steps, 20), args.warmup)
    ms = 1e3 * float(np.median(ts))
    raw = profiled(lambda: k1.argsort(), 5)
    prof = merge_bucket_sorts(raw)
    roof, rows = roofline_of(prof, sort_alg_bytes(n1, n1))
    got, _ = k1.argsort().to_numpy()
    res = {"workload": "stable ascending argsort of 1,000,000 full-range int64 keys",
           "ms": round(ms, 4), "rows_per_s": n1 / (ms * 1e-3),
           "effective_gbs_12B_per_row": round(12 * n1 / (ms * 1e-3) / 1e9, 1),
           "passes": sort_passes(per_step(raw, 5)), "roofline": roof, "kernels_roofline": rows,
           "parity_vs_arrow": None, "cpu_baseline": None}
    if not args.skip_cpu:
        ts_cpu, aidx = [], None
        for _ in range(5):
            s, aidx = arrow_order_by(keys, None, None, descending=False)
            ts_cpu.append(s)
        s = float(np.median(ts_cpu))
        res["parity_vs_arrow"] = bool(np.array_equal(got.astype(np.uint64), aidx))
        res["cpu_baseline"] = {"value": n1 / s, "unit": "rows/s", "cores": 1, "kind": "reference",
                               "sample": f"Arrow C++ array_sort_indices (ascending) of the same 1,000,000 keys, "
                                         f"median of 5 runs, {s * 1e3:.1f} ms; one thread"}
    log(f"argsort 1M: {ms:.4f} ms, parity vs arrow {res['parity_vs_arrow']}")
    return res


def order_by_f64_leg(args, log, keys, valid, p):
    """The same ORDER BY with the keys as float64, 0.1% NaN and 0.1% -0.0 (seeded positions), total
    order (-NaN < -inf < ... < -0.0 < +0.0 < ... < +inf < +NaN), descending, nulls last, then the same
    gather."""
    n = len(keys)
    sel = unit(splitmix64(0xA11CE504, n))
    kf = keys.astype(np.float64)
    kf[sel < 0.001] = np.nan
    kf[(sel >= 0.001) & (sel < 0.002)] = -0.0
    del sel
    kfm = am.MetalArray.from_numpy(kf, valid)

    def step():
        idx = kfm.argsort(True, "at_end", "total")
        out = p.take(idx)
        del idx, out

    ts = timed(step, args.steps, args.warmup)
    ms = 1e3 * float(np.median(ts))
    raw = profiled(step, 3)
    prof = merge_bucket_sorts(raw)
    m = int(valid.sum())
    roof, rows = roofline_of(prof, orderby_alg_bytes(n, m))
    res = {"workload": "order_by_50M_float64_total_order_0.1pct_nan_0.1pct_negzero_desc_nulls_last_then_gather",
           "ms_per_step": round(ms, 4), "rows_per_s": n / (ms * 1e-3),
           "passes": sort_passes(per_step(raw, 3)), "roofline": roof, "kernels_roofline": rows,
           "kernels_ms": kernels_ms(prof), "parity_vs_oracle": None, "cpu_baseline": None}
    if not args.skip_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_py as O   # the checker only
        got, _ = kfm.argsort(True, "at_end", "total").to_numpy()
        res["parity_vs_oracle"] = bool(np.array_equal(got, O.argsort(kf, valid, True, 0, 1)))
        ns = min(n, args.cpu_rows_numpy)
        t0 = time.perf_counter()
        order = total_order_argsort_numpy(kf[:ns], valid[:ns])
        s = time.perf_counter() - t0
        gs, _ = kfm.slice(0, ns).argsort(True, "at_end", "total").to_numpy()
        res["parity_vs_numpy_recipe"] = bool(np.array_equal(gs, order.astype(np.uint32)))
        res["cpu_baseline"] = {"value": ns / s, "unit": "rows/s", "cores": 1, "kind": "reference",
                               "sample": f"NumPy stable argsort of the total-order key (the reference's recipe, "
                                         f"python/tests/test_sort_options.py:59-78), descending, nulls last, over "
                                         f"the first {ns:,} rows, {s:.2f} s; one thread (gather not included)"}
    log(f"order-by float64 total order: {ms:.3f} ms, parity {res['parity_vs_oracle']}")
    del kfm
    return res


def group_by_leg(args, log):
    n = args.rows
    gkeys, ivals, fvals = group_by_inputs(n)
    gk = am.MetalArray.from_numpy(gkeys)
    gi = am.MetalArray.from_numpy(ivals)
    gf = am.MetalArray.from_numpy(fvals)
    threads = cpu_threads()

    def int_step(keys_arr):
        def step():
            g = am.group_by([keys_arr])
            a = g.sum(gi)
            b = g.count_all()
            c = g.count(gi)
            del a, b, c, g
        return step

    def float_step():
        g = am.group_by([gk])
        a = g.min(gf)
        b = g.max(gf)
        c = g.sum(gf)
        del a, b, c, g

    def gpu_result(keys_arr):
        g = am.group_by([keys_arr])
        return (g.key_arrays()[0].to_numpy()[0], g.sum(gi).to_numpy()[0], g.count_all().to_numpy()[0],
                g.count(gi).to_numpy()[0], int(g.group_count))

    def parity_and_cpu(keys_np, keys_arr, label):
        out = {}
        if args.skip_cpu:
            return out
        import pyarrow as pa
        pa.set_cpu_count(threads)
        arrow_group_by(keys_np[:1_000_000], ivals[:1_000_000])        # warm the thread pool
        s, ref = arrow_group_by(keys_np, ivals)
        k_, s_, ca_, c_, K = gpu_result(keys_arr)
        out["parity_vs_arrow"] = bool(
            K == ref.num_rows and np.array_equal(k_, ref.column("k").to_numpy())
            and np.array_equal(s_, ref.column(1).to_numpy()) and np.array_equal(ca_, ref.column(2).to_numpy())
            and np.array_equal(c_, ref.column(3).to_numpy()))
        out["cpu_baseline"] = {"value": n / s, "unit": "rows/s", "cores": threads, "kind": "reference",
                               "sample": f"Arrow C++ (pyarrow) Table.group_by('k').aggregate(sum, count_all, count) "
                                         f"over the same {n:,} rows, threaded (pa.set_cpu_count({threads})), "
                                         f"one run, {s:.2f} s"}
        log(f"group-by {label}: arrow {s:.2f} s on {threads} threads, parity {out['parity_vs_arrow']}")
        return out

    step = int_step(gk)
    tg = timed(step, args.steps, args.warmup)
    tf = timed(float_step, max(3, args.steps // 2), 1)
    raw = profiled(step, 3)
    fraw = profiled(float_step, 1)
    log("group-by kernels:", json.dumps(raw))
    alg = group_by_alg_bytes(n)
    roof, rows = roofline_of(raw, alg)
    froof, frows = roofline_of(fraw, alg)
    gms, fms = 1e3 * float(np.median(tg)), 1e3 * float(np.median(tf))
    g = am.group_by([gk])
    res = {
        "workload": "group_by_50M_int64_zipf1.1_100k_keys",
        "groups": int(g.group_count),
        "ms_per_step": round(gms, 4),
        "rows_per_s": n / (gms * 1e-3),
        "step": "am_group_by_keys + am_group_agg_ex sum, count_all, count (int64 values)",
        "b_alg_bytes": 16 * n,
        "step_hbm_frac_of_b_alg": (16 * n) / (gms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "effective_gbs_16B_per_row": round(16 * n / (gms * 1e-3) / 1e9, 1),
        "roofline": roof, "kernels_roofline": rows, "kernels_ms": kernels_ms(per_step(raw, 3)),
        "float_step_ms": round(fms, 4),
        "float_step": "am_group_by_keys + float64 min, max, correctly rounded sum",
        "float_step_roofline": froof, "float_kernels_roofline": frows,
    }
    del g
    res.update(parity_and_cpu(gkeys, gk, "zipf"))
    if not args.skip_cpu:
        s, _ = arrow_group_by(gkeys, None, fvals)
        res["float_step_cpu_baseline"] = {
            "value": n / s, "unit": "rows/s", "cores": threads, "kind": "reference",
            "sample": f"Arrow C++ Table.group_by('k').aggregate(min, max, sum) of the float64 column over the "
                      f"same {n:,} rows, threaded, {s:.2f} s (Arrow's float sum is not correctly rounded)"}
    log(f"group-by: {gms:.3f} ms, float step {fms:.3f} ms, groups {res['groups']}")
    del gf, gk
    # The second run: 200 distinct keys, uniform.
    k200 = uniform_200_keys(n)
    gk200 = am.MetalArray.from_numpy(k200)
    step200 = int_step(gk200)
    t200 = timed(step200, args.steps, args.warmup)
    ms200 = 1e3 * float(np.median(t200))
    raw200 = profiled(step200, 3)
    roof200, rows200 = roofline_of(raw200, alg)
    u = {"workload": "group_by_50M_int64_uniform_200_keys", "ms_per_step": round(ms200, 4),
         "rows_per_s": n / (ms200 * 1e-3),
         "step": "the same step over 50M int64 keys uniform over 200 distinct values",
         "step_hbm_frac_of_b_alg": (16 * n) / (ms200 * 1e-3) / 1e9 / HBM_PEAK_GBS,
         "roofline": roof200, "kernels_roofline": rows200, "kernels_ms": kernels_ms(per_step(raw200, 3))}
    u.update(parity_and_cpu(k200, gk200, "200 keys"))
    res["uniform_200_keys"] = u
    log(f"group-by, 200 uniform keys: {ms200:.3f} ms")
    del gk200
    # Variant R: the same Zipf(1.1) rows with the key = its rank 1..100,000 (a small span: the range path,
    # occupancy + scan + rank, no hashing).
    kr = (zipf_ranks(n, 100_000, 1.1, 0x21F0).astype(np.int64) + 1)
    gkr = am.MetalArray.from_numpy(kr)
    stepr = int_step(gkr)
    tr = timed(stepr, args.steps, args.warmup)
    msr = 1e3 * float(np.median(tr))
    rawr = profiled(stepr, 3)
    algr = dict(alg)
    algr["gb_range_occ"] = (8 * n, "range path: occupancy of the span; read key; 8 B/row")
    algr["gb_ids"] = (12 * n, "row to ascending id: read key, write id; 12 B/row")
    roofr, rowsr = roofline_of(rawr, algr)
    vr = {"workload": "group_by_50M_int64_zipf1.1_keys_are_ranks_1_to_100k", "ms_per_step": round(msr, 4),
          "rows_per_s": n / (msr * 1e-3),
          "step": "the same step; keys = Zipf rank (span 100,000: the range path)",
          "step_hbm_frac_of_b_alg": (16 * n) / (msr * 1e-3) / 1e9 / HBM_PEAK_GBS,
          "roofline": roofr, "kernels_roofline": rowsr, "kernels_ms": kernels_ms(per_step(rawr, 3))}
    vr.update(parity_and_cpu(kr, gkr, "variant R"))
    res["variant_r_rank_keys"] = vr
    log(f"group-by, variant R (rank keys): {msr:.3f} ms")
    del gkr, gi
    return res


def batched_leg(args, log):
    """All batches in one graph-replayed kernel sequence (am_batched_run) against the per-batch loop of
    the plain entry points (am_slice, am_argsort_ex2, am_group_by_keys, keys, sum, count_all)."""
    offs, sk, gk, gv = batched_inputs(args.batches)
    n = int(offs[-1])
    nb = len(offs) - 1
    msk, mgk, mgv = (am.MetalArray.from_numpy(x) for x in (sk, gk, gv))
    log(f"batched inputs resident: {nb} batches, {n} rows")
    plan = am.Batched(offs, hash_capacity=2048)
    t0 = time.perf_counter()
    cfg = plan.tune(msk, mgk, mgv)
    tune_s = time.perf_counter() - t0
    log(f"batched tuned in {tune_s:.2f} s: {cfg}")
    step = lambda: plan.run(msk, mgk, mgv, use_graph=True)
    tg = timed(step, args.steps, args.warmup)
    te = timed(lambda: plan.run(msk, mgk, mgv, use_graph=False), max(3, args.steps // 2), 1)
    gms, ems = 1e3 * float(np.median(tg)), 1e3 * float(np.median(te))
    bprof = merge_bucket_sorts(profiled(lambda: plan.run(msk, mgk, mgv, use_graph=False), 1))
    log(f"batched: graph {gms:.3f} ms, eager {ems:.3f} ms; kernels {json.dumps(bprof)}")
    config = plan.config

    def loop(batches):
        outs = []
        for b in batches:
            a, e = int(offs[b]), int(offs[b + 1])
            s_k, s_g, s_v = msk.slice(a, e - a), mgk.slice(a, e - a), mgv.slice(a, e - a)
            idx = s_k.argsort(False, "at_end", "ieee")
            g = am.group_by([s_g])
            outs.append((idx, g.key_arrays()[0], g.sum(s_v), g.count_all()))
            del g
        return outs

    loop(range(min(nb, 16)))   # warm the per-call path
    t0 = time.perf_counter()
    outs = loop(range(nb))
    loop_ms = 1e3 * (time.perf_counter() - t0)
    log(f"per-batch loop: {loop_ms:.1f} ms")
    # The batched results must equal the per-call results batch for batch (checked on a sample).
    idx_all, _ = plan.result("indices").to_numpy()
    keys_all, _ = plan.result("keys").to_numpy()
    sums_all, _ = plan.result("sums").to_numpy()
    cnts_all, _ = plan.result("counts").to_numpy()
    go, _ = plan.result("group_offsets").to_numpy()
    same = True
    sample = sorted(set(np.linspace(0, nb - 1, 128).astype(int).tolist()))
    for b in sample:
        a, e = int(offs[b]), int(offs[b + 1])
        lo, hi = int(go[b]), int(go[b + 1])
        i_, k_, s_, c_ = (x.to_numpy()[0] for x in outs[b])
        same &= bool(np.array_equal(idx_all[a:e], i_) and np.array_equal(keys_all[lo:hi], k_) and
                     np.array_equal(sums_all[lo:hi], s_) and np.array_equal(cnts_all[lo:hi], c_))
    del outs
    # Algorithmic bytes of the path that ran. MSD sort of 64-bit keys: map (read key, write mapped key:
    # 16 B/row), the first level over every row (histogram 8 + partition 24), the bucket sort (16); the
    # group-by (read key + value: 16). Batches above 524,288 rows take further levels, so this is a
    # lower bound there. LSD sort: the onesweep model of 188 B/row, plus 16 for the group-by.
    msd = config["sort_path"] == "msd"
    b_row = (16 + 8 + 24 + 16 + 16) if msd else (188 + 16)
    b_alg = b_row * n
    alg = {"seg_msd_map": (16 * n, "map keys: read key, write mapped key; 16 B/row"),
           "bg_agg": (16 * n, "group-by per chunk: read key + value; 16 B/row"),
           "sort_msd_local": (16 * n, "bucket sorts: read key + row, write index; 16 B/row")}
    roof, rows = roofline_of(bprof, alg)
    res = {
        "workload": f"batched_{nb}_lognormal_median20k_int64_argsort_plus_groupby_1000_keys",
        "batches": nb,
        "rows": n,
        "ms_per_step": round(gms, 4),
        "rows_per_s": n / (gms * 1e-3),
        "eager_ms_per_step": round(ems, 4),
        "per_batch_loop_ms": round(loop_ms, 3),
        "per_batch_loop_rows_per_s": n / (loop_ms * 1e-3),
        "speedup_vs_per_batch_loop": loop_ms / gms,
        "step": "am_batched_run (graph replay): stable ascending argsort + group-by sum/count_all per batch",
        "tuned": config,
        "digit_bits_note": "8, fixed: the MSD levels and the LSD passes rank 256 bins per tile (one per thread of "
                           "a 256-thread workgroup); an 11-bit digit needs 2,048-bin look-backs",
        "tune_s": round(tune_s, 3),
        "b_alg_bytes_per_row": b_row,
        "b_alg_bytes": b_alg,
        "step_hbm_frac_of_b_alg": b_alg / (gms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "roofline": roof, "kernels_roofline": rows,
        "same_as_per_batch_loop": same,
        "batches_compared": len(sample),
        "kernels_ms": kernels_ms(bprof),
        "cpu_baseline": None,
    }
    if not args.skip_cpu:
        import pyarrow as pa
        import pyarrow.compute as pc
        pa.set_cpu_count(cpu_threads())
        cs = sample[::4]
        rows_cpu, t_cpu = 0, 0.0
        for b in cs:
            a, e = int(offs[b]), int(offs[b + 1])
            ka, ga, va = pa.array(sk[a:e]), pa.array(gk[a:e]), pa.array(gv[a:e])
            t0 = time.perf_counter()
            pc.array_sort_indices(ka)
            pa.table({"k": ga, "v": va}).group_by("k").aggregate([("v", "sum"), ("v", "count",
                                                                  pc.CountOptions(mode="all"))])
            t_cpu += time.perf_counter() - t0
            rows_cpu += e - a
        res["cpu_baseline"] = {"value": rows_cpu / t_cpu, "unit": "rows/s", "cores": cpu_threads(),
                               "kind": "reference",
                               "sample": f"Arrow C++ per batch: array_sort_indices + group_by(sum, count_all), "
                                         f"{len(cs)} batches spread over the lengths ({rows_cpu:,} rows), "
                                         f"{t_cpu:.2f} s; sort one thread, group_by threaded"}
    del sk, gk, gv
    return res


def serving_inputs(n: int = 50_000_000, seed: int = 0x5E4E):
    """Resident columns of the served mix: group keys Zipf(1.1) over 200, 10,000 and 1,000,000
    distinct int64 values; int64 values uniform in [-1e9, 1e9]; full-range int64 sort keys; the two
    unique keys uniform over 10^6 and 10^3 (10M rows)."""
    gkeys = {}
    for support in (200, 10_000, 1_000_000):
        table = splitmix64(seed + support, support).view(np.int64)
        gkeys[support] = table[zipf_ranks(n, support, 1.1, seed + 7 * support)]
    gvals = (splitmix64(seed + 1, n) % np.uint64(2_000_000_001)).astype(np.int64) - 1_000_000_000
    skeys = splitmix64(seed + 2, n).view(np.int64)
    nu = 10_000_000
    ua = (splitmix64(seed + 3, nu) % np.uint64(1_000_000)).astype(np.int64)
    ub = (splitmix64(seed + 4, nu) % np.uint64(1_000)).astype(np.int64)
    return gkeys, gvals, skeys, ua, ub


def serving_leg(args, log):
    """Poisson arrivals of a mixed request stream on resident columns, served by am_server (slots,
    admission by footprint) and, on the same schedule, by one thread making the plain calls; then the
    same comparison saturated (every request queued at once)."""
    import threading
    gkeys, gvals, skeys, ua, ub = serving_inputs()
    dev_g = {k: am.MetalArray.from_numpy(v) for k, v in gkeys.items()}
    dev_v = am.MetalArray.from_numpy(gvals)
    dev_s = am.MetalArray.from_numpy(skeys)
    dev_a, dev_b = am.MetalArray.from_numpy(ua), am.MetalArray.from_numpy(ub)
    n, nu = len(gvals), len(ua)
    del gkeys, gvals, skeys, ua, ub
    rng = np.random.default_rng(0x5E4F)
    count, rate = args.requests, args.rate
    reqs = []
    for _ in range(count):
        u = rng.random()
        if u < 0.6:
            rows = int(np.exp(rng.uniform(np.log(1e6), np.log(n))))
            support = int(rng.choice([200, 10_000, 1_000_000]))
            off = int(rng.integers(0, n - rows + 1))
            reqs.append(("group_by", [dev_g[support].slice(off, rows)], dev_v.slice(off, rows), rows))
        elif u < 0.9:
            rows = int(np.exp(rng.uniform(np.log(250_000), np.log(n))))
            off = int(rng.integers(0, n - rows + 1))
            reqs.append(("argsort", [dev_s.slice(off, rows)], None, rows))
        else:
            rows = int(np.exp(rng.uniform(np.log(1e6), np.log(nu))))
            off = int(rng.integers(0, nu - rows + 1))
            reqs.append(("unique", [dev_a.slice(off, rows), dev_b.slice(off, rows)], None, rows))

    def plain_call(kind, cols, values):
        if kind == "argsort":
            return [cols[0].argsort()]
        if kind == "group_by":
            g = am.group_by(cols)
            return g.key_arrays() + [g.sum(values), g.count_all()]
        rows = am.unique_rows(cols)
        return [rows] + [c.take(rows) for c in cols]

    for kind, cols, values, _ in reqs[:20]:   # warm both paths (pool, code objects)
        plain_call(kind, cols, values)

    def run(requests, arrivals):
        """(plain latencies, plain span, sampled plain results, served latencies, served span, sampled
        served results, server stats) for the requests arriving at `arrivals` (seconds)."""
        total = len(requests)
        lat_plain, plain_out = [], []
        t0 = time.perf_counter()
        for i, ((kind, cols, values, _), at) in enumerate(zip(requests, arrivals)):
            wait = at - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            res = plain_call(kind, cols, values)
            lat_plain.append(time.perf_counter() - t0 - at)
            plain_out.append(res if i % 7 == 0 else None)   # a sample is kept for the comparison
            del res
        plain_span = time.perf_counter() - t0
        srv = am.Server(slots=args.slots, budget_bytes=args.budget_gb << 30)
        for t in [srv.submit(k, c, v) for k, c, v, _ in requests[:20]]:
            srv.wait(t)
        tickets = [None] * total
        lateness = [0.0] * total
        start = [0.0]

        def submitter():
            for i, ((kind, cols, values, _), at) in enumerate(zip(requests, arrivals)):
                wait = at - (time.perf_counter() - start[0])
                if wait > 0:
                    time.sleep(wait)
                lateness[i] = max(0.0, time.perf_counter() - start[0] - at)
                tickets[i] = srv.submit(kind, cols, values)

        start[0] = time.perf_counter()
        th = threading.Thread(target=submitter)
        th.start()
        lat_served, served_out = [], []
        for i in range(total):
            while tickets[i] is None:
                time.sleep(0.0001)
            res, times = srv.wait(tickets[i])
            lat_served.append(lateness[i] + times[2] * 1e-3)
            served_out.append(res if i % 7 == 0 else None)
            del res
        th.join()
        served_span = time.perf_counter() - start[0]
        st = srv.stats
        srv.close()
        return lat_plain, plain_span, plain_out, lat_served, served_span, served_out, st

    def identical(served_out, plain_out):
        ok = True
        for a_list, b_list in zip(served_out[::7], plain_out[::7]):
            for a, b in zip(a_list, b_list):
                va, ma = a.to_numpy()
                vb, mb = b.to_numpy()
                ok &= bool(np.array_equal(va, vb)) and ((ma is None) == (mb is None))
        return ok

    pct = lambda xs, q: 1e3 * float(np.percentile(np.asarray(xs), q))
    total_rows = sum(r[3] for r in reqs)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
    lp, ps, po, ls, ss, so, st = run(reqs, arrivals)
    res = {
        "workload": f"served_mix_{count}_requests_poisson_{rate:g}_per_s",
        "requests": count,
        "rate_per_s": rate,
        "mix": "60% group-by sum/count (1M-50M rows, Zipf(1.1) over 200 / 10k / 1M keys), 30% int64 argsort "
               "(250k-50M rows), 10% unique over two int64 keys (1M-10M rows, 10^6 x 10^3)",
        "total_rows": total_rows,
        "offered_rows_per_s": total_rows / float(arrivals[-1]),
        "served": {"p50_ms": pct(ls, 50), "p99_ms": pct(ls, 99), "mean_ms": 1e3 * float(np.mean(ls)),
                   "rows_per_s": total_rows / ss, "span_s": ss, "slots": args.slots,
                   "budget_bytes": args.budget_gb << 30, "peak_admitted_bytes": st["peak_admitted_bytes"],
                   "peak_running": st["peak_running"]},
        "plain": {"p50_ms": pct(lp, 50), "p99_ms": pct(lp, 99), "mean_ms": 1e3 * float(np.mean(lp)),
                  "rows_per_s": total_rows / ps, "span_s": ps},
        "served_equals_plain": identical(so, po),
        "compared_requests": len(so[::7]),
    }
    log(f"serving at {rate}/s: served p50 {res['served']['p50_ms']:.2f} p99 {res['served']['p99_ms']:.2f} ms "
        f"{res['served']['rows_per_s']:.3e} rows/s; plain p50 {res['plain']['p50_ms']:.2f} p99 "
        f"{res['plain']['p99_ms']:.2f} ms {res['plain']['rows_per_s']:.3e} rows/s")
    # Saturated: every request is queued at time 0, so throughput is what each form can sustain.
    sat = reqs[:args.saturated_requests]
    sat_rows = sum(r[3] for r in sat)
    lp, ps, po, ls, ss, so, st = run(sat, np.zeros(len(sat)))
    res["saturated"] = {
        "requests": len(sat), "total_rows": sat_rows,
        "served": {"rows_per_s": sat_rows / ss, "span_s": ss, "p50_ms": pct(ls, 50), "p99_ms": pct(ls, 99),
                   "peak_running": st["peak_running"]},
        "plain": {"rows_per_s": sat_rows / ps, "span_s": ps, "p50_ms": pct(lp, 50), "p99_ms": pct(lp, 99)},
        "served_over_plain_throughput": ps / ss,
        "served_equals_plain": identical(so, po),
    }
    log(f"serving saturated: served {sat_rows / ss:.3e} rows/s, plain {sat_rows / ps:.3e} rows/s")
    return res


# ==================================================================== main

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows", type=int, default=50_000_000)
    ap.add_argument("--cpu-rows", type=int, default=50_000_000,
                    help="rows of the Arrow C++ ORDER BY baseline (a prefix of the same columns)")
    ap.add_argument("--cpu-rows-numpy", type=int, default=10_000_000,
                    help="rows of the NumPy float64 total-order baseline")
    ap.add_argument("--skip-cpu", action="store_true", help="omit the CPU baselines and parity checks")
    ap.add_argument("--skip-groupby", action="store_true")
    ap.add_argument("--skip-f64", action="store_true", help="omit the float64 total-order ORDER BY leg")
    ap.add_argument("--skip-batched", action="store_true")
    ap.add_argument("--batches", type=int, default=2048)
    ap.add_argument("--skip-serving", action="store_true")
    ap.add_argument("--requests", type=int, default=1500)
    ap.add_argument("--saturated-requests", type=int, default=600)
    ap.add_argument("--rate", type=float, default=700.0, help="Poisson arrival rate of the served mix (1/s)")
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--budget-gb", type=int, default=8)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(nothing here launches on a second GPU)", file=sys.stderr)
        sys.exit(2)
    n = args.rows
    log = lambda *a: print("[bench]", *a, file=sys.stderr, flush=True)
    log("device:", am.device_name())

    ob, keys, valid, k, p = order_by_leg(args, log)
    pob = plan_order_by_leg(args, log, keys, valid, order_by_inputs_payload(n), k, p)
    a1 = argsort_1m_leg(args, log)
    f64_res = None if args.skip_f64 else order_by_f64_leg(args, log, keys, valid, p)
    del keys, valid, k, p
    gb_res = None if args.skip_groupby else group_by_leg(args, log)
    bt_res = None if args.skip_batched else batched_leg(args, log)
    sv_res = None if args.skip_serving else serving_leg(args, log)

    ms = ob["ms"]
    result = {
        "metric": METRIC,
        "value": n / (ms * 1e-3),
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "best_ms_per_step": round(ob["best"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (SplitMix64 streams; keys uniform over [-2^62, 2^62), 2% nulls, float64 payload)",
        "config": {"workload": "order_by_50M_int64_desc_nulls_last_then_gather_f64", "rows": n,
                   "null_fraction": round(1 - ob["m"] / n, 5), "descending": True, "null_placement": "at_end",
                   "float_order": "ieee", "passes": ob["passes"]},
        "step_hbm_frac_of_b_alg": ob["b_alg_step_bytes"] / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "b_alg_step_bytes": ob["b_alg_step_bytes"],
        "effective_gbs_12B_per_row": round(12 * n / (ms * 1e-3) / 1e9, 1),
        "roofline": ob["roofline"],
        "kernels_roofline": ob["kernels_roofline"],
        "kernels_ms": ob["kernels_ms"],
        "parity_vs_oracle": ob["parity_vs_oracle"],
        "parity_vs_arrow": ob["parity_vs_arrow"],
        "cpu_baseline": ob["cpu_baseline"],
        "host": host_info(),
        "plan_order_by": pob,
        "argsort_1M": a1,
        "order_by_f64_total": f64_res,
        "group_by": gb_res,
        "batched": bt_res,
        "serving": sv_res,
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()

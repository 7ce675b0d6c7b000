"""Benchmark of the argsort / group-by path on one MI355X.

Workload (the metric's ORDER BY configuration): 50,000,000 rows of int64 keys uniform over
[-2^62, 2^62) with 2% nulls (an independent validity stream), a stable descending argsort with nulls
last (am_argsort_ex2), then the gather of a float64 payload uniform in [-1e6, 1e6] (am_take).
One step = argsort + gather over the HBM-resident columns; rows/s = rows / step time.

Also measured and reported beside it (same JSON line):
  * argsort_1M: the plain ascending argsort of 1,000,000 full-range int64 keys;
  * order_by_f64_total: the same ORDER BY with the keys as float64, 0.1% NaN and 0.1% -0.0, under
    the total order, with its own parity check against the oracle;
  * group_by_50M_zipf_100k: am_group_by_keys over 50M int64 keys drawn Zipf(1.1) from 100,000
    distinct values, then int64 sum, count_all and count (am_group_agg_ex ops 0, 1, 2), and the float64
    min / max / sum of a second value column; the same integer step over 200 uniform distinct keys;
  * roofline: the dominant kernel of the step (the longest mean launch), its algorithmic bytes per
    launch (orderby_alg_bytes) over its mean launch time, measured with HIP events on the library's
    own stream (am_prof_*); kernels_roofline gives the same for every kernel of the step;
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


def batched_leg(args, log):
    """All batches in one graph-replayed kernel sequence (am_batched_run) against the per-batch loop of
    the plain entry points (am_slice, am_argsort_ex2, am_group_by_keys, keys, sum, count_all)."""
    offs, sk, gk, gv = batched_inputs(args.batches)
    n = int(offs[-1])
    nb = len(offs) - 1
    msk, mgk, mgv = (am.MetalArray.from_numpy(x) for x in (sk, gk, gv))
    del sk, gk, gv
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
    am.prof_reset()
    am.prof_enable(True)
    plan.run(msk, mgk, mgv, use_graph=False)
    am.prof_enable(False)
    bprof = am.prof_report()
    log(f"batched: graph {gms:.3f} ms, eager {ems:.3f} ms; kernels {json.dumps(bprof)}")

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
    b_alg = int((8 + (8 + 12) + 6 * 24 + 16 + 16) * n)   # sort 188 B/row + group-by 16 B/row
    return {
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
        "tuned": plan.config,
        "tune_s": round(tune_s, 3),
        "b_alg_bytes": b_alg,
        "hbm_frac_of_b_alg": b_alg / (gms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "same_as_per_batch_loop": same,
        "batches_compared": len(sample),
        "kernels_ms": {kname: round(v[1] / max(v[0], 1), 4) for kname, v in bprof.items()},
    }


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
    """Poisson arrivals of a mixed request stream on resident columns, served by am_server (4 slots,
    admission by footprint) and, on the same schedule, by one thread making the plain calls."""
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
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
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
    total_rows = sum(r[3] for r in reqs)

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

    # Plain: one caller, in arrival order, each call synchronous.
    lat_plain = []
    plain_out = []
    t0 = time.perf_counter()
    for i, ((kind, cols, values, _), at) in enumerate(zip(reqs, arrivals)):
        wait = at - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        res = plain_call(kind, cols, values)
        lat_plain.append(time.perf_counter() - t0 - at)
        plain_out.append(res if i % 7 == 0 else None)   # a sample is kept for the comparison
        del res
    plain_span = time.perf_counter() - t0

    srv = am.Server(slots=args.slots, budget_bytes=args.budget_gb << 30)
    warm = [srv.submit(k, c, v) for k, c, v, _ in reqs[:20]]
    for t in warm:
        srv.wait(t)
    tickets = [None] * count
    lateness = [0.0] * count

    def submitter():
        for i, ((kind, cols, values, _), at) in enumerate(zip(reqs, arrivals)):
            wait = at - (time.perf_counter() - t1)
            if wait > 0:
                time.sleep(wait)
            lateness[i] = max(0.0, time.perf_counter() - t1 - at)
            tickets[i] = srv.submit(kind, cols, values)

    t1 = time.perf_counter()
    th = threading.Thread(target=submitter)
    th.start()
    lat_served, served_out = [], []
    for i in range(count):
        while tickets[i] is None:
            time.sleep(0.0001)
        res, times = srv.wait(tickets[i])
        lat_served.append(lateness[i] + times[2] * 1e-3)
        served_out.append(res if i % 7 == 0 else None)
        del res
    th.join()
    served_span = time.perf_counter() - t1
    st = srv.stats
    srv.close()
    # Every served result must equal the plain call's.
    identical = True
    for a_list, b_list in zip(served_out[::7], plain_out[::7]):
        for a, b in zip(a_list, b_list):
            va, ma = a.to_numpy()
            vb, mb = b.to_numpy()
            identical &= bool(np.array_equal(va, vb)) and ((ma is None) == (mb is None))
    pct = lambda xs, q: 1e3 * float(np.percentile(np.asarray(xs), q))
    res = {
        "workload": f"served_mix_{count}_requests_poisson_{rate:g}_per_s",
        "requests": count,
        "rate_per_s": rate,
        "mix": "60% group-by sum/count (1M-50M rows, Zipf(1.1) over 200 / 10k / 1M keys), 30% int64 argsort "
               "(250k-50M rows), 10% unique over two int64 keys (1M-10M rows, 10^6 x 10^3)",
        "total_rows": total_rows,
        "offered_rows_per_s": total_rows / float(arrivals[-1]),
        "served": {"p50_ms": pct(lat_served, 50), "p99_ms": pct(lat_served, 99), "mean_ms": 1e3 * float(np.mean(lat_served)),
                   "rows_per_s": total_rows / served_span, "span_s": served_span,
                   "slots": args.slots, "budget_bytes": args.budget_gb << 30,
                   "peak_admitted_bytes": st["peak_admitted_bytes"], "peak_running": st["peak_running"]},
        "plain": {"p50_ms": pct(lat_plain, 50), "p99_ms": pct(lat_plain, 99), "mean_ms": 1e3 * float(np.mean(lat_plain)),
                  "rows_per_s": total_rows / plain_span, "span_s": plain_span},
        "served_equals_plain": identical,
        "compared_requests": len(served_out[::7]),
    }
    log(f"serving: served p50 {res['served']['p50_ms']:.2f} p99 {res['served']['p99_ms']:.2f} ms; "
        f"plain p50 {res['plain']['p50_ms']:.2f} p99 {res['plain']['p99_ms']:.2f} ms; stats {st}")
    return res


def timed(fn, steps: int, warmup: int):
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(steps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return ts


PMC_SUMMARY = os.path.join(ROOT, "profiles", "orderby_groupby_50M_pmc_traffic.txt")


def committed_traffic(prefix: str):
    """HBM bytes per launch of the first kernel whose name starts with `prefix`, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-kernel calibrated), or (None, None)."""
    try:
        with open(PMC_SUMMARY) as fh:
            for line in fh:
                if line.startswith("traffic " + prefix):
                    name, value = line[len("traffic "):].rsplit(" ", 1)
                    return int(value), name
    except OSError:
        pass
    return None, None


def orderby_alg_bytes(n: int, m: int):
    """Algorithmic HBM bytes per launch of each kernel of the ORDER BY step (n rows, m non-null): every
    input byte read once and every output byte written once, with a description."""
    vb = (n + 7) // 8
    return {
        "take": (20 * n, "gather: read u32 index + f64 value, write f64; 20 B/row"),
        "sort_upsweep": (8 * n + vb, "digit histograms: read key + validity; 8.125 B/row"),
        "sort_msd_first": (8 * n + vb + 12 * m + 4 * (n - m),
                           "top-digit partition: read key + validity, write key + row (null rows: index); ~20 B/row"),
        "sort_msd_hist": (8 * m, "one level's digit histograms of the large buckets: read key; 8 B/row"),
        "sort_msd_scatter": (24 * m, "one level's segmented partition: read key + row, write key + row; 24 B/row"),
        "sort_msd_local": (16 * m, "LDS bucket sort: read key + row, write index; 16 B/row"),
        "sort_msd_group": (16 * m, "LDS sort of groups of small buckets: read key + row, write index; 16 B/row"),
        "sort_first": (8 * n + vb + 12 * m + 4 * (n - m), "LSD first pass: read key + validity, write key + row"),
        "sort_scatter": (24 * m, "LSD middle pass: read key + row, write key + row; 24 B/row"),
        "sort_scatter_last": (16 * m, "LSD last pass: read key + row, write index; 16 B/row"),
    }


# rocprofv3's kernel names (as shortened in the committed PMC summary) of the profiled launch names.
ROCPROF_NAME = {
    "take": "take_kernel<u64, u32>",
    "sort_upsweep": "sort_upsweep<u64, 8>",
    "sort_msd_first": "msd_part<u64, 8, 16>",
    "sort_msd_hist": "lvl_hist<u64>",
    "sort_msd_scatter": "lvl_pass<u64, 20>",
    "sort_msd_local": "hyb_local<u64, 8>",
    "sort_first": "sort_first<u64, 8, 16>",
    "sort_scatter": "sort_mid<u64, 20, 0>",
    "sort_scatter_last": "sort_mid<u64, 20, 1>",
}


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
    ap.add_argument("--skip-f64", action="store_true", help="omit the float64 total-order ORDER BY leg")
    ap.add_argument("--skip-batched", action="store_true")
    ap.add_argument("--batches", type=int, default=2048)
    ap.add_argument("--skip-serving", action="store_true")
    ap.add_argument("--requests", type=int, default=1500)
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
    if "sort_msd_local" in prof and "sort_msd_group" in prof:
        # The two bucket sorts share the rows: one entry, their summed time per step.
        (c1, t1), (c2, t2) = prof.pop("sort_msd_local"), prof.pop("sort_msd_group")
        prof["sort_msd_local"] = (max(c1, c2), t1 + t2)
    # Roofline of the dominant kernel (the longest mean launch of the step), and of every other kernel
    # of the step with an algorithmic byte count.
    alg = orderby_alg_bytes(n, m)
    rows_roof = {}
    for kname, (count, total) in prof.items():
        if kname in alg and count:
            mean_ms = total / count
            rows_roof[kname] = {"mean_ms": round(mean_ms, 4), "alg_bytes": alg[kname][0],
                                "achieved_gbs": round(alg[kname][0] / (mean_ms * 1e-3) / 1e9, 1)}
    dom = max(rows_roof, key=lambda kname: rows_roof[kname]["mean_ms"]) if rows_roof else None
    dom_ms, dom_n = kernel_ms(dom, prof) if dom else (None, 0)
    bytes_per_launch = alg[dom][0] if dom else None
    traffic, traffic_kernel = committed_traffic(ROCPROF_NAME[dom]) if dom in ROCPROF_NAME else (None, None)
    achieved = bytes_per_launch / (dom_ms * 1e-3) / 1e9 if dom_ms else None

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

    # ---------------------------------------------------------------- ORDER BY, float64 keys, total order
    # The same keys as float64 with 0.1% NaN and 0.1% -0.0 (seeded positions), total order (-NaN < -inf
    # < ... < -0.0 < +0.0 < ... < +inf < +NaN), descending, nulls last, then the same gather.
    f64_res = None
    if not args.skip_f64:
        sel = unit(splitmix64(0xA11CE504, n))
        kf = keys.astype(np.float64)
        kf[sel < 0.001] = np.nan
        kf[(sel >= 0.001) & (sel < 0.002)] = -0.0
        kfm = am.MetalArray.from_numpy(kf, valid)

        def order_by_f64_step():
            idx = kfm.argsort(True, "at_end", "total")
            out = p.take(idx)
            del idx, out

        tf64 = timed(order_by_f64_step, args.steps, args.warmup)
        f64_ms = 1e3 * float(np.median(tf64))
        f64_parity = None
        if not args.skip_cpu:
            sys.path.insert(0, os.path.join(ROOT, "tests"))
            import oracle_py as O   # the checker only
            got, _ = kfm.argsort(True, "at_end", "total").to_numpy()
            f64_parity = bool(np.array_equal(got, O.argsort(kf, valid, True, 0, 1)))
        am.prof_reset()
        am.prof_enable(True)
        order_by_f64_step()
        am.prof_enable(False)
        f64_prof = am.prof_report()
        f64_res = {"workload": "order_by_50M_float64_total_order_0.1pct_nan_0.1pct_negzero_desc_nulls_last_then_gather",
                   "ms_per_step": round(f64_ms, 4), "rows_per_s": n / (f64_ms * 1e-3),
                   "parity_vs_oracle": f64_parity,
                   "kernels_ms": {kname: round(v[1] / max(v[0], 1), 4) for kname, v in f64_prof.items()}}
        log(f"order-by float64 total order: {f64_ms:.3f} ms, parity {f64_parity}")
        del kfm, kf, sel

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
        del g, gk, gf
        # The second run: 200 distinct keys, uniform.
        table200 = splitmix64(0x6B0DE6, 200).view(np.int64)
        gk = am.MetalArray.from_numpy(table200[(splitmix64(0x21F1, n) % np.uint64(200)).astype(np.int64)])
        t200 = timed(gb_step, args.steps, args.warmup)
        ms200 = 1e3 * float(np.median(t200))
        gb_res["uniform_200_keys"] = {"ms_per_step": round(ms200, 4), "rows_per_s": n / (ms200 * 1e-3),
                                      "step": "the same step over 50M int64 keys uniform over 200 distinct values"}
        log(f"group-by, 200 uniform keys: {ms200:.3f} ms")
        del gk, gi

    # ---------------------------------------------------------------- batched (2,048 record batches)
    bt_res = None
    if not args.skip_batched:
        bt_res = batched_leg(args, log)

    sv_res = None
    if not args.skip_serving:
        sv_res = serving_leg(args, log)

    # Algorithmic bytes of one step: the sort kernels the profiled steps launched (3 steps), and the gather.
    b_sort = sum(alg[kname][0] * prof[kname][0] // 3 for kname in rows_roof if kname != "take")
    b_take = 20 * n
    hybrid = "sort_msd_local" in prof or "sort_msd_group" in prof
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
                   "float_order": "ieee", "digit_bits": 8,
                   "sort": ("hybrid: top-digit partition, segmented second-digit partition, LDS bucket sort"
                            if hybrid else "LSD onesweep, 8 passes")},
        "step_hbm_frac_of_b_alg": (b_sort + b_take) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "b_alg_bytes": {"sort": b_sort, "gather": b_take},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": traffic,
                     "traffic_source": (f"{os.path.relpath(PMC_SUMMARY, ROOT)} ({traffic_kernel}, FETCH_SIZE + "
                                        "WRITE_SIZE scaled by a 1 GiB copy)") if traffic else None,
                     "kernel": f"{dom} ({alg[dom][1]})" if dom else None,
                     "bytes_per_launch": bytes_per_launch, "mean_launch_ms": dom_ms,
                     "launches": dom_n},
        "kernels_roofline": rows_roof,
        "kernels_ms": {kname: round(v[1] / max(v[0], 1), 4) for kname, v in prof.items()},
        "parity_vs_oracle": parity,
        "cpu_baseline": cpu,
        "argsort_1M": {"ms": round(ms1, 4), "rows_per_s": 1e6 / (ms1 * 1e-3),
                       "workload": "stable ascending argsort of 1,000,000 full-range int64 keys"},
        "order_by_f64_total": f64_res,
        "group_by": gb_res,
        "batched": bt_res,
        "serving": sv_res,
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()

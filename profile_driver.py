"""Fixed workload for hardware-counter profiles (rocprofv3 --pmc), run under the profiler.

Runs, in order: three ORDER BY steps over 50M int64 keys with 2% nulls (argsort descending, nulls
last, then the float64 gather), three ORDER BY steps of the same (key, payload) table through the plan
engine (am_plan_run: both columns come back sorted), two group-by steps over 50M Zipf(1.1) keys from
100k distinct values (keys, int64 sum, count_all, count), and three device-to-device copies of 1 GiB,
which calibrate the FETCH_SIZE / WRITE_SIZE counters against a kernel of known traffic (read 1 GiB,
write 1 GiB).

With --workload orderby it stops after the ORDER BY steps; with --workload batched it instead runs the
batched plan (2,048 lognormal batches, ~191M rows: tune, then three graph replays of the per-batch
argsort + group-by sum / count_all).

With --summarize FETCH_DIR WRITE_DIR it runs nothing on the GPU: it reads the counter CSVs of the two
rocprofv3 --pmc runs (FETCH_SIZE, then WRITE_SIZE) and prints the per-kernel traffic summary that
bench.py reads (copy-kernel calibrated, beside each kernel's algorithmic bytes). With --kernel-stats
STATS_DIR it prints the kernel table of a rocprofv3 --kernel-trace --stats run.

Usage: python3 profile_driver.py [--rows N] [--workload orderby_groupby|orderby|batched]
       python3 profile_driver.py --summarize FETCH_DIR WRITE_DIR
       python3 profile_driver.py --kernel-stats STATS_DIR
"""
from __future__ import annotations

import argparse
import ctypes
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import arrowmetal_amd as am  # noqa: E402
import bench  # noqa: E402


def device_copy(nbytes: int, times: int) -> None:
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    for _ in range(times):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0   # device to device
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def _short(kernel_name: str) -> str:
    """A kernel's name without namespace, return type and parameters, with u64 / u32 for the integer types."""
    name = kernel_name.replace("am::(anonymous namespace)::", "").replace("void ", "").split("(")[0]
    return name.replace("unsigned long long", "u64").replace("unsigned int", "u32")


def _counters(d: str, counter: str):
    """{kernel name: [per-dispatch counter values in KiB]} from a rocprofv3 counter CSV directory."""
    import csv
    import glob
    out = {}
    for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as fh:
            for r in csv.DictReader(fh):
                if r["Counter_Name"] != counter:
                    continue
                name = _short(r["Kernel_Name"])
                out.setdefault(name, {}).setdefault(r["Dispatch_Id"], 0.0)
                out[name][r["Dispatch_Id"]] += float(r["Counter_Value"])
    return {k: list(v.values()) for k, v in out.items()}


def _large(values):
    """The dispatch values of at least 1/8 of the largest one."""
    if not values:
        return values
    top = max(values)
    return [v for v in values if v >= top / 8]


def summarize(fetch_dir: str, write_dir: str, n: int = 50_000_000) -> str:
    fetch, write = _counters(fetch_dir, "FETCH_SIZE"), _counters(write_dir, "WRITE_SIZE")
    copy = "__amd_rocclr_copyBuffer"
    gib = float(1 << 30)
    # The three 1 GiB copies: the largest copy dispatches of each run.
    f_copy = sorted(fetch.get(copy, []))[-3:]
    w_copy = sorted(write.get(copy, []))[-3:]
    if len(f_copy) < 3 or len(w_copy) < 3:
        raise SystemExit("profile_driver: the calibration copies are missing from the counter files")
    f_factor = gib / (sum(f_copy) / 3 * 1024)
    w_factor = gib / (sum(w_copy) / 3 * 1024)
    keys, valid, _ = bench.order_by_inputs(n)
    m = int(valid.sum())
    del keys, valid
    alg = {}
    for table in (bench.orderby_alg_bytes(n, m), bench.plan_order_by_alg_bytes(n, m), bench.group_by_alg_bytes(n)):
        for launch, (b, _) in table.items():
            if launch in bench.ROCPROF_NAME:
                alg[bench.ROCPROF_NAME[launch]] = b
    lines = [
        "# rocprofv3 PMC: HBM traffic per dispatch, ORDER BY 50M (two calls and through the plan) + group-by 50M",
        "",
        "commands (from the repo root, R = repo root; each counter in its own run):",
        "  cd /tmp && timeout -s KILL 240 rocprofv3 --pmc FETCH_SIZE --output-format csv -d $R/gpurun_out/pmc_fetch "
        "-o run -- python3 $R/profile_driver.py",
        "  cd /tmp && timeout -s KILL 240 rocprofv3 --pmc WRITE_SIZE --output-format csv -d $R/gpurun_out/pmc_write "
        "-o run -- python3 $R/profile_driver.py",
        "  python3 profile_driver.py --summarize gpurun_out/pmc_fetch gpurun_out/pmc_write",
        "",
        "workload: profile_driver.py (3 ORDER BY steps over 50M int64 keys with 2% nulls, 3 ORDER BY steps of the "
        "same table through am_plan_run, 2 group-by steps over 50M Zipf(1.1) keys from 100k values, then three "
        "1 GiB device-to-device copies).",
        "",
        f"calibration ({copy}, the three 1 GiB device-to-device copies): FETCH_SIZE reads "
        f"{1 / f_factor:.3f}x the copied bytes, WRITE_SIZE {1 / w_factor:.3f}x;",
        f"so bytes = FETCH_SIZE KiB * 1024 * {f_factor:.3f} + WRITE_SIZE KiB * 1024 * {w_factor:.3f}.",
        "per kernel: the mean over its dispatches of at least 1/8 of its largest (the 50M-row launches; the group-by's",
        "sorts of its 100k distinct keys launch the same sort kernels on 500x fewer rows); disp counts those dispatches.",
        "",
        f"{'kernel':40s} {'disp':>5s} {'FETCH_KiB':>12s} {'WRITE_KiB':>12s} {'traffic_MB':>11s} {'alg_MB':>9s} {'ratio':>6s}",
    ]
    traffic = {}
    for name in fetch:
        # The group-by's sorts of its distinct keys launch the same sort kernels on far fewer rows: a
        # kernel's traffic is the mean over its large dispatches (at least 1/8 of its largest).
        fv, wv = _large(fetch[name]), _large(write.get(name, []))
        if not fv or not wv:
            continue
        fk, wk = sum(fv) / len(fv), sum(wv) / len(wv)
        t = fk * 1024 * f_factor + wk * 1024 * w_factor
        if t < 1e6:
            continue
        traffic[name] = t
        a = alg.get(name)
        lines.append(f"{name[:40]:40s} {len(fv):5d} {fk:12.0f} {wk:12.0f} {t / 1e6:11.1f} "
                     f"{(a / 1e6 if a else float('nan')):9.1f} {(t / a if a else float('nan')):6.2f}")
    lines += ["", "machine-readable: traffic_bytes_per_launch"]
    lines += [f"traffic {k} {int(v)}" for k, v in traffic.items()]
    return "\n".join(lines) + "\n"


def kernel_stats(d: str, top: int = 60) -> str:
    """The kernel table of a rocprofv3 --kernel-trace --stats run (its *kernel_stats.csv), by total time."""
    import csv
    import glob
    rows = []
    for path in glob.glob(os.path.join(d, "**", "*kernel_stats.csv"), recursive=True):
        with open(path) as fh:
            for r in csv.DictReader(fh):
                name = _short(r["Name"])
                rows.append((float(r["TotalDurationNs"]), name, int(r["Calls"]), float(r["AverageNs"]),
                             float(r["MinNs"]), float(r["MaxNs"])))
    rows.sort(reverse=True)
    out = [f"{'kernel':45s} {'calls':>6s} {'avg_us':>10s} {'min_us':>10s} {'max_us':>10s} {'total_ms':>10s}"]
    for tot, name, calls, avg, lo, hi in rows[:top]:
        out.append(f"{name[:45]:45s} {calls:6d} {avg / 1e3:10.1f} {lo / 1e3:10.1f} {hi / 1e3:10.1f} {tot / 1e6:10.3f}")
    return "\n".join(out) + "\n"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rows", type=int, default=50_000_000)
    ap.add_argument("--workload", default="orderby_groupby", choices=["orderby_groupby", "orderby", "batched"])
    ap.add_argument("--summarize", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"))
    ap.add_argument("--kernel-stats", metavar="STATS_DIR")
    args = ap.parse_args()
    if args.kernel_stats:
        sys.stdout.write(kernel_stats(args.kernel_stats))
        return
    if args.summarize:
        sys.stdout.write(summarize(*args.summarize, n=args.rows))
        return
    if args.workload == "batched":
        offs, sk, gk, gv = bench.batched_inputs(2048)
        msk, mgk, mgv = (am.MetalArray.from_numpy(x) for x in (sk, gk, gv))
        plan = am.Batched(offs, hash_capacity=2048)
        print("tuned", plan.tune(msk, mgk, mgv), flush=True)
        for i in range(3):
            plan.run(msk, mgk, mgv, use_graph=True)
            print(f"batched step {i} done", flush=True)
        return
    n = args.rows
    keys, valid, payload = bench.order_by_inputs(n)
    k = am.MetalArray.from_numpy(keys, valid)
    p = am.MetalArray.from_numpy(payload)
    for i in range(3):
        idx = k.argsort(True, "at_end", "ieee")
        out = p.take(idx)
        del idx, out
        print(f"order-by step {i} done", flush=True)
    import json
    src = am.PlanSource("t", {"key": k, "payload": p})
    plan = json.dumps({"op": "sort", "by": [["key", True]], "input": {"op": "scan", "source": "t"}})
    for i in range(3):
        out = am.plan_run(plan, [src])
        del out
        print(f"plan order-by step {i} done", flush=True)
    del src, k, p, keys, valid, payload
    if args.workload == "orderby":
        return
    gkeys, ivals, _ = bench.group_by_inputs(n)
    gk = am.MetalArray.from_numpy(gkeys)
    gi = am.MetalArray.from_numpy(ivals)
    for i in range(2):
        g = am.group_by([gk])
        a, b, c = g.sum(gi), g.count_all(), g.count(gi)
        del a, b, c, g
        print(f"group-by step {i} done", flush=True)
    device_copy(1 << 30, 3)
    print("copies done", flush=True)


if __name__ == "__main__":
    main()

"""Fixed workload for hardware-counter profiles (rocprofv3 --pmc), run under the profiler.

Runs, in order: three ORDER BY steps over 50M int64 keys with 2% nulls (argsort descending, nulls
last, then the float64 gather), two group-by steps over 50M Zipf(1.1) keys from 100k distinct values
(keys, int64 sum, count_all, count), and three device-to-device copies of 1 GiB, which calibrate the
FETCH_SIZE / WRITE_SIZE counters against a kernel of known traffic (read 1 GiB, write 1 GiB).

With --workload orderby it stops after the ORDER BY steps; with --workload batched it instead runs the batched plan (2,048 lognormal batches, ~191M rows: tune,
then three graph replays of the per-batch argsort + group-by sum / count_all).

Usage: python3 profile_driver.py [--rows N] [--workload orderby_groupby|orderby|batched]
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rows", type=int, default=50_000_000)
    ap.add_argument("--workload", default="orderby_groupby", choices=["orderby_groupby", "orderby", "batched"])
    args = ap.parse_args()
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
    del k, p, keys, valid, payload
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

"""GPU parity of the batched plan (am_batched_*): every batch's argsort and group-by sum / count_all
must equal the oracle run on that batch alone (the per-call semantics of am_argsort_ex2 and
am_group_by_keys + am_group_agg_ex), for the graph-replayed and the eager sequence, after tuning,
and when batches overflow the LDS hash capacity and are regrouped by the general path."""
import numpy as np
import pytest

import arrowmetal_amd as am
import oracle_py as O

pytestmark = pytest.mark.gpu


def lognormal_offsets(nb, median, sigma, cap, seed, zero_every=0):
    rng = np.random.default_rng(seed)
    lens = np.minimum(np.exp(rng.normal(np.log(median), sigma, nb)), cap).astype(np.int64)
    if zero_every:
        lens[::zero_every] = 0
    return np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)


def check_sort(idx, vals, offs, desc, forder, label):
    for b in range(len(offs) - 1):
        a, e = int(offs[b]), int(offs[b + 1])
        want = O.argsort(vals[a:e], None, desc, 0, forder)
        got = idx[a:e]
        if not np.array_equal(got, want):
            k = int(np.nonzero(got != want)[0][0])
            raise AssertionError(f"{label}: batch {b} (rows {a}..{e}) position {k}: got {got[k]} want {want[k]}")


def check_groups(plan, keys, vals, offs, label):
    gk, _ = plan.result("keys").to_numpy()
    gs, _ = plan.result("sums").to_numpy()
    gc, _ = plan.result("counts").to_numpy()
    go, _ = plan.result("group_offsets").to_numpy()
    assert go[0] == 0 and go[-1] == len(gk) == len(gs) == len(gc), label
    for b in range(len(offs) - 1):
        a, e = int(offs[b]), int(offs[b + 1])
        ids, K, rep = O.group_keys([keys[a:e]], [None])
        lo, hi = int(go[b]), int(go[b + 1])
        assert hi - lo == K, f"{label}: batch {b}: {hi - lo} groups, want {K}"
        if K == 0:
            continue
        np.testing.assert_array_equal(gk[lo:hi], keys[a:e][rep], err_msg=f"{label}: batch {b} keys")
        sb, _ = O.group_agg(ids, K, vals[a:e], 0)
        cb, _ = O.group_agg(ids, K, None, 1)
        np.testing.assert_array_equal(gs[lo:hi].view(np.uint64), sb, err_msg=f"{label}: batch {b} sums")
        np.testing.assert_array_equal(gc[lo:hi].view(np.uint64), cb, err_msg=f"{label}: batch {b} counts")


@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("desc", [False, True])
def test_batched_int64_matches_per_batch_oracle(use_graph, desc):
    offs = lognormal_offsets(300, 2000, 1.78, 200_000, seed=4 + desc, zero_every=37)
    n = int(offs[-1])
    print(f"batched int64 n={n} batches={len(offs) - 1} graph={use_graph} desc={desc}", flush=True)
    sk = O.splitmix64(0xA11CE501, n).view(np.int64)
    rng = np.random.default_rng(7)
    gk = rng.integers(0, 1000, n).astype(np.int64) * 1_000_003 - 400_000_000
    gv = rng.integers(-10**9, 10**9 + 1, n).astype(np.int64)
    plan = am.Batched(offs, hash_capacity=2048)
    msk, mgk, mgv = (am.MetalArray.from_numpy(x) for x in (sk, gk, gv))
    for rep in range(2):   # the second run replays the captured graph
        plan.run(msk, mgk, mgv, descending=desc, use_graph=use_graph)
        idx, _ = plan.result("indices").to_numpy()
        check_sort(idx, sk, offs, desc, 0, f"run {rep}")
        check_groups(plan, gk, gv, offs, f"run {rep}")
    assert plan.config["regrouped_batches"] == 0


@pytest.mark.parametrize("dtype,forder", [(np.float64, 0), (np.float64, 1), (np.float64, 2), (np.int32, 0),
                                          (np.uint16, 0), (np.float32, 0)])
def test_batched_sort_key_types(dtype, forder):
    offs = lognormal_offsets(64, 3000, 1.2, 50_000, seed=11)
    n = int(offs[-1])
    print(f"batched sort {np.dtype(dtype).name} forder={forder} n={n}", flush=True)
    rng = np.random.default_rng(12)
    if dtype in (np.float64, np.float32):
        v = np.round(rng.normal(0, 100, n)).astype(dtype)
        v[rng.random(n) < 0.01] = np.nan
        v[rng.random(n) < 0.01] = -0.0
        bits = v.view(np.uint64 if dtype == np.float64 else np.uint32)
        bits[rng.random(n) < 0.003] |= (1 << (63 if dtype == np.float64 else 31))   # negative NaNs too
        v[rng.random(n) < 0.005] = np.inf
        v[rng.random(n) < 0.005] = -np.inf
    else:
        info = np.iinfo(dtype)
        v = rng.integers(info.min, int(info.max) + 1, n).astype(dtype)
    plan = am.Batched(offs)
    mv = am.MetalArray.from_numpy(v)
    for desc in (False, True):
        plan.run(mv, descending=desc, float_order=("ieee", "total", "nan_largest")[forder])
        idx, _ = plan.result("indices").to_numpy()
        check_sort(idx, v, offs, desc, forder, f"{np.dtype(dtype).name} desc={desc}")


def test_batched_msd_levels_large_batches():
    """64-bit sort keys take the MSD radix sort: batches above 2048 rows are partitioned by digit 7, and
    batches above ~524K rows a second time by digit 6, before the LDS bucket sorts (groups of small
    sibling buckets, larger buckets one per workgroup, LSD passes for runs of equal keys). Each batch
    must match the oracle, with full-range keys, keys of few distinct values, and float64 keys."""
    lens = [1, 0, 2047, 2048, 2049, 5000, 31_000, 200_000, 700_000, 1_300_000, 40, 600_000, 3]
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    n = int(offs[-1])
    rng = np.random.default_rng(21)
    sk = O.splitmix64(0xB16B00B5, n).view(np.int64).copy()
    a, e = int(offs[8]), int(offs[9])   # the 700K-row batch: 3000 distinct keys
    sk[a:e] = rng.integers(-(2**62), 2**62, 3000)[rng.integers(0, 3000, e - a)]
    a, e = int(offs[11]), int(offs[12])   # the 600K-row batch: keys differing only in their top 3 bytes
    sk[a:e] = (rng.integers(0, 1 << 24, e - a) << 40) - (1 << 62)
    plan = am.Batched(offs)
    mk = am.MetalArray.from_numpy(sk)
    for desc in (False, True):
        print(f"batched msd levels int64 n={n} desc={desc}", flush=True)
        plan.run(mk, descending=desc, use_graph=True)
        assert plan.config["sort_path"] == "msd"
        idx, _ = plan.result("indices").to_numpy()
        check_sort(idx, sk, offs, desc, 0, f"int64 desc={desc}")
    v = np.round(rng.normal(0, 1e6, n), 1)
    v[rng.random(n) < 0.002] = np.nan
    v[rng.random(n) < 0.002] = -0.0
    mv = am.MetalArray.from_numpy(v)
    print(f"batched msd levels float64 n={n}", flush=True)
    plan.run(mv, descending=True, float_order="total", use_graph=True)
    assert plan.config["sort_path"] == "msd"
    idx, _ = plan.result("indices").to_numpy()
    check_sort(idx, v, offs, True, 1, "float64 total desc")


def test_batched_msd_list_overflow_falls_back(monkeypatch):
    """A batched run whose MSD bucket-sort list runs out of room (forced here by capping the list at 64
    entries when the plan is made) is sorted again by the segmented LSD passes, which the plan keeps
    from then on; every batch still matches the oracle."""
    monkeypatch.setenv("ARROWMETAL_MSD_LIST_CAP", "64")
    offs = lognormal_offsets(200, 20_000, 1.0, 200_000, seed=31)
    plan = am.Batched(offs)
    monkeypatch.delenv("ARROWMETAL_MSD_LIST_CAP")
    n = int(offs[-1])
    sk = O.splitmix64(0x0FF1CE, n).view(np.int64)
    mk = am.MetalArray.from_numpy(sk)
    for rep in range(2):
        print(f"batched msd list overflow n={n} run {rep}", flush=True)
        plan.run(mk, use_graph=True)
        assert plan.config["sort_path"] == "lsd"
        idx, _ = plan.result("indices").to_numpy()
        check_sort(idx, sk, offs, False, 0, f"run {rep}")


def test_batched_group_chunk_boundaries():
    """Batches at the group-by's 32,768-row chunk boundary: a batch of at most one chunk stores its
    LDS table as the batch's table, a larger one merges its chunks' tables into it."""
    lens = [32_767, 32_768, 32_769, 65_536, 65_537, 1, 0, 98_304, 5]
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    n = int(offs[-1])
    print(f"batched chunk boundaries n={n}", flush=True)
    sk = O.splitmix64(0xB0B0, n).view(np.int64)
    rng = np.random.default_rng(11)
    gk = rng.integers(0, 900, n).astype(np.int64) * 7_919 - 3_000_000
    # The smallest int64 (the LDS tables' empty-slot word, kept in a slot of its own) in a batch of one
    # chunk and in one of two.
    gk[offs[0]:offs[1]:5] = np.iinfo(np.int64).min
    gk[offs[2]:offs[3]:3] = np.iinfo(np.int64).min
    gv = rng.integers(-10**9, 10**9 + 1, n).astype(np.int64)
    plan = am.Batched(offs, hash_capacity=2048)
    msk, mgk, mgv = (am.MetalArray.from_numpy(x) for x in (sk, gk, gv))
    for rep in range(2):
        plan.run(msk, mgk, mgv, descending=False, use_graph=False)
        check_groups(plan, gk, gv, offs, f"chunk boundaries run {rep}")
    assert plan.config["regrouped_batches"] == 0


def test_batched_overflow_regroups():
    """Batches with more distinct keys than the hash capacity are regrouped by the general path."""
    offs = np.array([0, 100, 5100, 5100, 9000, 40_000, 40_050], dtype=np.int64)
    n = int(offs[-1])
    rng = np.random.default_rng(3)
    gk = rng.integers(-10**12, 10**12, n).astype(np.int64)    # nearly all distinct
    gk[:100] = 5
    gk[40_000:] = rng.integers(0, 3, 50)
    gv = rng.integers(-100, 100, n).astype(np.int64)
    plan = am.Batched(offs, hash_capacity=256)
    plan.run(None, am.MetalArray.from_numpy(gk), am.MetalArray.from_numpy(gv))
    assert plan.config["regrouped_batches"] == 3
    check_groups(plan, gk, gv, offs, "overflow")
    # The replay after a regroup captures again and still agrees.
    plan.run(None, am.MetalArray.from_numpy(gk), am.MetalArray.from_numpy(gv))
    check_groups(plan, gk, gv, offs, "overflow replay")


def test_batched_narrow_group_keys_and_tune():
    offs = lognormal_offsets(200, 5000, 1.0, 100_000, seed=21, zero_every=50)
    n = int(offs[-1])
    rng = np.random.default_rng(22)
    gk = rng.integers(-500, 500, n).astype(np.int32)
    gv = rng.integers(-(2**31), 2**31, n).astype(np.int64)
    sk = rng.integers(-(2**62), 2**62, n).astype(np.int64)
    plan = am.Batched(offs, hash_capacity=256)
    msk, mgk, mgv = (am.MetalArray.from_numpy(x) for x in (sk, gk, gv))
    cfg = plan.tune(msk, mgk, mgv)
    print("tuned", cfg, flush=True)
    assert cfg["tile_items"] in (8, 16, 20) and cfg["hash_capacity"] >= 1024
    plan.run(msk, mgk, mgv)
    idx, _ = plan.result("indices").to_numpy()
    check_sort(idx, sk, offs, False, 0, "tuned")
    check_groups(plan, gk, gv, offs, "tuned")
    assert plan.result("keys").format == "i"


def test_batched_errors():
    offs = np.array([0, 10, 20], dtype=np.int64)
    plan = am.Batched(offs)
    short = am.MetalArray.from_numpy(np.arange(5, dtype=np.int64))
    with pytest.raises(am.ArrowMetalError, match="batches' total"):
        plan.run(short)
    nulls = am.MetalArray.from_numpy(np.arange(20, dtype=np.int64), np.arange(20) % 3 != 0)
    with pytest.raises(am.ArrowMetalError, match="nulls"):
        plan.run(nulls)
    with pytest.raises(am.ArrowMetalError, match="no sort keys"):
        plan.result("indices")
    with pytest.raises(am.ArrowMetalError, match="offsets must not decrease"):
        am.Batched(np.array([0, 10, 5], dtype=np.int64))

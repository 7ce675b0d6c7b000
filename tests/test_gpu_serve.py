"""The request server (am_server_*): every request's results are bit-identical to the same request run
alone through the plain entry points, whatever else runs beside it; admission by memory footprint
and slots; errors surface at the waiting caller. Small requests are also checked against the oracle."""
import numpy as np
import pytest

import arrowmetal_amd as am
import oracle_py as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def resident():
    rng = np.random.default_rng(55)
    n = 2_000_000
    table = O.splitmix64(0x5E7, 1000).view(np.int64)
    w = 1.0 / np.arange(1, 1001) ** 1.1
    ranks = np.minimum(np.searchsorted(np.cumsum(w) / w.sum(), rng.random(n)), 999)
    gk = table[ranks]
    gv = rng.integers(-10**9, 10**9 + 1, n).astype(np.int64)
    sk = O.splitmix64(0x5E8, n).view(np.int64)
    sv = rng.random(n) >= 0.02
    ua = rng.integers(0, 10**6, n).astype(np.int64)
    ub = rng.integers(0, 10**3, n).astype(np.int64)
    host = dict(gk=gk, gv=gv, sk=sk, sv=sv, ua=ua, ub=ub)
    dev = dict(gk=am.MetalArray.from_numpy(gk), gv=am.MetalArray.from_numpy(gv),
               sk=am.MetalArray.from_numpy(sk, sv), ua=am.MetalArray.from_numpy(ua),
               ub=am.MetalArray.from_numpy(ub))
    return host, dev


def make_mix(dev, count, seed):
    rng = np.random.default_rng(seed)
    n = len(dev["gk"])
    reqs = []
    for i in range(count):
        kind = rng.choice(["group_by", "argsort", "unique"], p=[0.6, 0.3, 0.1])
        length = int(rng.integers(1000, n // 2))
        off = int(rng.integers(0, n - length))
        if kind == "group_by":
            reqs.append(("group_by", [dev["gk"].slice(off, length)], dev["gv"].slice(off, length), {}))
        elif kind == "argsort":
            opts = {"descending": bool(rng.integers(0, 2)), "null_placement": ("at_end", "at_start")[rng.integers(0, 2)]}
            reqs.append(("argsort", [dev["sk"].slice(off, length)], None, opts))
        else:
            reqs.append(("unique", [dev["ua"].slice(off, length), dev["ub"].slice(off, length)], None, {}))
    return reqs


def plain(kind, cols, values, opts):
    """The same request through the plain per-call entry points."""
    if kind == "argsort":
        return [cols[0].argsort(opts.get("descending", False), opts.get("null_placement", "at_end"))]
    if kind == "group_by":
        g = am.group_by(cols)
        return g.key_arrays() + [g.sum(values), g.count_all()]
    rows = am.unique_rows(cols)
    return [rows] + [c.take(rows) for c in cols]


def same(a, b):
    va, ma = a.to_numpy()
    vb, mb = b.to_numpy()
    return a.format == b.format and np.array_equal(va.view(np.uint8), vb.view(np.uint8)) and (
        (ma is None and mb is None) or (ma is not None and mb is not None and np.array_equal(ma, mb)))


@pytest.mark.parametrize("slots,budget", [(4, 64 << 30), (2, 1 << 30), (4, 1)])
def test_results_same_alone_and_in_the_mix(resident, slots, budget):
    _, dev = resident
    reqs = make_mix(dev, 40, seed=slots * 7 + (budget & 0xFF))
    srv = am.Server(slots=slots, budget_bytes=budget)
    tickets = [srv.submit(k, c, v, **o) for k, c, v, o in reqs]
    outs = [srv.wait(t) for t in tickets]
    st = srv.stats
    srv.close()
    print("server stats", st, flush=True)
    assert st["submitted"] == st["completed"] == len(reqs) and st["failed"] == 0
    assert st["peak_running"] <= slots
    if budget == 1:
        assert st["peak_running"] == 1   # every request exceeds the budget: each runs alone
    for i, ((k, c, v, o), (res, times)) in enumerate(zip(reqs, outs)):
        want = plain(k, c, v, o)
        assert len(res) == len(want), f"request {i} ({k})"
        for j, (a, b) in enumerate(zip(res, want)):
            assert same(a, b), f"request {i} ({k}) result {j} differs from the plain call"
        assert times[2] >= times[1] >= 0 and times[0] >= 0


def test_served_results_match_oracle(resident):
    host, dev = resident
    srv = am.Server(slots=3)
    off, length = 12_345, 150_000
    t1 = srv.submit("argsort", [dev["sk"].slice(off, length)], descending=True)
    t2 = srv.submit("group_by", [dev["gk"].slice(off, length)], dev["gv"].slice(off, length))
    t3 = srv.submit("unique", [dev["ua"].slice(off, length), dev["ub"].slice(off, length)])
    (idx,), _ = srv.wait(t1)
    gres, _ = srv.wait(t2)
    ures, _ = srv.wait(t3)
    srv.close()
    sk, sv = host["sk"][off:off + length], host["sv"][off:off + length]
    np.testing.assert_array_equal(idx.to_numpy()[0], O.argsort(sk, sv, True, 0, 0))
    gk, gv = host["gk"][off:off + length], host["gv"][off:off + length]
    ids, K, rep = O.group_keys([gk], [None])
    keys, sums, cnts = (r.to_numpy()[0] for r in gres)
    np.testing.assert_array_equal(keys, gk[rep])
    np.testing.assert_array_equal(sums.view(np.uint64), O.group_agg(ids, K, gv, 0)[0])
    np.testing.assert_array_equal(cnts.view(np.uint64), O.group_agg(ids, K, None, 1)[0])
    ua, ub = host["ua"][off:off + length], host["ub"][off:off + length]
    rows = O.unique_rows([ua, ub])
    np.testing.assert_array_equal(ures[0].to_numpy()[0], rows)
    np.testing.assert_array_equal(ures[1].to_numpy()[0], ua[rows])
    np.testing.assert_array_equal(ures[2].to_numpy()[0], ub[rows])


def test_server_errors(resident):
    _, dev = resident
    srv = am.Server(slots=2)
    with pytest.raises(am.ArrowMetalError, match="unknown float_order"):
        srv.submit("argsort", [dev["sk"]], float_order=7)
    with pytest.raises(am.ArrowMetalError, match="needs a value column"):
        srv.submit("group_by", [dev["gk"]])
    bad = srv.submit("group_by", [dev["gk"].slice(0, 100)], dev["gv"].slice(0, 99))
    good = srv.submit("argsort", [dev["sk"].slice(0, 1000)])
    with pytest.raises(am.ArrowMetalError, match="length mismatch"):
        srv.wait(bad)
    (idx,), _ = srv.wait(good)
    assert len(idx) == 1000
    with pytest.raises(am.ArrowMetalError, match="unknown ticket"):
        srv.wait(bad)
    st = srv.stats
    srv.close()
    assert st["failed"] == 1 and st["completed"] == 2


def test_submit_from_another_thread_while_waiting(resident):
    """A submitter thread keeps adding requests (enough to rehash the ticket map many times) while the
    main thread blocks in wait(); every result still equals the plain call's, and a ticket can be
    waited for once only."""
    import threading
    host, dev = resident
    srv = am.Server(slots=4, budget_bytes=8 << 30)
    n = len(dev["gk"])
    reqs = []
    for i in range(96):
        length = 1000 + 997 * i
        off = (7919 * i) % (n - length)
        reqs.append(("argsort", [dev["sk"].slice(off, length)], None, {"descending": bool(i & 1)}))
    tickets = [None] * len(reqs)
    first = srv.submit(*reqs[0][:2], values=reqs[0][2], **reqs[0][3])
    tickets[0] = first

    def submitter():
        for i in range(1, len(reqs)):
            kind, cols, values, opts = reqs[i]
            tickets[i] = srv.submit(kind, cols, values=values, **opts)

    th = threading.Thread(target=submitter)
    th.start()
    got0, _ = srv.wait(first)
    th.join()
    results = [got0] + [srv.wait(t)[0] for t in tickets[1:]]
    with pytest.raises(am.ArrowMetalError, match="unknown ticket"):
        srv.wait(first)
    srv.close()
    for (kind, cols, values, opts), got in zip(reqs, results):
        want = plain(kind, cols, values, opts)
        assert len(got) == len(want)
        assert all(same(a, b) for a, b in zip(got, want))

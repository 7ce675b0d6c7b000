"""ctypes wrapper over the CPU oracle (oracle/liboracle.so). Test infrastructure only.

Only tests, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import this module, and
only to check (or time) against the HIP library; the product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

# Element type codes of oracle/oracle.c.
I8, I16, I32, I64, U8, U16, U32, U64, F32, F64 = range(10)
NP_TYPE = {I8: np.int8, I16: np.int16, I32: np.int32, I64: np.int64, U8: np.uint8, U16: np.uint16,
           U32: np.uint32, U64: np.uint64, F32: np.float32, F64: np.float64}
TYPE_OF = {np.dtype(v): k for k, v in NP_TYPE.items()}

_lib = None


def lib():
    """Loads the oracle, building it with its Makefile when it is missing (gcc only)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i64 = ctypes.c_int64
        c_int = ctypes.c_int
        L.or_splitmix64_fill.argtypes = [ctypes.c_uint64, P, i64]
        L.or_argsort.argtypes = [P, c_int, P, i64, c_int, c_int, c_int, P]
        L.or_take.argtypes = [P, c_int, P, i64, P, c_int, P, i64, P, P]
        L.or_group_keys.argtypes = [c_int, P, P, P, i64, P, P]
        L.or_group_keys.restype = i64
        L.or_unique_values.argtypes = [c_int, P, P, i64, c_int, P]
        L.or_unique_values.restype = i64
        L.or_group_agg.argtypes = [P, i64, i64, P, c_int, P, c_int, P, P]
        L.or_group_by_dense.argtypes = [P, c_int, P, i64, i64, P, c_int, P, c_int, P, P]
        L.or_lexsort.argtypes = [c_int, P, P, P, i64, P, P, P, P]
        L.or_unique_rows.argtypes = [c_int, P, P, P, i64, P]
        L.or_unique_rows.restype = i64
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def pack_validity(valid_bool):
    """Arrow LSB-first bitmap from a boolean mask (None stays None)."""
    if valid_bool is None:
        return None
    return np.packbits(np.asarray(valid_bool, dtype=bool), bitorder="little")


def splitmix64(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint64)
    lib().or_splitmix64_fill(ctypes.c_uint64(seed & (2**64 - 1)), _ptr(out), n)
    return out


def argsort(values: np.ndarray, valid=None, descending=False, null_placement=0, float_order=0):
    values = np.ascontiguousarray(values)
    bm = pack_validity(valid)
    out = np.empty(len(values), dtype=np.uint32)
    lib().or_argsort(_ptr(values), TYPE_OF[values.dtype], _ptr(bm), len(values), int(descending),
                     int(null_placement), int(float_order), _ptr(out))
    return out


def top_k(values, k, valid=None, largest=True, null_placement=0, float_order=0):
    """The first k rows of the stable argsort with descending = largest (the reference's definition,
    Kernels/Sort.swift:95-112)."""
    return argsort(values, valid, largest, null_placement, float_order)[:max(k, 0)]


def lexsort(columns, valids=None, descending=None, null_placement=None, float_order=None):
    """Multi-key stable argsort with options per key (MultiSort.swift:28-65)."""
    columns = [np.ascontiguousarray(c) for c in columns]
    nc, n = len(columns), len(columns[0])
    valids = valids or [None] * nc
    bms = [pack_validity(v) for v in valids]
    opt = lambda v: np.ascontiguousarray([0] * nc if v is None else [int(x) for x in v], dtype=np.int32)
    types = np.array([TYPE_OF[c.dtype] for c in columns], dtype=np.int32)
    vptrs = (ctypes.c_void_p * nc)(*[c.ctypes.data for c in columns])
    bptrs = (ctypes.c_void_p * nc)(*[0 if b is None else b.ctypes.data for b in bms])
    d, p, f = opt(descending), opt(null_placement), opt(float_order)
    out = np.empty(max(n, 1), dtype=np.uint32)
    lib().or_lexsort(nc, _ptr(types), ctypes.cast(vptrs, ctypes.c_void_p), ctypes.cast(bptrs, ctypes.c_void_p),
                     n, _ptr(d), _ptr(p), _ptr(f), _ptr(out))
    return out[:n].copy()


def take(values: np.ndarray, indices: np.ndarray, valid=None, idx_valid=None):
    """Returns (out, out_valid) or raises IndexError like the reference's range check."""
    values = np.ascontiguousarray(values)
    indices = np.ascontiguousarray(indices)
    bm, ibm = pack_validity(valid), pack_validity(idx_valid)
    out = np.empty(len(indices), dtype=values.dtype)
    ov = np.empty(len(indices), dtype=np.uint8)
    rc = lib().or_take(_ptr(values), values.dtype.itemsize, _ptr(bm), len(values), _ptr(indices),
                       TYPE_OF[indices.dtype], _ptr(ibm), len(indices), _ptr(out), _ptr(ov))
    if rc:
        raise IndexError(f"take: index out of range (array length {len(values)})")
    return out, ov.astype(bool)


def group_keys(columns, valids=None):
    """Dense ids (ascending key order, null group last) and representative rows for key columns."""
    columns = [np.ascontiguousarray(c) for c in columns]
    n = len(columns[0])
    valids = valids or [None] * len(columns)
    bms = [pack_validity(v) for v in valids]
    types = np.array([TYPE_OF[c.dtype] for c in columns], dtype=np.int32)
    vptrs = (ctypes.c_void_p * len(columns))(*[c.ctypes.data for c in columns])
    bptrs = (ctypes.c_void_p * len(columns))(*[0 if b is None else b.ctypes.data for b in bms])
    ids = np.empty(n, dtype=np.int32)
    rep = np.empty(max(n, 1), dtype=np.uint32)
    K = lib().or_group_keys(len(columns), _ptr(types), ctypes.cast(vptrs, ctypes.c_void_p),
                            ctypes.cast(bptrs, ctypes.c_void_p), n, _ptr(ids), _ptr(rep))
    return ids, int(K), rep[:K].copy()


def group_agg(ids: np.ndarray, K: int, values, op: int, valid=None):
    """One grouped aggregate over dense ids. Returns (bits uint64[K], valid bool[K])."""
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    t = TYPE_OF[values.dtype] if values is not None else I64
    if values is not None:
        values = np.ascontiguousarray(values)
    bm = pack_validity(valid)
    out = np.zeros(max(K, 1), dtype=np.uint64)
    ov = np.zeros(max(K, 1), dtype=np.uint8)
    rc = lib().or_group_agg(_ptr(ids), len(ids), K, _ptr(values), t, _ptr(bm), op, _ptr(out), _ptr(ov))
    if rc:
        raise ValueError(f"oracle: op {op} is not defined for this type")
    return out[:K], ov[:K].astype(bool)


def group_by_dense(keys: np.ndarray, K: int, values, agg: int, key_valid=None, valid=None):
    keys = np.ascontiguousarray(keys)
    t = TYPE_OF[values.dtype] if values is not None else I64
    if values is not None:
        values = np.ascontiguousarray(values)
    kbm, bm = pack_validity(key_valid), pack_validity(valid)
    out = np.zeros(max(K, 1), dtype=np.uint64)
    ov = np.zeros(max(K, 1), dtype=np.uint8)
    rc = lib().or_group_by_dense(_ptr(keys), TYPE_OF[keys.dtype], _ptr(kbm), len(keys), K, _ptr(values), t,
                                 _ptr(bm), agg, _ptr(out), _ptr(ov))
    if rc:
        raise ValueError(f"oracle: agg {agg} is not defined for this type")
    return out[:K], ov[:K].astype(bool)


def unique_rows(columns, valids=None):
    columns = [np.ascontiguousarray(c) for c in columns]
    n = len(columns[0])
    valids = valids or [None] * len(columns)
    bms = [pack_validity(v) for v in valids]
    types = np.array([TYPE_OF[c.dtype] for c in columns], dtype=np.int32)
    vptrs = (ctypes.c_void_p * len(columns))(*[c.ctypes.data for c in columns])
    bptrs = (ctypes.c_void_p * len(columns))(*[0 if b is None else b.ctypes.data for b in bms])
    out = np.empty(max(n, 1), dtype=np.uint32)
    m = lib().or_unique_rows(len(columns), _ptr(types), ctypes.cast(vptrs, ctypes.c_void_p),
                             ctypes.cast(bptrs, ctypes.c_void_p), n, _ptr(out))
    return out[:m].copy()


def unique_values(values, valid=None, order=0):
    """Rows whose values form Arrow `unique` (order 0 first appearance with the null once, order 1
    ascending without nulls); gather them to get the result."""
    values = np.ascontiguousarray(values)
    n = len(values)
    bm = pack_validity(valid)
    out = np.empty(max(n, 1), dtype=np.uint32)
    m = lib().or_unique_values(TYPE_OF[values.dtype], _ptr(values), None if bm is None else _ptr(bm), n,
                               order, _ptr(out))
    return out[:m].copy()


# ---------------------------------------------------------------- the plan engine's subset

_AGG_OP = {"sum": 0, "count": 2, "mean": 3, "min": 4, "max": 5}


def _agg_dtype(op, values):
    if op == "count":
        return np.dtype(np.int64)
    if op == "mean":
        return np.dtype(np.float64)
    if op == "sum":
        if values.dtype.kind == "u":
            return np.dtype(np.uint64)
        return np.dtype(np.float64) if values.dtype.kind == "f" else np.dtype(np.int64)
    return values.dtype


def _from_bits(bits, dtype):
    dtype = np.dtype(dtype)
    if dtype.itemsize == 8:
        return bits.view(dtype)
    mask = np.uint64((1 << (8 * dtype.itemsize)) - 1)
    return (bits & mask).astype(np.dtype(f"u{dtype.itemsize}")).view(dtype)


def _col_name(sexpr):
    s = sexpr.strip()
    assert s.startswith("(col") and s.endswith(")"), sexpr
    return s[4:-1].strip().strip('"')


def _sort_keys(node):
    """(column, descending, nulls_first, float_order) per key, with the sort-level defaults
    (Engine/PlanJSON.swift:96-118)."""
    dnf = node.get("nulls") == "first"
    dfo = {"ieee": 0, "total": 1, "nan_largest": 2}[node.get("float_order") or "ieee"]
    out = []
    for k in node.get("by", []):
        if isinstance(k, dict):
            col, desc, opts = k["column"], bool(k.get("descending", False)), k
        else:
            col, desc = k[0], bool(k[1]) if len(k) > 1 else False
            opts = k[2] if len(k) > 2 and isinstance(k[2], dict) else {}
        nf = dnf if opts.get("nulls") is None else opts["nulls"] == "first"
        fo = dfo if opts.get("float_order") is None else {"ieee": 0, "total": 1, "nan_largest": 2}[opts["float_order"]]
        out.append((col, desc, nf, fo))
    return out


def _take_table(t, idx):
    out = {}
    for name, (v, valid) in t.items():
        vals, ov = take(v, idx.astype(np.uint32), valid)
        out[name] = (vals, ov if valid is not None else None)
    return out


def plan_run(plan, sources):
    """CPU restatement of the plan engine's subset (Engine/Executor.swift:46-160, 243-350;
    PhysicalPlan.swift:118-126): `sources` maps a source name to {column: (values, valid or None)};
    returns {column: (values, valid or None)} in output order. Composed from the oracle's own argsort,
    lexsort, take, group keys, aggregates and keep-first unique."""
    op = plan["op"]
    if op == "scan":
        src = sources[plan["source"]]
        cols = plan.get("columns")
        return dict(src) if cols is None else {c: src[c] for c in cols}
    if op == "sort":
        t = plan_run(plan["input"], sources)
        keys = _sort_keys(plan)
        n = len(next(iter(t.values()))[0]) if t else 0
        if not keys or n == 0:
            return t
        idx = lexsort([t[k[0]][0] for k in keys], [t[k[0]][1] for k in keys], [k[1] for k in keys],
                      [1 if k[2] else 0 for k in keys], [k[3] for k in keys])
        return _take_table(t, idx)
    if op == "limit":
        # A top-k selection is the first k rows of the stable sort, so the sort-then-slice form gives
        # the same rows.
        t = plan_run(plan["input"], sources)
        n = len(next(iter(t.values()))[0]) if t else 0
        lo = min(max(int(plan.get("offset", 0)), 0), n)
        count = plan.get("count", 2**63 - 1)
        ln = max(min(int(count), n - lo), 0)
        return {k: (v[lo:lo + ln], None if va is None else va[lo:lo + ln]) for k, (v, va) in t.items()}
    if op == "unique":
        t = plan_run(plan["input"], sources)
        names = plan.get("subset") or list(t)
        n = len(next(iter(t.values()))[0]) if t else 0
        if n == 0:
            return t
        keep = unique_rows([t[c][0] for c in names], [t[c][1] for c in names])
        return _take_table(t, keep)
    if op == "group_by":
        t = plan_run(plan["input"], sources)
        keys = [(nm, _col_name(e)) for nm, e in plan.get("keys", [])]
        aggs = [(a[0], a[1], _col_name(a[2]) if len(a) > 2 and a[2] else None) for a in plan.get("aggs", [])]
        kcols = [t[c][0] for _, c in keys]
        kvalid = [t[c][1] for _, c in keys]
        ids, K, rep = group_keys(kcols, kvalid)
        out = {}
        for (nm, c), v, va in zip(keys, kcols, kvalid):
            kv, kva = take(v, rep.astype(np.uint32), va)
            out[nm] = (kv, kva if va is not None else None)
        for aop, nm, c in aggs:
            if aop == "count" and c is None:
                bits, ok = group_agg(ids, K, None, 1)
                out[nm] = (bits.view(np.int64), None)
                continue
            v, va = t[c]
            bits, ok = group_agg(ids, K, v, 2 if aop == "count" else _AGG_OP[aop], va)
            out[nm] = (_from_bits(bits, _agg_dtype(aop, v)), None if ok.all() else ok)
        return out
    raise ValueError(f"oracle plan: operator {op} is outside the subset")

"""ctypes binding of libArrowMetalC.so (the MI355X build of the ArrowMetal C ABI).

It mirrors the reference's Python package (python/arrowmetal/__init__.py: the ctypes loader at
:26-78, MetalArray at :251-475, GroupByKeys / group_by at :2344-2510, batch at :787-797) for the
argsort / take / group-by path, so that tests read like the reference's own. Arrays cross the C Data
Interface: pyarrow arrays through ``_export_to_c`` / ``_import_from_c`` as the reference does, NumPy
arrays through hand-built ArrowArray structs (no pyarrow needed).

The library is looked up in ``$ARROWMETAL_LIB`` (an explicit path; a wrong value is an error, never a
fall-through), then next to this file. There is no fallback implementation: without the HIP library,
or without a ROCm device, every call raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_LIB_NAME = "libArrowMetalC.so"
_HERE = os.path.dirname(os.path.abspath(__file__))


def library_path() -> str:
    env = os.environ.get("ARROWMETAL_LIB")
    if env:
        if not os.path.exists(env):
            raise OSError(f"arrowmetal: $ARROWMETAL_LIB is set to {env}, which does not exist")
        return env
    p = os.path.join(_HERE, _LIB_NAME)
    if not os.path.exists(p):
        raise OSError(f"arrowmetal: {p} not found; build it with `python arrowmetal_amd/build.py`")
    return p


_P = ctypes.c_void_p
_lib = None


class ArrowMetalError(RuntimeError):
    pass


class _ArrowSchema(ctypes.Structure):
    pass


_ArrowSchema._fields_ = [("format", ctypes.c_char_p), ("name", ctypes.c_char_p), ("metadata", ctypes.c_char_p),
                         ("flags", ctypes.c_int64), ("n_children", ctypes.c_int64),
                         ("children", ctypes.POINTER(ctypes.POINTER(_ArrowSchema))),
                         ("dictionary", ctypes.POINTER(_ArrowSchema)),
                         ("release", ctypes.CFUNCTYPE(None, ctypes.POINTER(_ArrowSchema))),
                         ("private_data", ctypes.c_void_p)]


class _ArrowArray(ctypes.Structure):
    pass


_ArrowArray._fields_ = [("length", ctypes.c_int64), ("null_count", ctypes.c_int64), ("offset", ctypes.c_int64),
                        ("n_buffers", ctypes.c_int64), ("n_children", ctypes.c_int64),
                        ("buffers", ctypes.POINTER(ctypes.c_void_p)),
                        ("children", ctypes.POINTER(ctypes.POINTER(_ArrowArray))),
                        ("dictionary", ctypes.POINTER(_ArrowArray)),
                        ("release", ctypes.CFUNCTYPE(None, ctypes.POINTER(_ArrowArray))),
                        ("private_data", ctypes.c_void_p)]


class _ArrowDeviceArray(ctypes.Structure):
    _fields_ = [("array", _ArrowArray), ("device_id", ctypes.c_int64), ("device_type", ctypes.c_int32),
                ("reserved", ctypes.c_int64 * 3), ("sync_event", ctypes.c_void_p)]


# Exported symbols of include/arrowmetal.h and their ctypes signatures.
_SIGNATURES = {
    "am_version": ([], ctypes.c_char_p),
    "am_device_name": ([], ctypes.c_char_p),
    "am_last_error": ([], ctypes.c_char_p),
    "am_import": ([_P, _P, ctypes.POINTER(_P)], ctypes.c_int),
    "am_import_device": ([_P, _P, ctypes.POINTER(_P)], ctypes.c_int),
    "am_export": ([_P, _P, _P], ctypes.c_int),
    "am_export_device": ([_P, _P, _P], ctypes.c_int),
    "am_release": ([_P], None),
    "am_length": ([_P], ctypes.c_int64),
    "am_null_count": ([_P], ctypes.c_int64),
    "am_format": ([_P], ctypes.c_char_p),
    "am_compute_format": ([_P], ctypes.c_char_p),
    "am_take": ([_P, _P, ctypes.POINTER(_P)], ctypes.c_int),
    "am_argsort": ([_P, ctypes.c_int, ctypes.POINTER(_P)], ctypes.c_int),
    "am_argsort_ex": ([_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_P)], ctypes.c_int),
    "am_argsort_ex2": ([_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_P)], ctypes.c_int),
    "am_group_by": ([_P, ctypes.c_int64, ctypes.c_int, _P, ctypes.POINTER(_P)], ctypes.c_int),
    "am_sort": ([_P, ctypes.c_int, ctypes.POINTER(_P)], ctypes.c_int),
    "am_sort_ex": ([_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_P)], ctypes.c_int),
    "am_sort_ex2": ([_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_P)], ctypes.c_int),
    "am_top_k": ([_P, ctypes.c_int64, ctypes.c_int, ctypes.POINTER(_P)], ctypes.c_int),
    "am_top_k_ex": ([_P, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_P)], ctypes.c_int),
    "am_lexsort": ([ctypes.POINTER(_P), _P, ctypes.c_int64, ctypes.POINTER(_P)], ctypes.c_int),
    "am_lexsort_ex": ([ctypes.POINTER(_P), _P, ctypes.c_int64, ctypes.c_int, ctypes.POINTER(_P)], ctypes.c_int),
    "am_lexsort_ex2": ([ctypes.POINTER(_P), _P, _P, _P, ctypes.c_int64, ctypes.POINTER(_P)], ctypes.c_int),
    "am_batch_begin": ([], ctypes.c_int),
    "am_batch_end": ([], ctypes.c_int),
    "am_group_by_keys": ([ctypes.POINTER(_P), ctypes.c_int64, ctypes.POINTER(_P)], ctypes.c_int),
    "am_group_by_group_count": ([_P], ctypes.c_int64),
    "am_group_by_keys_result": ([_P, ctypes.c_int64, ctypes.POINTER(_P)], ctypes.c_int),
    "am_group_by_ids": ([_P, ctypes.POINTER(_P)], ctypes.c_int),
    "am_group_by_release": ([_P], None),
    "am_group_agg_ex": ([_P, _P, ctypes.c_int, ctypes.c_double, ctypes.POINTER(_P)], ctypes.c_int),
    "am_unique_rows": ([ctypes.POINTER(_P), ctypes.c_int64, ctypes.POINTER(_P)], ctypes.c_int),
    "am_unique_ex": ([_P, ctypes.c_int, ctypes.POINTER(_P)], ctypes.c_int),
    "am_slice": ([_P, ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(_P)], ctypes.c_int),
    "am_batched_create": ([ctypes.POINTER(ctypes.c_int64), ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                           ctypes.POINTER(_P)], ctypes.c_int),
    "am_batched_tune": ([_P, _P, _P, _P, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "am_batched_run": ([_P, _P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "am_batched_result": ([_P, ctypes.c_int, ctypes.POINTER(_P)], ctypes.c_int),
    "am_batched_config": ([_P, ctypes.POINTER(ctypes.c_int64), ctypes.c_int], ctypes.c_int),
    "am_batched_release": ([_P], None),
    "am_server_create": ([ctypes.c_int, ctypes.c_int64, ctypes.POINTER(_P)], ctypes.c_int),
    "am_server_submit": ([_P, ctypes.c_int, ctypes.POINTER(_P), ctypes.c_int64, _P, ctypes.c_int, ctypes.c_int,
                          ctypes.c_int, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "am_server_wait": ([_P, ctypes.c_int64, ctypes.POINTER(_P), ctypes.c_int64, ctypes.POINTER(ctypes.c_double),
                        ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "am_server_stats": ([_P, ctypes.POINTER(ctypes.c_int64), ctypes.c_int], ctypes.c_int),
    "am_server_release": ([_P], None),
    "am_plan_source_create": ([ctypes.c_char_p, ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_char_p), ctypes.c_int64,
                               ctypes.POINTER(_P)], ctypes.c_int),
    "am_plan_source_release": ([_P], None),
    "am_plan_run": ([ctypes.c_char_p, ctypes.POINTER(_P), ctypes.c_int64, ctypes.c_int, ctypes.POINTER(_P)],
                    ctypes.c_int),
    "am_plan_explain": ([ctypes.c_char_p, ctypes.POINTER(_P), ctypes.c_int64, ctypes.c_int], ctypes.c_char_p),
    "am_plan_column_count": ([_P], ctypes.c_int64),
    "am_plan_row_count": ([_P], ctypes.c_int64),
    "am_plan_column_name": ([_P, ctypes.c_int64], ctypes.c_char_p),
    "am_plan_column": ([_P, ctypes.c_int64, ctypes.POINTER(_P)], ctypes.c_int),
    "am_plan_result_release": ([_P], None),
    "am_group_by_noalloc": ([_P, ctypes.c_int64, ctypes.c_int, _P, _P, _P, _P], ctypes.c_int),
    "am_import_chunks": ([_P, _P, ctypes.c_int64, ctypes.POINTER(_P)], ctypes.c_int),
    "am_import_chunks_supported": ([_P], ctypes.c_int),
    "am_query": ([ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_char_p), ctypes.c_int64, ctypes.c_char_p,
                  ctypes.POINTER(_P)], ctypes.c_int),
    "am_query_column_count": ([_P], ctypes.c_int64),
    "am_query_column_name": ([_P, ctypes.c_int64], ctypes.c_char_p),
    "am_query_column": ([_P, ctypes.c_int64, ctypes.POINTER(_P)], ctypes.c_int),
    "am_query_scalar_count": ([_P], ctypes.c_int64),
    "am_query_scalar_name": ([_P, ctypes.c_int64], ctypes.c_char_p),
    "am_query_scalar": ([_P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double),
                         ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "am_query_result_release": ([_P], None),
    "am_query_canonical": ([ctypes.c_char_p], ctypes.c_char_p),
    "am_stream_last_error": ([], ctypes.c_char_p),
    "am_stream_from_c_stream": ([_P, ctypes.c_int], _P),
    "am_stream_release": ([_P], None),
    "am_stream_group_by": ([_P, ctypes.POINTER(ctypes.c_char_p), ctypes.c_int64, ctypes.POINTER(ctypes.c_int),
                            ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_char_p), ctypes.c_int64,
                            ctypes.c_int64, ctypes.c_int, ctypes.POINTER(_P)], ctypes.c_int),
    "am_stream_result_column_count": ([_P], ctypes.c_int64),
    "am_stream_result_column_name": ([_P, ctypes.c_int64], ctypes.c_char_p),
    "am_stream_result_column": ([_P, ctypes.c_int64, ctypes.POINTER(_P)], ctypes.c_int),
    "am_stream_result_scalar_count": ([_P], ctypes.c_int64),
    "am_stream_result_scalar_name": ([_P, ctypes.c_int64], ctypes.c_char_p),
    "am_stream_result_scalar": ([_P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "am_stream_result_rows_out": ([_P], ctypes.c_int64),
    "am_stream_result_stats": ([_P] + [ctypes.POINTER(ctypes.c_int64)] * 3 + [ctypes.POINTER(ctypes.c_double)] * 5,
                               ctypes.c_int),
    "am_stream_result_release": ([_P], None),
    "am_set_max_row_index": ([ctypes.c_int64], ctypes.c_int64),
    "am_prof_enable": ([ctypes.c_int], ctypes.c_int),
    "am_prof_reset": ([], ctypes.c_int),
    "am_prof_report": ([ctypes.c_char_p, ctypes.c_int64], ctypes.c_int64),
}


def lib():
    """Loads libArrowMetalC.so (once) and declares the signatures of every entry point."""
    global _lib
    if _lib is None:
        L = ctypes.CDLL(library_path())
        for name, (args, res) in _SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def version():
    return lib().am_version().decode()


def device_name():
    return lib().am_device_name().decode()


def _check(rc):
    if rc == 2:
        raise ArrowMetalError("bad argument (NULL handle)")
    if rc != 0:
        raise ArrowMetalError((lib().am_last_error() or b"unknown error").decode())


def _call(fn, *args):
    out = _P()
    _check(fn(*args, ctypes.byref(out)))
    return MetalArray(out)


_NP_FORMAT = {np.dtype(np.int8): "c", np.dtype(np.uint8): "C", np.dtype(np.int16): "s",
              np.dtype(np.uint16): "S", np.dtype(np.int32): "i", np.dtype(np.uint32): "I",
              np.dtype(np.int64): "l", np.dtype(np.uint64): "L", np.dtype(np.float32): "f",
              np.dtype(np.float64): "g"}
_FORMAT_NP = {v: k for k, v in _NP_FORMAT.items()}


def _np_of_format(fmt: str):
    if fmt in _FORMAT_NP:
        return _FORMAT_NP[fmt]
    if fmt in ("tdD", "tts", "ttm", "tiM"):
        return np.dtype(np.int32)
    if fmt.startswith("t"):
        return np.dtype(np.int64)
    raise ArrowMetalError(f"no NumPy type for Arrow format {fmt!r}")


# Producer-side release callbacks for arrays built from NumPy. The library moves the struct (it sets
# the caller's release to NULL) and calls this on its own copy once it has copied the buffers.
_keep = {}


@ctypes.CFUNCTYPE(None, ctypes.POINTER(_ArrowArray))
def _release_array(p):
    _keep.pop(p.contents.private_data, None)
    p.contents.release = ctypes.CFUNCTYPE(None, ctypes.POINTER(_ArrowArray))()


def _c_array(values: np.ndarray, valid=None, offset: int = 0, length=None):
    """A C Data Interface array over NumPy buffers (bit-packed validity from a bool mask or bytes)."""
    values = np.ascontiguousarray(values)
    n = len(values) - offset if length is None else length
    bitmap = None
    nulls = 0
    if valid is not None:
        valid = np.asarray(valid)
        if valid.dtype == bool:
            nulls = int(len(valid) - np.count_nonzero(valid[offset:offset + n])) if n else 0
            bitmap = np.packbits(valid, bitorder="little")
        else:
            bitmap = np.ascontiguousarray(valid, dtype=np.uint8)
            nulls = -1
    bufs = (ctypes.c_void_p * 2)(None if bitmap is None else bitmap.ctypes.data, values.ctypes.data)
    arr = _ArrowArray()
    arr.length = n
    arr.null_count = nulls
    arr.offset = offset
    arr.n_buffers = 2
    arr.n_children = 0
    arr.buffers = ctypes.cast(bufs, ctypes.POINTER(ctypes.c_void_p))
    token = id(arr)
    arr.private_data = token
    arr.release = _release_array
    _keep[token] = (values, bitmap, bufs)
    return arr


class MetalArray:
    """An Arrow array resident in HBM. Immutable; every operation returns a new array."""

    def __init__(self, handle):
        self._h = handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.am_release(h)
            self._h = None

    # ---- interop
    @classmethod
    def from_numpy(cls, values, valid=None, fmt=None, offset=0, length=None):
        """Imports a NumPy array (copied into HBM once). `valid` is a bool mask (True = valid) or an
        Arrow bitmap; `fmt` overrides the Arrow format (temporal types)."""
        values = np.ascontiguousarray(values)
        fmt = fmt or _NP_FORMAT[values.dtype]
        schema = _ArrowSchema()
        schema.format = fmt.encode()
        schema.name = b""
        schema.flags = 2
        arr = _c_array(values, valid, offset, length)
        out = _P()
        rc = lib().am_import(ctypes.addressof(schema), ctypes.addressof(arr), ctypes.byref(out))
        if arr.release:           # not moved (an argument error): release it ourselves
            arr.release(ctypes.byref(arr))
        _check(rc)
        return cls(out)

    @classmethod
    def from_chunks(cls, chunks, fmt=None):
        """One array from several chunks of one type (`am_import_chunks`): a list of (values, valid,
        offset, length) tuples of NumPy buffers (valid: a bool mask, an Arrow bitmap with null_count -1,
        or None), or a pyarrow ChunkedArray. Returns None when the library does not take the type (3)."""
        try:
            import pyarrow as pa
            is_chunked = isinstance(chunks, pa.ChunkedArray)
        except ImportError:
            is_chunked = False
        n = len(chunks.chunks) if is_chunked else len(chunks)
        arrs = (_ArrowArray * max(n, 1))()
        schema = _ArrowSchema()
        extra = []
        try:
            if is_chunked:
                for i, c in enumerate(chunks.chunks):
                    sc = schema if i == 0 else _ArrowSchema()
                    c._export_to_c(ctypes.addressof(arrs[i]), ctypes.addressof(sc))
                    if i:
                        extra.append(sc)
                if n == 0:
                    chunks.type._export_to_c(ctypes.addressof(schema))
            else:
                first = np.ascontiguousarray(chunks[0][0]) if n else np.zeros(0, dtype=np.int64)
                schema.format = (fmt or _NP_FORMAT[first.dtype]).encode()
                schema.name = b""
                schema.flags = 2
                for i, (v, valid, off, ln) in enumerate(chunks):
                    a = _c_array(v, valid, off, ln)
                    extra.append(a)   # its token keys the buffers in _keep until the release runs
                    ctypes.memmove(ctypes.addressof(arrs[i]), ctypes.addressof(a), ctypes.sizeof(_ArrowArray))
            out = _P()
            rc = lib().am_import_chunks(ctypes.addressof(schema), ctypes.addressof(arrs), n, ctypes.byref(out))
            if rc == 3:
                return None
            _check(rc)
            return cls(out)
        finally:
            for i in range(n):
                if arrs[i].release:
                    arrs[i].release(ctypes.byref(arrs[i]))
            if is_chunked:
                for sc in [schema] + extra:
                    if sc.release:
                        sc.release(ctypes.byref(sc))

    @classmethod
    def from_arrow(cls, obj):
        """Imports anything pyarrow can export through the C Data Interface."""
        import pyarrow as pa
        if isinstance(obj, pa.ChunkedArray):
            obj = obj.combine_chunks()
        if not isinstance(obj, pa.Array):
            obj = pa.array(obj)
        schema = _ArrowSchema()
        array = _ArrowArray()
        obj._export_to_c(ctypes.addressof(array), ctypes.addressof(schema))
        try:
            return _call(lib().am_import, ctypes.addressof(schema), ctypes.addressof(array))
        finally:
            if schema.release:
                schema.release(ctypes.byref(schema))
            if array.release:
                array.release(ctypes.byref(array))

    def to_arrow(self):
        import pyarrow as pa
        schema = _ArrowSchema()
        array = _ArrowArray()
        _check(lib().am_export(self._h, ctypes.addressof(schema), ctypes.addressof(array)))
        return pa.Array._import_from_c(ctypes.addressof(array), ctypes.addressof(schema))

    def to_numpy(self):
        """(values, valid) as NumPy arrays; `valid` is a bool mask, or None when there are no nulls."""
        schema = _ArrowSchema()
        array = _ArrowArray()
        _check(lib().am_export(self._h, ctypes.addressof(schema), ctypes.addressof(array)))
        try:
            fmt = schema.format.decode()
            dt = _np_of_format(fmt)
            n = array.length
            vals = np.empty(n, dtype=dt)
            if n:
                ctypes.memmove(vals.ctypes.data, array.buffers[1] + array.offset * dt.itemsize, n * dt.itemsize)
            valid = None
            if array.buffers[0] and array.null_count != 0 and n:
                nbytes = (array.offset + n + 7) // 8
                raw = np.empty(nbytes, dtype=np.uint8)
                ctypes.memmove(raw.ctypes.data, array.buffers[0], nbytes)
                valid = np.unpackbits(raw, bitorder="little")[array.offset:array.offset + n].astype(bool)
            return vals, valid
        finally:
            array.release(ctypes.byref(array))
            schema.release(ctypes.byref(schema))

    def device_buffers(self):
        """(values_ptr, validity_ptr, offset, length, null_count) of the HBM buffers (am_export_device)."""
        schema = _ArrowSchema()
        darr = _ArrowDeviceArray()
        _check(lib().am_export_device(self._h, ctypes.addressof(schema), ctypes.addressof(darr)))
        a = darr.array
        out = (a.buffers[1], a.buffers[0], a.offset, a.length, a.null_count, darr.device_type)
        a.release(ctypes.byref(a))
        schema.release(ctypes.byref(schema))
        return out

    # ---- properties
    def __len__(self):
        return lib().am_length(self._h)

    @property
    def null_count(self):
        return lib().am_null_count(self._h)

    @property
    def format(self):
        return lib().am_format(self._h).decode()

    # ---- selection and sorting
    def take(self, indices):
        if not isinstance(indices, MetalArray):
            indices = MetalArray.from_numpy(np.asarray(indices))
        return _call(lib().am_take, self._h, indices._h)

    def argsort(self, descending=False, null_placement="at_end", float_order="ieee"):
        """uint32 indices that sort the array stably (Arrow `array_sort_indices` with the reference's
        null placement and float orders)."""
        if null_placement == "at_end" and float_order == "ieee":
            return _call(lib().am_argsort, self._h, 1 if descending else 0)
        return _call(lib().am_argsort_ex2, self._h, 1 if descending else 0,
                     NULL_PLACEMENTS.index(null_placement), _float_order(float_order))

    def sort(self, descending=False, null_placement="at_end", float_order="ieee"):
        """The values in argsort's order (`am_sort_ex2`), same type; -0.0 and NaN bits kept."""
        return _call(lib().am_sort_ex2, self._h, 1 if descending else 0, NULL_PLACEMENTS.index(null_placement),
                     _float_order(float_order))

    def top_k(self, k, largest=True, null_placement="at_end", float_order="ieee"):
        """The first k indices argsort(descending=largest, ...) gives (`am_top_k_ex`)."""
        return _call(lib().am_top_k_ex, self._h, int(k), 1 if largest else 0, NULL_PLACEMENTS.index(null_placement),
                     _float_order(float_order))

    def group_by(self, key_count):
        """Dense-key group-by (`am_group_by`): this array holds keys in [0, key_count)."""
        return GroupBy(self, key_count)

    def slice(self, offset, length):
        """Zero-copy window of `length` elements from `offset` (`am_slice`)."""
        return _call(lib().am_slice, self._h, int(offset), int(length))

    def unique(self, order="first_appearance"):
        """Arrow `unique` (`am_unique_ex`): "first_appearance" keeps the null once, at the first null
        row; "sorted" is ascending with nulls dropped."""
        if order not in ("first_appearance", "sorted"):
            raise ValueError(f"unknown value order {order!r}")
        return _call(lib().am_unique_ex, self._h, 1 if order == "sorted" else 0)


NULL_PLACEMENTS = ("at_end", "at_start")
FLOAT_ORDERS = ("ieee", "total", "nan_largest")


def _float_order(v):
    if isinstance(v, int):
        return v
    return FLOAT_ORDERS.index(v)


_AGG = {"sum": 0, "count": 1, "min": 2, "max": 3, "mean": 4, "count_values": 5}


class GroupBy:
    """am_group_by over dense keys in [0, key_count)."""

    def __init__(self, keys, key_count):
        self.keys, self.key_count = keys, key_count

    def _agg(self, name, values):
        return _call(lib().am_group_by, self.keys._h, self.key_count, _AGG[name],
                     values._h if values is not None else None)

    def count(self): return self._agg("count", None)
    def sum(self, values): return self._agg("sum", values)
    def min(self, values): return self._agg("min", values)
    def max(self, values): return self._agg("max", values)
    def mean(self, values): return self._agg("mean", values)
    def count_values(self, values): return self._agg("count_values", values)


_GROUP_AGG = {"sum": 0, "count_all": 1, "count": 2, "mean": 3, "min": 4, "max": 5, "min_max": 6}


class GroupByKeys:
    """Arrow `hash_*` aggregation over arbitrary key columns (am_group_by_keys + am_group_agg_ex).
    Groups come out ascending by key, the null group last."""

    def __init__(self, columns):
        cols = [c if isinstance(c, MetalArray) else MetalArray.from_numpy(np.asarray(c)) for c in columns]
        if not cols:
            raise ArrowMetalError("group_by needs at least one key column")
        self._columns = cols
        handles = (_P * len(cols))(*[c._h for c in cols])
        out = _P()
        _check(lib().am_group_by_keys(handles, len(cols), ctypes.byref(out)))
        self._h = out
        self.group_count = lib().am_group_by_group_count(self._h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.am_group_by_release(h)
            self._h = None

    def __len__(self):
        return self.group_count

    def key_arrays(self):
        out = []
        for i in range(len(self._columns)):
            h = _P()
            _check(lib().am_group_by_keys_result(self._h, i, ctypes.byref(h)))
            out.append(MetalArray(h))
        return out

    def keys(self):
        return [k.to_numpy() for k in self.key_arrays()]

    def ids(self):
        return _call(lib().am_group_by_ids, self._h)

    def _agg(self, name, values, p1=0.0):
        column = None if values is None else (values if isinstance(values, MetalArray)
                                              else MetalArray.from_numpy(np.asarray(values)))
        out = _P()
        _check(lib().am_group_agg_ex(self._h, None if column is None else column._h, _GROUP_AGG[name],
                                     float(p1), ctypes.byref(out)))
        return MetalArray(out)

    def sum(self, values): return self._agg("sum", values)
    def count(self, values): return self._agg("count", values)
    def count_all(self): return self._agg("count_all", None)
    def mean(self, values): return self._agg("mean", values)
    def min(self, values): return self._agg("min", values)
    def max(self, values): return self._agg("max", values)
    def min_max(self, values): return self._agg("min_max", values)


def group_by(keys):
    if not isinstance(keys, (list, tuple)):
        keys = [keys]
    return GroupByKeys(keys)


def _int_array(vals, count):
    if vals is None:
        return None
    if len(vals) != count:
        raise ValueError(f"{len(vals)} options for {count} columns")
    return (ctypes.c_int * count)(*[int(v) for v in vals])


def lexsort(columns, descending=None, null_placement=None, float_order=None):
    """Stable multi-key argsort (`am_lexsort_ex2`): options per key, or None for the defaults."""
    handles = (_P * len(columns))(*[c._h for c in columns])
    n = len(columns)
    place = None if null_placement is None else [NULL_PLACEMENTS.index(p) if isinstance(p, str) else p
                                                  for p in null_placement]
    orders = None if float_order is None else [_float_order(f) for f in float_order]
    d, p, f = _int_array(descending, n), _int_array(place, n), _int_array(orders, n)
    out = _P()
    _check(lib().am_lexsort_ex2(handles, d, p, f, n, ctypes.byref(out)))
    return MetalArray(out)


def unique_rows(columns):
    """Keep-first unique rows over key columns (the plan engine's `unique`), as uint32 row numbers."""
    handles = (_P * len(columns))(*[c._h for c in columns])
    out = _P()
    _check(lib().am_unique_rows(handles, len(columns), ctypes.byref(out)))
    return MetalArray(out)


class PlanSource:
    """A registered table the plan engine scans by name (`am_plan_source_create`; reference
    python/arrowmetal/lazy.py:650-680). `columns` maps names to MetalArray (or NumPy arrays, imported
    once)."""

    def __init__(self, name, columns):
        self._h = None
        self.name = name
        self._cols = {k: (v if isinstance(v, MetalArray) else MetalArray.from_numpy(v)) for k, v in columns.items()}
        names = list(self._cols)
        handles = (_P * max(len(names), 1))(*[self._cols[k]._h for k in names])
        cnames = (ctypes.c_char_p * max(len(names), 1))(*[k.encode() for k in names])
        out = _P()
        _check(lib().am_plan_source_create(name.encode(), handles, cnames, len(names), ctypes.byref(out)))
        self._h = out

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h and _lib is not None:
            _lib.am_plan_source_release(h)


def query(columns, text):
    """am_query over named columns ({name: MetalArray or NumPy array}); returns {name: MetalArray}."""
    cols = {k: (v if isinstance(v, MetalArray) else MetalArray.from_numpy(v)) for k, v in columns.items()}
    names = list(cols)
    handles = (_P * max(len(names), 1))(*[cols[k]._h for k in names])
    cnames = (ctypes.c_char_p * max(len(names), 1))(*[k.encode() for k in names])
    out = _P()
    L = lib()
    _check(L.am_query(handles, cnames, len(names), text.encode(), ctypes.byref(out)))
    try:
        res = {}
        for i in range(L.am_query_column_count(out)):
            h = _P()
            _check(L.am_query_column(out, i, ctypes.byref(h)))
            res[L.am_query_column_name(out, i).decode()] = MetalArray(h)
        return res
    finally:
        L.am_query_result_release(out)


def query_canonical(text):
    r = lib().am_query_canonical(text.encode())
    if r is None:
        raise ArrowMetalError((lib().am_last_error() or b"unknown error").decode())
    return r.decode()


class _ArrowArrayStream(ctypes.Structure):
    _fields_ = [("get_schema", ctypes.c_void_p), ("get_next", ctypes.c_void_p), ("get_last_error", ctypes.c_void_p),
                ("release", ctypes.c_void_p), ("private_data", ctypes.c_void_p)]


STREAM_OPS = {"sum": 0, "count": 1, "min": 2, "max": 3, "mean": 4, "variance": 5, "stddev": 6,
              "count_distinct_approx": 7}


def stream_group_by(reader, keys, aggs, dense_key_count=0):
    """Streaming group-by (`am_stream_from_c_stream` + `am_stream_group_by`) over a pyarrow
    RecordBatchReader (or anything with `_export_to_c`). `aggs` is a list of (op, column or None, name).
    Returns ({name: MetalArray}, stats dict)."""
    cs = _ArrowArrayStream()
    reader._export_to_c(ctypes.addressof(cs))
    L = lib()
    s = L.am_stream_from_c_stream(ctypes.addressof(cs), 0)
    if not s:
        raise ArrowMetalError((L.am_stream_last_error() or b"").decode())
    try:
        k = (ctypes.c_char_p * len(keys))(*[x.encode() for x in keys])
        ops = (ctypes.c_int * len(aggs))(*[STREAM_OPS[a[0]] if isinstance(a[0], str) else a[0] for a in aggs])
        cols = (ctypes.c_char_p * len(aggs))(*[None if a[1] is None else a[1].encode() for a in aggs])
        names = (ctypes.c_char_p * len(aggs))(*[a[2].encode() for a in aggs])
        out = _P()
        rc = L.am_stream_group_by(s, k, len(keys), ops, cols, names, len(aggs), dense_key_count, 0, ctypes.byref(out))
        if rc == 2:
            raise ArrowMetalError("bad argument (NULL handle)")
        if rc != 0:
            raise ArrowMetalError((L.am_stream_last_error() or b"").decode())
        try:
            res = {}
            for i in range(L.am_stream_result_column_count(out)):
                h = _P()
                _check(L.am_stream_result_column(out, i, ctypes.byref(h)))
                res[L.am_stream_result_column_name(out, i).decode()] = MetalArray(h)
            ints = [ctypes.c_int64() for _ in range(3)]
            dbl = [ctypes.c_double() for _ in range(5)]
            _check(L.am_stream_result_stats(out, *[ctypes.byref(x) for x in ints], *[ctypes.byref(x) for x in dbl]))
            stats = dict(zip(("batches", "rows", "bytes_read"), [x.value for x in ints]))
            stats.update(zip(("wall_s", "read_s", "gpu_s", "merge_s", "overlap"), [x.value for x in dbl]))
            stats["rows_out"] = L.am_stream_result_rows_out(out)
            return res, stats
        finally:
            L.am_stream_result_release(out)
    finally:
        L.am_stream_release(s)


def _plan_text(plan):
    import json
    return plan if isinstance(plan, str) else json.dumps(plan)


def _source_array(sources):
    sources = list(sources)
    return (_P * max(len(sources), 1))(*[s._h for s in sources]), len(sources)


def plan_run(plan, sources, optimize=True):
    """Runs a JSON plan (a dict or its text) with `am_plan_run`; returns the result as an ordered dict
    of column name -> MetalArray."""
    arr, n = _source_array(sources)
    out = _P()
    _check(lib().am_plan_run(_plan_text(plan).encode(), arr, n, 1 if optimize else 0, ctypes.byref(out)))
    try:
        L = lib()
        cols = {}
        for i in range(L.am_plan_column_count(out)):
            name = L.am_plan_column_name(out, i).decode()
            h = _P()
            _check(L.am_plan_column(out, i, ctypes.byref(h)))
            cols[name] = MetalArray(h)
    finally:
        lib().am_plan_result_release(out)
    return cols


def plan_explain(plan, sources=(), optimize=True):
    """The physical plan `am_plan_explain` prints (raises with am_last_error() when it does not parse)."""
    arr, n = _source_array(sources)
    r = lib().am_plan_explain(_plan_text(plan).encode(), arr, n, 1 if optimize else 0)
    if r is None:
        raise ArrowMetalError((lib().am_last_error() or b"unknown error").decode())
    return r.decode()


class Batched:
    """A plan over record batches stored back to back (`am_batched_*`): batch b is rows
    [offsets[b], offsets[b+1]) of every column. One run sorts every batch and groups every batch in a
    fixed kernel sequence, captured in a hipGraph and replayed."""

    RESULTS = {"indices": 0, "keys": 1, "sums": 2, "counts": 3, "group_offsets": 4}

    def __init__(self, offsets, hash_capacity=2048, tile_items=16):
        offs = np.ascontiguousarray(offsets, dtype=np.int64)
        self._h = None
        h = _P()
        _check(lib().am_batched_create(offs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(offs) - 1,
                                       int(hash_capacity), int(tile_items), ctypes.byref(h)))
        self._h = h
        self.nbatches = len(offs) - 1

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.am_batched_release(self._h)
            self._h = None

    @staticmethod
    def _handles(sort_keys, group_keys, values):
        return tuple(None if a is None else a._h for a in (sort_keys, group_keys, values))

    def tune(self, sort_keys=None, group_keys=None, values=None, descending=False, float_order="ieee"):
        _check(lib().am_batched_tune(self._h, *self._handles(sort_keys, group_keys, values),
                                     1 if descending else 0, _float_order(float_order)))
        return self.config

    def run(self, sort_keys=None, group_keys=None, values=None, descending=False, float_order="ieee",
            use_graph=True):
        _check(lib().am_batched_run(self._h, *self._handles(sort_keys, group_keys, values),
                                    1 if descending else 0, _float_order(float_order), 1 if use_graph else 0))

    def result(self, name):
        out = _P()
        _check(lib().am_batched_result(self._h, self.RESULTS[name], ctypes.byref(out)))
        return MetalArray(out)

    @property
    def config(self):
        v = (ctypes.c_int64 * 7)()
        _check(lib().am_batched_config(self._h, v, 7))
        return {"tile_items": v[0], "hash_capacity": v[1], "strategy": ("lds", "global")[v[2]],
                "digit_bits": v[3], "regrouped_batches": v[4], "sort_tiles": v[5],
                "sort_path": ("lsd", "msd")[v[6]]}


class Server:
    """Request server with continuous batching (`am_server_*`): `slots` workers, each on its own HIP
    stream, admit requests in arrival order as slots and the device-memory budget free up."""

    OPS = {"group_by": 0, "argsort": 1, "unique": 2}

    def __init__(self, slots=4, budget_bytes=64 << 30):
        self._h = None
        h = _P()
        _check(lib().am_server_create(int(slots), int(budget_bytes), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.am_server_release(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def submit(self, op, columns, values=None, descending=False, null_placement="at_end", float_order="ieee"):
        """Queues a request; returns its ticket. op: "group_by" (columns + values), "argsort" (one
        column) or "unique" (columns)."""
        if not isinstance(columns, (list, tuple)):
            columns = [columns]
        handles = (_P * len(columns))(*[c._h for c in columns])
        t = ctypes.c_int64()
        _check(lib().am_server_submit(self._h, self.OPS[op], handles, len(columns),
                                      None if values is None else values._h, 1 if descending else 0,
                                      NULL_PLACEMENTS.index(null_placement), _float_order(float_order),
                                      ctypes.byref(t)))
        return t.value

    def wait(self, ticket, cap=8):
        """(results, (queued_ms, running_ms, latency_ms)) of a request, blocking until it is done."""
        outs = (_P * cap)()
        times = (ctypes.c_double * 3)()
        k = ctypes.c_int64()
        _check(lib().am_server_wait(self._h, int(ticket), outs, cap, times, ctypes.byref(k)))
        return [MetalArray(outs[i]) for i in range(min(k.value, cap))], tuple(times)

    @property
    def stats(self):
        v = (ctypes.c_int64 * 8)()
        _check(lib().am_server_stats(self._h, v, 8))
        names = ("submitted", "completed", "failed", "peak_admitted_bytes", "peak_running", "budget_bytes",
                 "queued", "slots")
        return dict(zip(names, list(v)))


class batch:
    """Context manager: calls issued inside queue on the thread's stream and complete at exit."""

    def __enter__(self):
        _check(lib().am_batch_begin())
        return self

    def __exit__(self, exc_type, exc, tb):
        _check(lib().am_batch_end())
        return False


def set_max_row_index(max_row):
    """Test hook (the reference's Dispatch.maxRowIndex): the largest row number index-returning calls
    may produce; None or an out-of-range value restores 2^32 - 1. Returns the previous limit."""
    return lib().am_set_max_row_index(-1 if max_row is None else int(max_row))


def prof_enable(on=True):
    lib().am_prof_enable(1 if on else 0)


def prof_reset():
    _check(lib().am_prof_reset())


def prof_report():
    """{kernel name: (launches, total_ms)} measured with HIP events on the launching stream."""
    need = lib().am_prof_report(None, 0)
    buf = ctypes.create_string_buffer(int(max(need, 1)))
    lib().am_prof_report(buf, need)
    out = {}
    for line in buf.value.decode().splitlines():
        name, count, ms = line.split()
        out[name] = (int(count), float(ms))
    return out

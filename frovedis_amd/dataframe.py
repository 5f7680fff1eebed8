"""Dataframe layer on one rank's share (the reference's
src/frovedis/dataframe): group_by and grouped aggregation, the fused hash
group-by, unique_hashtable and equi_join.  Each call goes to one C-ABI entry
point of libfrovedis_hip."""
import ctypes

import numpy as np

from . import TYPE_CODE, FhipError, check, lib
from .core import DeviceArray, _sz, current_stream

__all__ = ["group_by", "grouped_agg", "group_by_agg", "AGG_OPS", "UniqueHashtable",
           "equi_join", "JOIN_STRATEGY"]

AGG_OPS = {"sum": 0, "count": 1, "avg": 2, "max": 3, "min": 4}


def group_by(col, local_idx=None, stream=None):
    """typed_dfcolumn<T>::group_by on one rank (dataframe/dfcolumn_groupby.cc:
    16-120).  Returns (local_grouped_idx, local_idx_split, group_keys); the
    index array is permuted in place when given."""
    dev = col.t.device
    if local_idx is None:
        import torch
        local_idx = DeviceArray(torch.arange(len(col), dtype=torch.int64, device=dev), np.uint64)
    if local_idx.dtype != np.dtype(np.uint64):
        raise FhipError("group_by: local_idx must be size_t (uint64)")
    n = len(local_idx)
    if n and int(local_idx.t.max()) >= len(col):
        raise FhipError("group_by: local_idx points beyond the column")
    split = DeviceArray.empty(n + 1, np.uint64, dev)
    keys = DeviceArray.empty(max(n, 1), col.dtype, dev)
    ns = ctypes.c_size_t(0)
    check(lib().fhip_dev_group_by(col.code, col.ptr, local_idx.ptr, _sz(n), split.ptr,
                                  ctypes.byref(ns), keys.ptr, stream or current_stream()))
    G = ns.value - 1
    return local_idx, split[: ns.value], keys[:G]


def grouped_agg(op, col, grouped_idx, split, nulls=None, stream=None):
    """Grouped sum / count / avg / max / min over one column
    (dataframe/dfcolumn_grouped_agg.cc:105-889, one rank)."""
    if op not in AGG_OPS:
        raise FhipError(f"unknown aggregation {op!r}")
    G = max(len(split) - 1, 0)
    dev = grouped_idx.t.device
    odt = {"count": np.uint64, "avg": np.float64}.get(op, col.dtype)
    out = DeviceArray.empty(G, odt, dev)
    nn = 0 if nulls is None else len(nulls)
    check(lib().fhip_dev_grouped_agg(AGG_OPS[op], col.code, col.ptr, _sz(len(col)),
                                     grouped_idx.ptr, _sz(len(grouped_idx)), split.ptr,
                                     _sz(len(split)),
                                     nulls.ptr if nn else None, _sz(nn), out.ptr,
                                     stream or current_stream()))
    return out


class _Agg(ctypes.Structure):
    _fields_ = [("op", ctypes.c_int), ("vtype", ctypes.c_int),
                ("col", ctypes.c_void_p), ("out", ctypes.c_void_p)]


def group_by_agg(keys, aggs, table_hint=0, capacity=None, stream=None):
    """Fused group_by(keys).select({keys}, aggs) on the hash path.

    ``aggs`` is a list of (op, column) with op in sum/count/avg/max/min
    (column may be None for count).  Returns (group_keys, [outputs]) as
    device arrays, groups ascending by key like grouped_dftable::select
    (dftable.cc:3724-3772)."""
    n = len(keys)
    dev = keys.t.device
    cap = int(capacity) if capacity is not None else max(1, min(n, 1 << 22))
    for _ in range(2):
        kout = DeviceArray.empty(cap, keys.dtype, dev)
        outs = []
        specs = (_Agg * max(1, len(aggs)))()
        for i, (op, col) in enumerate(aggs):
            if op not in AGG_OPS:
                raise FhipError(f"unknown aggregation {op!r}")
            odt = {"count": np.uint64, "avg": np.float64}.get(op, None if col is None else col.dtype)
            out = DeviceArray.empty(cap, odt, dev)
            outs.append(out)
            specs[i].op = AGG_OPS[op]
            specs[i].vtype = col.code if col is not None else 0
            specs[i].col = col.ptr.value if col is not None else None
            specs[i].out = out.ptr.value
        ng = ctypes.c_size_t(0)
        rc = lib().fhip_dev_group_by_agg(keys.code, keys.ptr, _sz(n), specs, len(aggs), kout.ptr,
                                         _sz(cap), ctypes.byref(ng), _sz(table_hint),
                                         stream or current_stream())
        if rc == 6 and ng.value > cap:  # FHIP_ERR_TOO_LARGE: resize once
            cap = ng.value
            continue
        check(rc)
        G = ng.value
        return kout[:G], [o[:G] for o in outs]
    raise FhipError("group_by_agg: output did not fit after resizing")


class UniqueHashtable:
    """unique_hashtable<K, size_t> (dataframe/hashtable.hpp:57-110) in HBM.

    ``UniqueHashtable(keys, vals, check_unique=True)`` mirrors the
    (k, v, int& is_unique_ok) constructor; vals=None gives the set
    constructor (k)."""

    def __init__(self, keys, vals=None, check_unique=False, stream=None):
        self.dtype = keys.dtype
        self.n = len(keys)
        self._h = ctypes.c_void_p(0)
        flag = ctypes.c_int(1)
        check(lib().fhip_dev_ht_build(keys.code, keys.ptr, vals.ptr if vals is not None else None,
                                      _sz(self.n), int(bool(check_unique)), ctypes.byref(flag),
                                      ctypes.byref(self._h), stream or current_stream()))
        self.is_unique_ok = bool(flag.value)
        ts, nc = ctypes.c_size_t(0), ctypes.c_size_t(0)
        check(lib().fhip_ht_info(self._h, ctypes.byref(ts), ctypes.byref(nc)))
        self.table_size, self.conflict_size = ts.value, nc.value

    def _q(self, keys):
        if keys.dtype != self.dtype:
            raise FhipError("unique_hashtable: key type differs from the table's")

    def lookup(self, keys, stream=None):
        """(values, miss) like lookup(k, miss) (:1442-1515)."""
        self._q(keys)
        dev = keys.t.device
        vals = DeviceArray.empty(len(keys), np.uint64, dev)
        miss = DeviceArray.empty(len(keys), np.uint64, dev)
        nm = ctypes.c_size_t(0)
        check(lib().fhip_dev_ht_lookup(self._h, keys.ptr, _sz(len(keys)), vals.ptr, miss.ptr,
                                       ctypes.byref(nm), stream or current_stream()))
        return vals, miss[: nm.value]

    def check_existence(self, keys, stream=None):
        self._q(keys)
        out = DeviceArray.empty(len(keys), np.int32, keys.t.device)
        check(lib().fhip_dev_ht_check_existence(self._h, keys.ptr, _sz(len(keys)), out.ptr,
                                                stream or current_stream()))
        return out

    def all_keys(self, stream=None):
        import torch
        out = DeviceArray.empty(max(self.n, 1), self.dtype, torch.device("cuda"))
        m = ctypes.c_size_t(0)
        check(lib().fhip_dev_ht_all_keys(self._h, out.ptr, ctypes.byref(m),
                                         stream or current_stream()))
        return out[: m.value]

    def clear(self):
        if self._h:
            check(lib().fhip_ht_destroy(self._h))
            self._h = ctypes.c_void_p(0)

    def __del__(self):
        try:
            self.clear()
        except Exception:
            pass


JOIN_STRATEGY = {"auto": 0, "hash": 1, "sortmerge": 2}


def equi_join(left, left_idx, right, right_idx, strategy="auto", capacity=None, stream=None):
    """equi_join<T> (dataframe/join.hpp:304-326): returns
    (left_idx_out, right_idx_out, used_hash)."""
    if left.dtype != right.dtype:
        raise FhipError("column types are different")
    dev = left.t.device
    cap = int(capacity) if capacity is not None else max(1, len(left))
    for _ in range(2):
        lo = DeviceArray.empty(cap, np.uint64, dev)
        ro = DeviceArray.empty(cap, np.uint64, dev)
        nout = ctypes.c_size_t(0)
        used = ctypes.c_int(0)
        rc = lib().fhip_dev_equi_join(left.code, left.ptr, left_idx.ptr, _sz(len(left)),
                                      right.ptr, right_idx.ptr, _sz(len(right)),
                                      JOIN_STRATEGY[strategy], lo.ptr, ro.ptr, _sz(cap),
                                      ctypes.byref(nout), ctypes.byref(used),
                                      stream or current_stream())
        if rc == 6 and nout.value > cap:
            cap = nout.value
            continue
        check(rc)
        m = nout.value
        return lo[:m], ro[:m], bool(used.value)
    raise FhipError("equi_join: output did not fit after resizing")

"""Dataframe layer on one rank's share (the reference's
src/frovedis/dataframe): group_by and grouped aggregation, the fused hash
group-by, unique_hashtable and equi_join.  Each call goes to one C-ABI entry
point of libfrovedis_hip."""
import ctypes

import numpy as np

from . import TYPE_CODE, FhipError, check, lib
from .core import DeviceArray, _sz, current_stream

__all__ = ["group_by", "grouped_agg", "group_by_agg", "AGG_OPS"]

AGG_OPS = {"sum": 0, "count": 1, "avg": 2, "max": 3, "min": 4}


def group_by(col, local_idx=None, stream=None):
    """typed_dfcolumn<T>::group_by on one rank (dataframe/dfcolumn_groupby.cc:
    16-120).  Returns (local_grouped_idx, local_idx_split, group_keys); the
    index array is permuted in place when given."""
    n = len(col)
    dev = col.t.device
    if local_idx is None:
        import torch
        local_idx = DeviceArray(torch.arange(n, dtype=torch.int64, device=dev), np.uint64)
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
    check(lib().fhip_dev_grouped_agg(AGG_OPS[op], col.code, col.ptr, grouped_idx.ptr,
                                     _sz(len(grouped_idx)), split.ptr, _sz(len(split)),
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

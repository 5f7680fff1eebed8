"""Dataframe layer on one rank's share (the reference's
src/frovedis/dataframe): group_by over one or several columns and grouped
aggregation, the fused hash group-by, unique_hashtable, equi_join and the
outer join, and the column operations around them (global_extract with
NULLs, filter, sort).  Each call goes to one C-ABI entry point of
libfrovedis_hip."""
import ctypes

import numpy as np

from . import TYPE_CODE, FhipError, check, lib
from .core import DeviceArray, _sz, current_stream

__all__ = ["group_by", "grouped_agg", "group_by_agg", "AGG_OPS", "UniqueHashtable",
           "equi_join", "outer_equi_join", "star_join", "JOIN_STRATEGY", "NULL_INDEX", "global_extract",
           "extract_nulls", "filter", "CMP_OPS", "sort", "sort_desc", "multi_group_by"]

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


def group_by_agg(keys, aggs, table_hint=0, capacity=None, stream=None, model=None):
    """Fused group_by(keys).select({keys}, aggs) on the hash path.

    ``aggs`` is a list of (op, column) with op in sum/count/avg/max/min
    (column may be None for count).  Returns (group_keys, [outputs]) as
    device arrays, groups ascending by key like grouped_dftable::select
    (dftable.cc:3724-3772).  ``model``: a CdfModel of the key column
    (trained on an earlier input): its cardinality estimate sizes the
    partitioned hash tables and its heavy keys fill the heavy-key table
    (fhip_dev_group_by_agg_model), in place of ``table_hint`` and of the
    sample the library would count; the result is the same."""
    n = len(keys)
    dev = keys.t.device
    cap = int(capacity) if capacity is not None else max(1, min(n, 1 << 22))
    odts = []
    for op, col in aggs:
        if op not in AGG_OPS:
            raise FhipError(f"unknown aggregation {op!r}")
        odts.append({"count": np.uint64, "avg": np.float64}.get(op, None if col is None else col.dtype))
    for _ in range(2):
        # the group keys and every output in one allocation
        kout, *outs = DeviceArray.empty_many(cap, [keys.dtype] + odts, dev)
        specs = (_Agg * max(1, len(aggs)))()
        for i, (op, col) in enumerate(aggs):
            out = outs[i]
            specs[i].op = AGG_OPS[op]
            specs[i].vtype = col.code if col is not None else 0
            specs[i].col = col.ptr.value if col is not None else None
            specs[i].out = out.ptr.value
        ng = ctypes.c_size_t(0)
        if model is not None:
            rc = lib().fhip_dev_group_by_agg_model(keys.code, keys.ptr, _sz(n), specs, len(aggs),
                                                   kout.ptr, _sz(cap), ctypes.byref(ng),
                                                   model._m, stream or current_stream())
        else:
            rc = lib().fhip_dev_group_by_agg(keys.code, keys.ptr, _sz(n), specs, len(aggs),
                                             kout.ptr, _sz(cap), ctypes.byref(ng),
                                             _sz(table_hint), stream or current_stream())
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


JOIN_STRATEGY = {"auto": 0, "hash": 1, "sortmerge": 2, "auto_unique": 3}


def equi_join(left, left_idx, right, right_idx, strategy="auto", capacity=None, stream=None,
              model=None):
    """equi_join<T> (dataframe/join.hpp:304-326): returns
    (left_idx_out, right_idx_out, used_hash).  With a CdfModel of the right
    key column (trained on an earlier input) and strategy "auto", the model
    picks the AUTO variant (fhip_learned_join_strategy); the result is the
    same."""
    if left.dtype != right.dtype:
        raise FhipError("column types are different")
    if model is not None and strategy == "auto":
        strategy = model.join_strategy(len(right))
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


def outer_equi_join(left, left_idx, right, right_idx, strategy="auto", capacity=None,
                    stream=None):
    """outer_equi_join<T> (dataframe/join.hpp:328-379): returns
    (left_idx_out, right_idx_out, only_left, used_hash); only_left are the
    left row ids without a match, whose right columns are NULL."""
    if left.dtype != right.dtype:
        raise FhipError("column types are different")
    dev = left.t.device
    cap = int(capacity) if capacity is not None else max(1, len(left))
    only = DeviceArray.empty(max(1, len(left)), np.uint64, dev)
    for _ in range(2):
        lo = DeviceArray.empty(cap, np.uint64, dev)
        ro = DeviceArray.empty(cap, np.uint64, dev)
        nout = ctypes.c_size_t(0)
        nonly = ctypes.c_size_t(0)
        used = ctypes.c_int(0)
        rc = lib().fhip_dev_outer_equi_join(left.code, left.ptr, left_idx.ptr, _sz(len(left)),
                                            right.ptr, right_idx.ptr, _sz(len(right)),
                                            JOIN_STRATEGY[strategy], lo.ptr, ro.ptr, _sz(cap),
                                            ctypes.byref(nout), only.ptr, ctypes.byref(nonly),
                                            ctypes.byref(used), stream or current_stream())
        if rc == 6 and nout.value > cap:
            cap = nout.value
            continue
        check(rc)
        m = nout.value
        return lo[:m], ro[:m], only[: nonly.value], bool(used.value)
    raise FhipError("outer_equi_join: output did not fit after resizing")


NULL_INDEX = (1 << 64) - 1
"""A row index that selects NULL (numeric_limits<T>::max()) in
global_extract: an outer join's missing right row."""


def global_extract(col, idx, stream=None):
    """col[idx] on one rank (global_extract, dfcolumn_misc.cc:714-761, with
    append_nulls :1052-1078 for indices equal to NULL_INDEX)."""
    dev = idx.t.device
    out = DeviceArray.empty(len(idx), col.dtype, dev)
    check(lib().fhip_dev_gather(col.code, col.ptr, _sz(len(col)), idx.ptr, _sz(len(idx)),
                                out.ptr, stream or current_stream()))
    return out


def extract_nulls(idx, src_nulls=None, src_len=None, stream=None):
    """Null positions of global_extract(col, idx): where idx is NULL_INDEX
    or points at a null row of the source (dfcolumn.cc:223-233)."""
    dev = idx.t.device
    out = DeviceArray.empty(max(1, len(idx)), np.uint64, dev)
    nn = 0 if src_nulls is None else len(src_nulls)
    if nn and src_len is None:
        raise FhipError("extract_nulls: src_len is needed with src_nulls")
    nout = ctypes.c_size_t(0)
    check(lib().fhip_dev_gather_nulls(idx.ptr, _sz(len(idx)), _sz(src_len or 0),
                                      src_nulls.ptr if nn else None, _sz(nn), out.ptr,
                                      ctypes.byref(nout), stream or current_stream()))
    return out[: nout.value]


CMP_OPS = {"eq": 0, "neq": 1, "lt": 2, "le": 3, "gt": 4, "ge": 5}


def filter(col, op, other, nulls=None, other_nulls=None, stream=None):
    """filter_<op> (other a DeviceArray of the same length) or
    filter_<op>_immed (other a numpy scalar; its dtype is the scalar's C++
    type) on one rank (dataframe/dfcolumn_filter_*.cc): the ascending row
    positions where the comparison holds, null rows removed."""
    if op not in CMP_OPS:
        raise FhipError(f"unknown comparison {op!r}")
    n = len(col)
    dev = col.t.device
    out = DeviceArray.empty(max(1, n), np.uint64, dev)
    nout = ctypes.c_size_t(0)
    nn = 0 if nulls is None else len(nulls)
    if isinstance(other, DeviceArray):
        if len(other) != n:
            raise FhipError("filter: columns of different sizes")
        on = 0 if other_nulls is None else len(other_nulls)
        check(lib().fhip_dev_filter_pair(col.code, col.ptr, other.code, other.ptr, _sz(n),
                                         CMP_OPS[op], nulls.ptr if nn else None, _sz(nn),
                                         other_nulls.ptr if on else None, _sz(on), out.ptr,
                                         ctypes.byref(nout), stream or current_stream()))
    else:
        sc = np.asarray(other)
        if sc.dtype not in TYPE_CODE:
            raise FhipError(f"unsupported scalar type {sc.dtype}")
        buf = np.array([sc], dtype=sc.dtype)
        check(lib().fhip_dev_filter(col.code, col.ptr, _sz(n), CMP_OPS[op], TYPE_CODE[sc.dtype],
                                    buf.ctypes.data_as(ctypes.c_void_p),
                                    nulls.ptr if nn else None, _sz(nn), out.ptr,
                                    ctypes.byref(nout), stream or current_stream()))
    return out[: nout.value]


def _sort(col, idx, descending, stream):
    if idx is not None and len(idx) and int(idx.t.max()) >= len(col):
        raise FhipError("sort: index points beyond the column")
    n = len(col) if idx is None else len(idx)
    dev = col.t.device
    keys = DeviceArray.empty(n, col.dtype, dev)
    out = DeviceArray.empty(n, np.uint64, dev)
    check(lib().fhip_dev_sort_index(col.code, col.ptr, _sz(len(col)),
                                    idx.ptr if idx is not None else None, _sz(n),
                                    int(descending), keys.ptr, out.ptr,
                                    stream or current_stream()))
    return keys, out


def sort(col, idx=None, stream=None):
    """dftable::sort (idx None: rows 0..n-1) or sorted_dftable::sort (idx:
    the current order) on one rank (dfcolumn_sort.cc:141-180, :299-340):
    returns (sorted_keys, row_order).  Stable, so sorting by c2 and then by
    c1 orders by (c1, c2)."""
    return _sort(col, idx, False, stream)


def sort_desc(col, idx=None, stream=None):
    """sort_desc (dfcolumn_sort.cc:254-296): stable descending order."""
    return _sort(col, idx, True, stream)


def multi_group_by(cols, local_idx=None, stream=None):
    """dftable::group_by over several key columns on one rank (dftable.cc:
    367-487): returns (local_grouped_idx, local_idx_split, [group keys of
    every column]); groups in lexicographic key order (column 0 first).
    The result feeds grouped_agg unchanged."""
    if not cols:
        raise FhipError("multi_group_by: no key column")
    dev = cols[0].t.device
    n0 = len(cols[0])
    if any(len(c) != n0 for c in cols):
        raise FhipError("multi_group_by: columns of different sizes")
    if local_idx is None:
        import torch
        local_idx = DeviceArray(torch.arange(n0, dtype=torch.int64, device=dev), np.uint64)
    if local_idx.dtype != np.dtype(np.uint64):
        raise FhipError("multi_group_by: local_idx must be size_t (uint64)")
    n = len(local_idx)
    if n and int(local_idx.t.max()) >= n0:
        raise FhipError("multi_group_by: local_idx points beyond the columns")
    split = DeviceArray.empty(n + 1, np.uint64, dev)
    keys = [DeviceArray.empty(max(n, 1), c.dtype, dev) for c in cols]
    k = len(cols)
    kt = (ctypes.c_int * k)(*[c.code for c in cols])
    cp = (ctypes.c_void_p * k)(*[c.ptr.value for c in cols])
    kp = (ctypes.c_void_p * k)(*[x.ptr.value for x in keys])
    ns = ctypes.c_size_t(0)
    check(lib().fhip_dev_multi_group_by(k, kt, cp, local_idx.ptr, _sz(n), split.ptr,
                                        ctypes.byref(ns), kp, stream or current_stream()))
    G = ns.value - 1
    return local_idx, split[: ns.value], [x[:G] for x in keys]


def star_join(left_cols, rights, left_idx=None, stream=None):
    """star_join on one rank (dftable.cc:295-333): left_cols[d] is joined
    to dimension d, rights[d] = (right key column, right ids) with unique
    keys.  Returns (left_rows, [right ids per dimension]) for the left rows
    found in every dimension, in left order."""
    nd = len(left_cols)
    if nd == 0 or nd != len(rights):
        raise FhipError("star_join: one right table per left key column")
    for lc, (rc, ri) in zip(left_cols, rights):
        if lc.dtype != rc.dtype:
            raise FhipError("star_join_eq: column types are different")
        if len(rc) != len(ri):
            raise FhipError("star_join: right keys and ids of different sizes")
    dev = left_cols[0].t.device
    nl = len(left_cols[0]) if left_idx is None else len(left_idx)
    if left_idx is not None and nl and int(left_idx.t.max()) >= min(len(c) for c in left_cols):
        raise FhipError("star_join: left index points beyond a column")
    lo = DeviceArray.empty(max(nl, 1), np.uint64, dev)
    ros = [DeviceArray.empty(max(nl, 1), np.uint64, dev) for _ in range(nd)]
    P = ctypes.c_void_p
    nout = ctypes.c_size_t(0)
    check(lib().fhip_dev_star_join(
        nd, (ctypes.c_int * nd)(*[c.code for c in left_cols]),
        (P * nd)(*[c.ptr.value for c in left_cols]),
        left_idx.ptr if left_idx is not None else None, _sz(nl),
        (P * nd)(*[r.ptr.value for r, _ in rights]), (P * nd)(*[i.ptr.value for _, i in rights]),
        (ctypes.c_size_t * nd)(*[len(r) for r, _ in rights]), lo.ptr,
        (P * nd)(*[r.ptr.value for r in ros]), ctypes.byref(nout), stream or current_stream()))
    m = nout.value
    return lo[:m], [r[:m] for r in ros]

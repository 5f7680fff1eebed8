"""CPU oracle (test infrastructure only).

numpy front end over ``liboracle.so`` (built from ``oracle/oracle.cpp``).  It
restates the reference's x86 algorithms for the dataframe execution core
(radix_sort, prefix_sum, set_separate, lower_bound, unique_hashtable,
equi_join, group_by and the grouped aggregates).  Only ``tests/``,
``__graft_entry__.smoke()`` and the cpu_baseline leg of ``bench.py`` use it,
as the checker; the product library never calls it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

TYPE_CODE = {np.dtype(np.int32): 0, np.dtype(np.uint32): 1,
             np.dtype(np.int64): 2, np.dtype(np.uint64): 3,
             np.dtype(np.float32): 4, np.dtype(np.float64): 5}
OP_CODE = {"sum": 0, "count": 1, "avg": 2, "max": 3, "min": 4}

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-f", os.path.join(_HERE, "Makefile")])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        vp, sz, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        L.orc_radix_sort.argtypes = [i, vp, i, vp, sz, i, i]
        L.orc_prefix_sum.argtypes = [i, vp, vp, sz]
        L.orc_set_separate.argtypes = [i, vp, sz, vp]
        L.orc_set_separate.restype = sz
        L.orc_set_is_unique.argtypes = [i, vp, sz]
        L.orc_lower_bound.argtypes = [i, vp, sz, vp, sz, vp, i]
        L.orc_ht_build.argtypes = [i, vp, vp, sz, i, ctypes.POINTER(i)]
        L.orc_ht_build.restype = vp
        L.orc_ht_table_size.argtypes = [vp]
        L.orc_ht_table_size.restype = sz
        L.orc_ht_conflict_size.argtypes = [vp]
        L.orc_ht_conflict_size.restype = sz
        L.orc_ht_lookup.argtypes = [vp, vp, sz, vp, vp]
        L.orc_ht_lookup.restype = sz
        L.orc_ht_check_existence.argtypes = [vp, vp, sz, vp]
        L.orc_ht_all_keys.argtypes = [vp, vp]
        L.orc_ht_all_keys.restype = sz
        L.orc_ht_free.argtypes = [vp]
        L.orc_equi_join.argtypes = [i, vp, vp, sz, vp, vp, sz, i,
                                    ctypes.POINTER(ctypes.POINTER(ctypes.c_size_t)),
                                    ctypes.POINTER(ctypes.POINTER(ctypes.c_size_t)),
                                    ctypes.POINTER(i)]
        L.orc_equi_join.restype = sz
        L.orc_free.argtypes = [vp]
        L.orc_group_by.argtypes = [i, vp, vp, sz, vp, vp]
        L.orc_group_by.restype = sz
        L.orc_grouped_agg.argtypes = [i, i, vp, vp, sz, vp, sz, vp, sz, vp]
        L.orc_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def radix_sort(keys, vals=None, positive_only=False, descending=False):
    """In-place stable sort like ``radix_sort(K*, V*, size_t, bool)``
    (core/radix_sort.hpp:197-219); returns (keys, vals) copies."""
    k = np.ascontiguousarray(keys).copy()
    v = None if vals is None else np.ascontiguousarray(vals).copy()
    vb = 0 if v is None else v.dtype.itemsize
    if v is not None and len(v) != len(k):
        raise RuntimeError("radix_sort: different size of arrays")
    rc = lib().orc_radix_sort(TYPE_CODE[k.dtype], _p(k), vb, _p(v), len(k),
                              int(positive_only), int(descending))
    assert rc == 0
    return k, v


def prefix_sum(a):
    a = np.ascontiguousarray(a)
    out = np.empty_like(a)
    lib().orc_prefix_sum(TYPE_CODE[a.dtype], _p(a), _p(out), len(a))
    return out


def set_separate(keys):
    k = np.ascontiguousarray(keys)
    out = np.empty(len(k) + 1, np.uint64)
    m = lib().orc_set_separate(TYPE_CODE[k.dtype], _p(k), len(k), _p(out))
    return out[:m].copy()


def set_is_unique(keys):
    k = np.ascontiguousarray(keys)
    return bool(lib().orc_set_is_unique(TYPE_CODE[k.dtype], _p(k), len(k)))


def lower_bound(sorted_keys, queries, upper=False):
    s = np.ascontiguousarray(sorted_keys)
    q = np.ascontiguousarray(queries, dtype=s.dtype)
    out = np.empty(len(q), np.uint64)
    lib().orc_lower_bound(TYPE_CODE[s.dtype], _p(s), len(s), _p(q), len(q),
                          _p(out), int(upper))
    return out


class UniqueHashtable:
    """``unique_hashtable<K, size_t>`` (dataframe/hashtable.hpp:57-110)."""

    def __init__(self, keys, vals=None, check_unique=False):
        self.k = np.ascontiguousarray(keys)
        v = None if vals is None else np.ascontiguousarray(vals, dtype=np.uint64)
        flag = ctypes.c_int(1)
        self.h = lib().orc_ht_build(TYPE_CODE[self.k.dtype], _p(self.k), _p(v),
                                    len(self.k), int(check_unique),
                                    ctypes.byref(flag))
        if not self.h:
            raise RuntimeError(lib().orc_last_error().decode())
        self.is_unique_ok = bool(flag.value)

    @property
    def table_size(self):
        return lib().orc_ht_table_size(self.h)

    @property
    def conflict_size(self):
        return lib().orc_ht_conflict_size(self.h)

    def lookup(self, keys):
        q = np.ascontiguousarray(keys, dtype=self.k.dtype)
        out = np.empty(len(q), np.uint64)
        miss = np.empty(len(q), np.uint64)
        m = lib().orc_ht_lookup(self.h, _p(q), len(q), _p(out), _p(miss))
        return out, miss[:m].copy()

    def check_existence(self, keys):
        q = np.ascontiguousarray(keys, dtype=self.k.dtype)
        out = np.empty(len(q), np.int32)
        lib().orc_ht_check_existence(self.h, _p(q), len(q), _p(out))
        return out

    def all_keys(self):
        out = np.empty(len(self.k), self.k.dtype)
        m = lib().orc_ht_all_keys(self.h, _p(out))
        return out[:m].copy()

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_ht_free(self.h)
            self.h = None


def equi_join(left, left_idx, right, right_idx, algo="ve"):
    """(left_idx_out, right_idx_out, used_hash).  algo "ve" follows
    dataframe/join.hpp:304-326, "x86" follows :25-45."""
    l = np.ascontiguousarray(left)
    r = np.ascontiguousarray(right, dtype=l.dtype)
    li = np.ascontiguousarray(left_idx, dtype=np.uint64)
    ri = np.ascontiguousarray(right_idx, dtype=np.uint64)
    lo = ctypes.POINTER(ctypes.c_size_t)()
    ro = ctypes.POINTER(ctypes.c_size_t)()
    used = ctypes.c_int(0)
    m = lib().orc_equi_join(TYPE_CODE[l.dtype], _p(l), _p(li), len(l), _p(r),
                            _p(ri), len(r), 1 if algo == "x86" else 0,
                            ctypes.byref(lo), ctypes.byref(ro), ctypes.byref(used))
    a = np.ctypeslib.as_array(lo, shape=(max(m, 1),))[:m].astype(np.uint64)
    b = np.ctypeslib.as_array(ro, shape=(max(m, 1),))[:m].astype(np.uint64)
    lib().orc_free(ctypes.cast(lo, ctypes.c_void_p))
    lib().orc_free(ctypes.cast(ro, ctypes.c_void_p))
    return a, b, bool(used.value)


def group_by(col, idx=None):
    """(grouped_idx, split, group_keys): dataframe/dfcolumn_groupby.cc:16-105
    on one rank."""
    c = np.ascontiguousarray(col)
    gi = (np.arange(len(c), dtype=np.uint64) if idx is None
          else np.ascontiguousarray(idx, dtype=np.uint64).copy())
    if len(gi) and int(gi.max()) >= len(c):
        raise IndexError("group_by: index beyond the column")
    n = len(gi)
    split = np.empty(n + 1, np.uint64)
    keys = np.empty(max(n, 1), c.dtype)
    m = lib().orc_group_by(TYPE_CODE[c.dtype], _p(c), _p(gi), n, _p(split),
                           _p(keys))
    return gi, split[:m].copy(), keys[:m - 1].copy()


def grouped_agg(op, col, grouped_idx, split, nulls=None):
    c = np.ascontiguousarray(col)
    gi = np.ascontiguousarray(grouped_idx, dtype=np.uint64)
    sp = np.ascontiguousarray(split, dtype=np.uint64)
    nu = (np.zeros(0, np.uint64) if nulls is None
          else np.ascontiguousarray(nulls, dtype=np.uint64))
    G = max(len(sp) - 1, 0)
    if op == "count":
        out = np.empty(G, np.uint64)
    elif op == "avg":
        out = np.empty(G, np.float64)
    else:
        out = np.empty(G, c.dtype)
    rc = lib().orc_grouped_agg(OP_CODE[op], TYPE_CODE[c.dtype], _p(c), _p(gi),
                               len(gi), _p(sp), len(sp),
                               _p(nu) if len(nu) else None, len(nu), _p(out))
    assert rc == 0
    return out

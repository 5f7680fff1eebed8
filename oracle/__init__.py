"""CPU oracle (test infrastructure only).

numpy front end over ``liboracle.so`` (built from ``oracle/oracle.cpp``).  It
restates the reference's x86 algorithms for the dataframe execution core
(radix_sort, prefix_sum, set_separate, lower_bound, unique_hashtable,
equi_join and the outer join, group_by over one or several columns, the
grouped aggregates, and the gather / filter / sort of dftable columns).  Only ``tests/``,
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
        psz = ctypes.POINTER(ctypes.POINTER(ctypes.c_size_t))
        L.orc_outer_equi_join.argtypes = [i, vp, vp, sz, vp, vp, sz, i, psz, psz, psz,
                                          ctypes.POINTER(sz), ctypes.POINTER(i)]
        L.orc_outer_equi_join.restype = sz
        L.orc_gather.argtypes = [i, vp, sz, vp, sz, vp]
        L.orc_gather_nulls.argtypes = [vp, sz, vp, sz, vp]
        L.orc_gather_nulls.restype = sz
        L.orc_filter.argtypes = [i, vp, sz, i, i, vp, vp, sz, vp]
        L.orc_filter.restype = sz
        L.orc_filter_pair.argtypes = [i, vp, i, vp, sz, i, vp, sz, vp, sz, vp]
        L.orc_filter_pair.restype = sz
        L.orc_sort_index.argtypes = [i, vp, vp, sz, i, vp, vp]
        L.orc_multi_group_by.argtypes = [i, vp, vp, vp, sz, vp, vp]
        L.orc_multi_group_by.restype = sz
        L.orc_star_join.argtypes = [i, vp, vp, vp, sz, vp, vp, vp, vp, vp]
        L.orc_star_join.restype = sz
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


JOIN_ALGO = {"ve": 0, "x86": 1, "ve_sortmerge": 2}


def equi_join(left, left_idx, right, right_idx, algo="ve"):
    """(left_idx_out, right_idx_out, used_hash).  algo "ve" follows
    dataframe/join.hpp:304-326, "x86" follows :25-45, "ve_sortmerge" the
    sort-merge branch of "ve"."""
    l = np.ascontiguousarray(left)
    r = np.ascontiguousarray(right, dtype=l.dtype)
    li = np.ascontiguousarray(left_idx, dtype=np.uint64)
    ri = np.ascontiguousarray(right_idx, dtype=np.uint64)
    lo = ctypes.POINTER(ctypes.c_size_t)()
    ro = ctypes.POINTER(ctypes.c_size_t)()
    used = ctypes.c_int(0)
    m = lib().orc_equi_join(TYPE_CODE[l.dtype], _p(l), _p(li), len(l), _p(r),
                            _p(ri), len(r), JOIN_ALGO[algo],
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


def _take(ptr, m):
    a = np.ctypeslib.as_array(ptr, shape=(max(m, 1),))[:m].astype(np.uint64)
    lib().orc_free(ctypes.cast(ptr, ctypes.c_void_p))
    return a


def outer_equi_join(left, left_idx, right, right_idx, algo="ve"):
    """(left_idx_out, right_idx_out, only_left, used_hash): outer_equi_join
    (dataframe/join.hpp:328-379 "ve", :73-99 "x86")."""
    l = np.ascontiguousarray(left)
    r = np.ascontiguousarray(right, dtype=l.dtype)
    li = np.ascontiguousarray(left_idx, dtype=np.uint64)
    ri = np.ascontiguousarray(right_idx, dtype=np.uint64)
    P = ctypes.POINTER(ctypes.c_size_t)
    lo, ro, ol = P(), P(), P()
    no = ctypes.c_size_t(0)
    used = ctypes.c_int(0)
    m = lib().orc_outer_equi_join(TYPE_CODE[l.dtype], _p(l), _p(li), len(l), _p(r), _p(ri),
                                  len(r), JOIN_ALGO[algo], ctypes.byref(lo), ctypes.byref(ro),
                                  ctypes.byref(ol), ctypes.byref(no), ctypes.byref(used))
    return _take(lo, m), _take(ro, m), _take(ol, no.value), bool(used.value)


NULL_INDEX = np.uint64(2**64 - 1)
CMP_CODE = {"eq": 0, "neq": 1, "lt": 2, "le": 3, "gt": 4, "ge": 5}


def gather(src, idx):
    """src[idx] with NULL (max of the type) where idx == NULL_INDEX."""
    s = np.ascontiguousarray(src)
    ix = np.ascontiguousarray(idx, dtype=np.uint64)
    out = np.empty(len(ix), s.dtype)
    if lib().orc_gather(TYPE_CODE[s.dtype], _p(s), len(s), _p(ix), len(ix), _p(out)):
        raise IndexError(lib().orc_last_error().decode())
    return out


def gather_nulls(idx, src_nulls=None):
    ix = np.ascontiguousarray(idx, dtype=np.uint64)
    nu = np.zeros(0, np.uint64) if src_nulls is None else np.ascontiguousarray(src_nulls,
                                                                                 dtype=np.uint64)
    out = np.empty(max(len(ix), 1), np.uint64)
    m = lib().orc_gather_nulls(_p(ix), len(ix), _p(nu) if len(nu) else None, len(nu), _p(out))
    return out[:m].copy()


def filter(col, op, scalar, nulls=None):
    """Row positions where ``col op scalar`` (scalar: a numpy scalar whose
    dtype is the scalar's C++ type), nulls removed."""
    c = np.ascontiguousarray(col)
    sc = np.array([scalar], dtype=np.asarray(scalar).dtype)
    nu = np.zeros(0, np.uint64) if nulls is None else np.ascontiguousarray(nulls, dtype=np.uint64)
    out = np.empty(max(len(c), 1), np.uint64)
    m = lib().orc_filter(TYPE_CODE[c.dtype], _p(c), len(c), CMP_CODE[op], TYPE_CODE[sc.dtype],
                         _p(sc), _p(nu) if len(nu) else None, len(nu), _p(out))
    if m == 2**64 - 1:
        raise RuntimeError(lib().orc_last_error().decode())
    return out[:m].copy()


def filter_pair(a, b, op, a_nulls=None, b_nulls=None):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    an = np.zeros(0, np.uint64) if a_nulls is None else np.ascontiguousarray(a_nulls,
                                                                             dtype=np.uint64)
    bn = np.zeros(0, np.uint64) if b_nulls is None else np.ascontiguousarray(b_nulls,
                                                                             dtype=np.uint64)
    out = np.empty(max(len(a), 1), np.uint64)
    m = lib().orc_filter_pair(TYPE_CODE[a.dtype], _p(a), TYPE_CODE[b.dtype], _p(b), len(a),
                              CMP_CODE[op], _p(an) if len(an) else None, len(an),
                              _p(bn) if len(bn) else None, len(bn), _p(out))
    if m == 2**64 - 1:
        raise RuntimeError(lib().orc_last_error().decode())
    return out[:m].copy()


def sort_index(col, idx=None, descending=False):
    """(sorted_keys, idx_out): dftable::sort / sort_desc (idx None) or
    sorted_dftable::sort (idx = the current order)."""
    c = np.ascontiguousarray(col)
    ix = None if idx is None else np.ascontiguousarray(idx, dtype=np.uint64)
    n = len(c) if ix is None else len(ix)
    if ix is not None and n and int(ix.max()) >= len(c):
        raise IndexError("sort_index: index beyond the column")
    keys = np.empty(max(n, 1), c.dtype)
    out = np.empty(max(n, 1), np.uint64)
    rc = lib().orc_sort_index(TYPE_CODE[c.dtype], _p(c), _p(ix) if ix is not None else None, n,
                              int(descending), _p(keys), _p(out))
    assert rc == 0
    return keys[:n].copy(), out[:n].copy()


def multi_group_by(cols, idx=None):
    """(grouped_idx, split, [group keys per column]): dftable::group_by over
    several columns on one rank."""
    cs = [np.ascontiguousarray(c) for c in cols]
    n0 = len(cs[0])
    gi = (np.arange(n0, dtype=np.uint64) if idx is None
          else np.ascontiguousarray(idx, dtype=np.uint64).copy())
    n = len(gi)
    split = np.empty(n + 1, np.uint64)
    keys = [np.empty(max(n, 1), c.dtype) for c in cs]
    kt = (ctypes.c_int * len(cs))(*[TYPE_CODE[c.dtype] for c in cs])
    cp = (ctypes.c_void_p * len(cs))(*[c.ctypes.data for c in cs])
    kp = (ctypes.c_void_p * len(cs))(*[k.ctypes.data for k in keys])
    m = lib().orc_multi_group_by(len(cs), kt, cp, _p(gi), n, _p(split), kp)
    return gi, split[:m].copy(), [k[:max(m - 1, 0)].copy() for k in keys]


def star_join(left_cols, rights, left_idx=None):
    """(left_rows, [right ids per dimension]): star_join on one rank.
    rights: list of (right key column, right ids)."""
    lcs = [np.ascontiguousarray(c) for c in left_cols]
    rcs = [np.ascontiguousarray(r, dtype=l.dtype) for l, (r, _) in zip(lcs, rights)]
    ris = [np.ascontiguousarray(i, dtype=np.uint64) for _, i in rights]
    nd = len(lcs)
    li = None if left_idx is None else np.ascontiguousarray(left_idx, dtype=np.uint64)
    nl = len(lcs[0]) if li is None else len(li)
    lo = np.empty(max(nl, 1), np.uint64)
    ros = [np.empty(max(nl, 1), np.uint64) for _ in range(nd)]
    P = ctypes.c_void_p
    m = lib().orc_star_join(nd, (ctypes.c_int * nd)(*[TYPE_CODE[c.dtype] for c in lcs]),
                            (P * nd)(*[c.ctypes.data for c in lcs]),
                            _p(li) if li is not None else None, nl,
                            (P * nd)(*[r.ctypes.data if len(r) else None for r in rcs]),
                            (P * nd)(*[r.ctypes.data if len(r) else None for r in ris]),
                            (ctypes.c_size_t * nd)(*[len(r) for r in rcs]), _p(lo),
                            (P * nd)(*[r.ctypes.data for r in ros]))
    if m == 2**64 - 1:
        raise RuntimeError(lib().orc_last_error().decode())
    return lo[:m].copy(), [r[:m].copy() for r in ros]

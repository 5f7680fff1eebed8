"""Pins the oracle's column operations (gather with NULLs, filter, sort,
multi-column group_by, outer and star joins) before the GPU is checked
against them.

The checks come from outside the oracle's own code: numpy take / isin /
stable argsort, comparisons evaluated in the common type C++ picks for the
operands (the usual arithmetic conversions, restated here from the C++
standard), Python tuple sorting, brute-force join pairs, and the reference
tutorial's known answers (doc/tutorial/tutorial.md:2321-2341 sort,
:2456-2466 outer join, :2535-2552 multi-column group_by; the data of
doc/tutorial/src/tut5.3-5.6).
"""
import itertools

import numpy as np
import pytest

import oracle

ALL_TYPES = [np.int32, np.uint32, np.int64, np.uint64, np.float32, np.float64]
NULL = oracle.NULL_INDEX


def _col(rng, dt, n, lo=-40, hi=40):
    dt = np.dtype(dt)
    v = rng.integers(lo, hi, n)
    if dt.kind == "u":
        v = np.abs(v)
    if dt.kind == "f":
        return (v * 0.5).astype(dt)
    return v.astype(dt)


# ---- gather / gather_nulls ---------------------------------------------------

@pytest.mark.parametrize("dt", ALL_TYPES)
def test_gather_is_take_with_max_for_null(dt):
    rng = np.random.default_rng(1)
    src = _col(rng, dt, 500)
    idx = rng.integers(0, 500, 1000).astype(np.uint64)
    idx[::7] = NULL
    got = oracle.gather(src, idx)
    nz = idx != NULL
    np.testing.assert_array_equal(got[nz], np.take(src, idx[nz].astype(np.int64)))
    mx = np.finfo(dt).max if np.dtype(dt).kind == "f" else np.iinfo(dt).max
    assert (got[~nz] == mx).all()


def test_gather_rejects_out_of_range():
    with pytest.raises(IndexError):
        oracle.gather(np.arange(4, dtype=np.int64), np.array([1, 4], np.uint64))


def test_gather_nulls_is_isin():
    rng = np.random.default_rng(2)
    idx = rng.integers(0, 300, 2000).astype(np.uint64)
    idx[::11] = NULL
    nulls = np.unique(rng.integers(0, 300, 40)).astype(np.uint64)
    want = np.flatnonzero(np.isin(idx, nulls) | (idx == NULL)).astype(np.uint64)
    np.testing.assert_array_equal(oracle.gather_nulls(idx, nulls), want)
    np.testing.assert_array_equal(oracle.gather_nulls(idx), np.flatnonzero(idx == NULL))


# ---- filter --------------------------------------------------------------------

def cxx_common(a, b):
    """The type C++'s usual arithmetic conversions give `a op b` for the
    six column types (C++17 [expr.arith.conv]): a floating operand wins
    (the wider one); between integers of equal rank the unsigned one wins;
    a wider signed type wins over a narrower unsigned one (it holds all of
    its values); otherwise the unsigned type of the wider rank."""
    a, b = np.dtype(a), np.dtype(b)
    if a.kind == "f" or b.kind == "f":
        fs = [t for t in (a, b) if t.kind == "f"]
        return max(fs, key=lambda t: t.itemsize)
    if a == b:
        return a
    if a.itemsize == b.itemsize:
        return a if a.kind == "u" else b
    wide, narrow = (a, b) if a.itemsize > b.itemsize else (b, a)
    return wide


def _np_cmp(x, y, op):
    return {"eq": x == y, "neq": x != y, "lt": x < y, "le": x <= y, "gt": x > y,
            "ge": x >= y}[op]


def test_cxx_common_examples():
    # fixed by the standard: int vs unsigned -> unsigned; int vs long -> long;
    # unsigned vs long -> long (64-bit long); long vs unsigned long ->
    # unsigned long; anything vs double -> double
    assert cxx_common(np.int32, np.uint32) == np.uint32
    assert cxx_common(np.int32, np.int64) == np.int64
    assert cxx_common(np.uint32, np.int64) == np.int64
    assert cxx_common(np.int64, np.uint64) == np.uint64
    assert cxx_common(np.uint64, np.float32) == np.float32
    assert cxx_common(np.float32, np.float64) == np.float64


@pytest.mark.parametrize("ct,st", list(itertools.product(ALL_TYPES, ALL_TYPES)))
def test_filter_scalar_in_cxx_common_type(ct, st):
    rng = np.random.default_rng(3)
    col = _col(rng, ct, 400)
    if np.dtype(ct).kind == "i":
        col[::9] = -col[::9] - 3  # negatives meet unsigned scalars
    nulls = np.unique(rng.integers(0, 400, 25)).astype(np.uint64)
    ctype = cxx_common(ct, st)
    for raw in [-3, 0, 5, 17]:
        if np.dtype(st).kind == "u" and raw < 0:
            continue
        sc = np.array([raw], dtype=st)[0] if np.dtype(st).kind != "f" else np.dtype(st).type(raw + 0.5)
        for op in oracle.CMP_CODE:
            hit = _np_cmp(col.astype(ctype), np.array([sc]).astype(ctype)[0], op)
            hit[nulls.astype(np.int64)] = False
            want = np.flatnonzero(hit).astype(np.uint64)
            np.testing.assert_array_equal(oracle.filter(col, op, sc, nulls), want,
                                          err_msg=f"{ct} {st} {op} {sc}")


def test_filter_negative_int_against_unsigned_is_converted():
    # C++: -1 == 4294967295u holds (int converts to unsigned)
    col = np.array([-1, 1], np.int32)
    np.testing.assert_array_equal(oracle.filter(col, "eq", np.uint32(0xFFFFFFFF)), [0])
    np.testing.assert_array_equal(oracle.filter(col, "gt", np.uint32(2)), [0])


@pytest.mark.parametrize("op", list(oracle.CMP_CODE))
def test_filter_pair_null_rule(op):
    rng = np.random.default_rng(4)
    a = _col(rng, np.int64, 300)
    b = _col(rng, np.float64, 300)
    an = np.unique(rng.integers(0, 300, 20)).astype(np.uint64)
    bn = np.unique(rng.integers(0, 300, 20)).astype(np.uint64)
    hit = _np_cmp(a.astype(np.float64), b, op)
    hit[an.astype(np.int64)] = False
    if op != "eq":  # dfcolumn_filter_eq.cc: eq does not look at b's nulls
        hit[bn.astype(np.int64)] = False
    np.testing.assert_array_equal(oracle.filter_pair(a, b, op, an, bn), np.flatnonzero(hit))


def test_filter_tutorial():
    # doc/tutorial/src/tut5.3: c1 int, c2 double, c4 double
    c1 = np.arange(1, 9, dtype=np.int32)
    c2 = np.arange(10, 90, 10).astype(np.float64)
    c4 = np.arange(9, 1, -1).astype(np.float64)
    np.testing.assert_array_equal(oracle.filter_pair(c1, c4, "eq"), [4])
    np.testing.assert_array_equal(oracle.filter_pair(c1, c4, "lt"), [0, 1, 2, 3])
    np.testing.assert_array_equal(oracle.filter_pair(c1, c4, "ge"), [4, 5, 6, 7])
    np.testing.assert_array_equal(oracle.filter(c2, "eq", np.float64(30.0)), [2])
    np.testing.assert_array_equal(oracle.filter(c2, "le", np.float64(30.0)), [0, 1, 2])


# ---- sort ----------------------------------------------------------------------

@pytest.mark.parametrize("dt", ALL_TYPES)
@pytest.mark.parametrize("n", [0, 1, 50, 63, 64, 3000])
def test_sort_is_stable_argsort(dt, n):
    rng = np.random.default_rng(5)
    col = _col(rng, dt, n)
    k, idx = oracle.sort_index(col)
    want = np.argsort(col, kind="stable")
    np.testing.assert_array_equal(idx, want)
    np.testing.assert_array_equal(k, col[want])
    kd, idd = oracle.sort_index(col, descending=True)
    want_d = np.array(sorted(range(n), key=lambda i: col[i], reverse=True), dtype=np.int64)
    np.testing.assert_array_equal(idd, want_d)
    np.testing.assert_array_equal(kd, col[want_d])


def test_sort_with_index_orders_by_later_key_first():
    rng = np.random.default_rng(6)
    a = rng.integers(0, 5, 2000).astype(np.int64)
    b = rng.integers(-3, 3, 2000).astype(np.int32)
    _, i1 = oracle.sort_index(b)
    _, i2 = oracle.sort_index(a, i1)
    want = sorted(range(2000), key=lambda i: (a[i], b[i]))
    np.testing.assert_array_equal(i2, want)


def test_sort_tutorial():
    # doc/tutorial/src/tut5.4/t.csv, t.sort("c2").sort("c1")
    c1 = np.array([1, 3, 3, 2, 2, 2, 1, 1], np.int64)
    c2 = np.array([50, 20, 30, 20, 30, 40, 40, 50], np.int64)
    c3 = "abcdefgh"
    _, i = oracle.sort_index(c2)
    k, i = oracle.sort_index(c1, i)
    assert "".join(c3[j] for j in i) == "gahdefbc"
    np.testing.assert_array_equal(k, [1, 1, 1, 2, 2, 2, 3, 3])


# ---- multi-column group_by -------------------------------------------------------

def _tuple_groups(cols, rows):
    keys = sorted(set(tuple(c[r] for c in cols) for r in rows))
    return keys


@pytest.mark.parametrize("types", [(np.int64, np.int32), (np.uint32, np.float64, np.int64),
                                   (np.float32,), (np.int32, np.int32, np.uint64, np.int64)])
def test_multi_group_by_is_tuple_sort(types):
    rng = np.random.default_rng(7)
    n = 3000
    cols = [_col(rng, t, n, -4, 4) for t in types]
    gi, split, keys = oracle.multi_group_by(cols)
    want_keys = _tuple_groups(cols, range(n))
    got_keys = list(zip(*[k.tolist() for k in keys]))
    assert got_keys == [tuple(x.item() for x in t) for t in want_keys]
    # every group's rows carry its key, and rows appear in stable order
    order = sorted(range(n), key=lambda i: tuple(c[i] for c in cols))
    np.testing.assert_array_equal(gi, order)
    assert split[0] == 0 and split[-1] == n and len(split) == len(want_keys) + 1


def test_multi_group_by_subset_index():
    rng = np.random.default_rng(8)
    a = _col(rng, np.int64, 500, 0, 5)
    b = _col(rng, np.int64, 500, 0, 3)
    idx = np.sort(rng.choice(500, 200, replace=False)).astype(np.uint64)
    gi, split, keys = oracle.multi_group_by([a, b], idx)
    order = sorted(idx.tolist(), key=lambda i: (a[i], b[i]))
    np.testing.assert_array_equal(gi, order)
    assert len(keys[0]) == len(set((a[i], b[i]) for i in idx))


def test_multi_group_by_tutorial():
    # doc/tutorial/src/tut5.6/t.csv grouped by (c1, c2), aggregated on c3;
    # the tutorial lists the rows in the order of its multi-rank run, one
    # rank gives the lexicographic order
    c1 = np.array([1, 3, 3, 2, 2, 2, 1, 1], np.int64)
    c2 = np.array([0, 0, 0, 0, 1, 1, 1, 1], np.int64)
    c3 = np.array([30, 20, 30, 20, 30, 40, 40, 50], np.int64)
    gi, split, (k1, k2) = oracle.multi_group_by([c1, c2])
    rows = set()
    agg = {op: oracle.grouped_agg(op, c3, gi, split) for op in ["sum", "count", "avg", "max", "min"]}
    for g in range(len(k1)):
        rows.add((int(k1[g]), int(k2[g]), int(agg["sum"][g]), int(agg["count"][g]),
                  float(agg["avg"][g]), int(agg["max"][g]), int(agg["min"][g])))
    assert rows == {(1, 0, 30, 1, 30.0, 30, 30), (2, 0, 20, 1, 20.0, 20, 20),
                    (3, 0, 50, 2, 25.0, 30, 20), (1, 1, 90, 2, 45.0, 50, 40),
                    (2, 1, 70, 2, 35.0, 40, 30)}
    assert list(zip(k1.tolist(), k2.tolist())) == [(1, 0), (1, 1), (2, 0), (2, 1), (3, 0)]


# ---- outer join ------------------------------------------------------------------

def _brute_pairs(l, li, r, ri):
    return sorted((int(li[i]), int(ri[j])) for i in range(len(l)) for j in range(len(r))
                  if l[i] == r[j])


@pytest.mark.parametrize("algo", ["ve", "x86", "ve_sortmerge"])
@pytest.mark.parametrize("unique", [True, False])
def test_outer_join_closed_form(algo, unique):
    rng = np.random.default_rng(9)
    nl, nr = 700, 300
    l = rng.integers(0, 600, nl).astype(np.int64)
    r = (rng.choice(600, nr, replace=False) if unique else rng.integers(0, 200, nr)).astype(np.int64)
    li = rng.permutation(nl).astype(np.uint64) + np.uint64(1000)
    ri = np.arange(nr, dtype=np.uint64)
    lo, ro, only, used = oracle.outer_equi_join(l, li, r, ri, algo)
    assert sorted(zip(lo.tolist(), ro.tolist())) == _brute_pairs(l, li, r, ri)
    missing = ~np.isin(l, r)
    if algo == "x86" or (algo == "ve" and used):
        want = li[missing]                       # left order
    else:
        o = np.argsort(l, kind="stable")         # stable key order of LEFT
        want = li[o][missing[o]]
    np.testing.assert_array_equal(only, want)
    if algo == "ve":
        assert used == unique


def test_outer_join_tutorial():
    # doc/tutorial/src/tut5.5: t1.outer_hash_join(t2, eq("t1c3", "t2c1"));
    # the row (7, 70, 4) has no partner and gets NULL right columns
    t1c1 = np.arange(1, 9, dtype=np.int64)
    t1c3 = np.array([3, 2, 1, 2, 3, 2, 4, 1], np.int64)
    t2c1 = np.array([1, 2, 3, 2], np.int64)
    lo, ro, only, used = oracle.outer_equi_join(t1c3, np.arange(8), t2c1, np.arange(4))
    assert not used
    np.testing.assert_array_equal(t1c1[only.astype(np.int64)], [7])
    # the materialised right key column: pairs, then NULL for the outer rows
    right_idx = np.concatenate([ro, np.full(len(only), NULL, np.uint64)])
    col = oracle.gather(t2c1, right_idx)
    assert col[-1] == np.iinfo(np.int64).max
    np.testing.assert_array_equal(oracle.gather_nulls(right_idx), [len(ro)])


def test_outer_join_empty_sides():
    l = np.array([3, 1, 2], np.int64)
    e = np.zeros(0, np.int64)
    for algo in ["ve", "ve_sortmerge"]:
        lo, ro, only, _ = oracle.outer_equi_join(l, np.arange(3), e, np.zeros(0), algo)
        assert len(lo) == 0
        assert sorted(only.tolist()) == [0, 1, 2]
        lo, ro, only, _ = oracle.outer_equi_join(e, np.zeros(0), l, np.arange(3), algo)
        assert len(lo) == 0 and len(only) == 0


# ---- star join -------------------------------------------------------------------

def test_star_join_brute_force():
    rng = np.random.default_rng(10)
    nl = 2000
    dims = []
    left = []
    for nr, t in [(300, np.int64), (50, np.int32), (700, np.uint64)]:
        keys = rng.choice(4 * nr, nr, replace=False).astype(t)
        ids = rng.permutation(nr).astype(np.uint64) + np.uint64(100)
        dims.append((keys, ids))
        left.append(rng.integers(0, 4 * nr, nl).astype(t))
    lidx = np.sort(rng.choice(nl, 1500, replace=False)).astype(np.uint64)
    lo, ros = oracle.star_join(left, dims, lidx)
    want_rows, want_ids = [], [[] for _ in dims]
    for i in lidx.tolist():
        pos = [np.flatnonzero(k == c[i]) for c, (k, _) in zip(left, dims)]
        if all(len(p) for p in pos):
            want_rows.append(i)
            for d, p in enumerate(pos):
                want_ids[d].append(int(dims[d][1][p[0]]))
    np.testing.assert_array_equal(lo, want_rows)
    for d in range(len(dims)):
        np.testing.assert_array_equal(ros[d], want_ids[d])


def test_star_join_requires_unique_right():
    with pytest.raises(RuntimeError, match="not unique"):
        oracle.star_join([np.arange(5, dtype=np.int64)],
                         [(np.array([1, 2, 2], np.int64), np.arange(3))])

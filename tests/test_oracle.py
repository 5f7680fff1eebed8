"""Pins the CPU oracle before anything is checked against it.

Each check comes from something other than the oracle's own formula: numpy's
stable sort / searchsorted / cumsum (closed forms), Python dicts and brute
force pair sets, math.fsum, and the reference tutorial's known answers
(doc/tutorial/tutorial.md:2355-2375, :2500-2531).
"""
import csv
import json
import math
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")
INT_TYPES = [np.int32, np.uint32, np.int64, np.uint64]


def _keys(rng, dt, n, kind):
    info = np.iinfo(dt)
    if kind == "full":
        b = rng.integers(0, np.iinfo(np.uint64).max, n, dtype=np.uint64, endpoint=True)
        if dt == np.uint64:
            return b
        if dt == np.int64:
            return b.view(np.int64)
        hi32 = (b >> np.uint64(32)).astype(np.uint32)
        return hi32 if dt == np.uint32 else hi32.view(np.int32)
    if kind == "small":
        return rng.integers(0, 300, n).astype(dt)
    if kind == "dup":
        lo = max(info.min, -50)
        return rng.integers(lo, 50, n).astype(dt)
    if kind == "twobyte":
        return rng.integers(0, 1 << 16, n).astype(dt)
    raise ValueError(kind)


def _stable_desc_order(k):
    # stable descending: ties keep original order
    _, inv = np.unique(k, return_inverse=True)
    return np.lexsort((np.arange(len(k)), -inv.astype(np.int64)))


@pytest.mark.parametrize("dt", INT_TYPES)
@pytest.mark.parametrize("n", [0, 1, 2, 63, 64, 65, 1000, 70000])
@pytest.mark.parametrize("kind", ["full", "small", "dup", "twobyte"])
def test_radix_sort_int_closed_form(dt, n, kind):
    rng = np.random.default_rng(n * 7 + len(kind))
    k = _keys(rng, dt, n, kind)
    v = np.arange(n, dtype=np.uint64)
    sk, sv = oracle.radix_sort(k, v)
    order = np.argsort(k, kind="stable")
    assert np.array_equal(sv, order)
    assert np.array_equal(sk, k[order])
    # key-only and 4-byte payload agree with the kv sort
    sk2, _ = oracle.radix_sort(k)
    assert np.array_equal(sk2, sk)
    _, sv4 = oracle.radix_sort(k, v.astype(np.uint32))
    assert np.array_equal(sv4, order.astype(np.uint32))
    # descending: ties keep original order (radix_sort.incl:242-480)
    dk, dv = oracle.radix_sort(k, v, descending=True)
    dorder = _stable_desc_order(k)
    assert np.array_equal(dv, dorder)


@pytest.mark.parametrize("dt", [np.int32, np.int64])
def test_radix_sort_positive_only_semantics(dt):
    rng = np.random.default_rng(3)
    k = rng.integers(0, 5000, 3000).astype(dt)
    v = np.arange(3000, dtype=np.uint64)
    sk, sv = oracle.radix_sort(k, v, positive_only=True)
    assert np.array_equal(sv, np.argsort(k, kind="stable"))
    # with negatives and positive_only the reference sorts only the low
    # get_max_key_size() bytes of the raw bits (radix_sort.hpp:206-208)
    k2 = k.copy()
    k2[::7] = -k2[::7] - 1
    sk, sv = oracle.radix_sort(k2, v, positive_only=True)
    mx = max(int(k2.max()), 0)
    nbytes = max(1, (mx.bit_length() + 7) // 8)
    bits = k2.view(np.uint32 if dt == np.int32 else np.uint64).astype(np.uint64)
    low = bits & np.uint64((1 << (8 * nbytes)) - 1)
    assert np.array_equal(sv, np.argsort(low, kind="stable"))


@pytest.mark.parametrize("ft,it", [(np.float32, np.uint32), (np.float64, np.uint64)])
def test_radix_sort_float_quirks(ft, it):
    rng = np.random.default_rng(5)
    n = 500
    # no negatives: keys are sorted as raw bits, so -0.0 comes after every
    # positive value (radix_sort.hpp:256-266: contain_negative uses < 0)
    k = rng.random(n).astype(ft) * 100
    k[10] = -0.0
    k[20] = 0.0
    v = np.arange(n, dtype=np.uint64)
    sk, sv = oracle.radix_sort(k, v)
    assert np.array_equal(sv, np.argsort(k.view(it), kind="stable"))
    assert sv[-1] == 10 and np.signbit(sk[-1])
    # with negatives: total order of the flipped bits, -0.0 before +0.0
    k[30] = -3.5
    sk, sv = oracle.radix_sort(k, v)
    b = k.view(it).astype(np.uint64)
    top = np.uint64(1) << np.uint64(8 * k.itemsize - 1)
    flipped = np.where(b & top, ~b & np.uint64((1 << (8 * k.itemsize)) - 1), b | top)
    assert np.array_equal(sv, np.argsort(flipped, kind="stable"))
    assert list(sv[:1]) == [30]
    pos_neg0 = list(sv).index(10)
    pos_pos0 = list(sv).index(20)
    assert pos_neg0 < pos_pos0
    # below 64 elements: insertion sort with `>`; -0.0 == 0.0 keep order
    ks = np.array([3.0, -0.0, 1.0, 0.0, -2.0, 0.0], ft)
    sk, sv = oracle.radix_sort(ks, np.arange(6, dtype=np.uint64))
    assert list(sv) == [4, 1, 3, 5, 2, 0]


@pytest.mark.parametrize("dt", INT_TYPES + [np.float64])
def test_prefix_sum_closed_form(dt):
    rng = np.random.default_rng(1)
    a = rng.integers(0, 100, 5000).astype(dt)
    out = oracle.prefix_sum(a)
    assert np.array_equal(out, np.cumsum(a, dtype=dt))
    assert len(oracle.prefix_sum(np.zeros(0, dt))) == 0


@pytest.mark.parametrize("n", [0, 1, 2, 10, 5000])
def test_set_separate_closed_form(n):
    rng = np.random.default_rng(n)
    k = np.sort(rng.integers(-5, 5, n)).astype(np.int64)
    sep = oracle.set_separate(k)
    if n == 0:
        assert list(sep) == [0]
        return
    starts = np.concatenate([[0], np.flatnonzero(np.diff(k) != 0) + 1, [n]])
    assert np.array_equal(sep, starts.astype(np.uint64))
    assert oracle.set_is_unique(k) == (len(np.unique(k)) == n)
    assert oracle.set_is_unique(np.arange(n, dtype=np.int64))


def test_lower_upper_bound_closed_form():
    rng = np.random.default_rng(2)
    s = np.sort(rng.integers(0, 1000, 3000)).astype(np.int64)
    q = rng.integers(-10, 1010, 4000).astype(np.int64)
    assert np.array_equal(oracle.lower_bound(s, q), np.searchsorted(s, q, "left"))
    assert np.array_equal(oracle.lower_bound(s, q, upper=True), np.searchsorted(s, q, "right"))


PRIMES = [8 + 3, 16 + 3, 32 + 5, 64 + 3, 128 + 3, 256 + 27, 512 + 9, 1024 + 9,
          2048 + 5, 4096 + 3, 8192 + 27, 16384 + 43, 32768 + 3, 65536 + 45,
          131072 + 29, 262144 + 3, 524288 + 21, 1048576 + 7]


def _is_prime(p):
    return p > 1 and all(p % d for d in range(2, int(p ** 0.5) + 1))


def test_hashtable_prime_list_is_prime():
    assert all(_is_prime(p) for p in PRIMES)


@pytest.mark.parametrize("dt", [np.int64, np.uint32, np.int32])
@pytest.mark.parametrize("n", [0, 1, 100, 20000])
def test_hashtable_dictionary_semantics(dt, n):
    rng = np.random.default_rng(n + 11)
    hi = 10 * max(n, 1)
    keys = rng.choice(hi, size=n, replace=False).astype(dt)
    if dt == np.int64 and n:
        keys[::3] = -keys[::3] - 1
    vals = rng.integers(0, 2**63, n, dtype=np.uint64)
    ht = oracle.UniqueHashtable(keys, vals, check_unique=True)
    assert ht.is_unique_ok
    ts = ht.table_size
    assert ts == min(p for p in PRIMES if p >= 3 * n)
    # a key owns its slot iff it is the last row hashing there, so the
    # conflict list holds n - (#distinct home slots) rows
    bits = keys.view(np.uint32 if keys.itemsize == 4 else np.uint64).astype(np.uint64)
    slots = bits % np.uint64(ts)
    assert ht.conflict_size == n - len(np.unique(slots))
    d = dict(zip(keys.tolist(), vals.tolist()))
    q = np.concatenate([keys, rng.integers(hi, 2 * hi + 1, n // 2 + 1).astype(dt)])
    rng.shuffle(q)
    got, miss = ht.lookup(q)
    exp_miss = [i for i, x in enumerate(q.tolist()) if x not in d]
    assert miss.tolist() == exp_miss
    for i, x in enumerate(q.tolist()):
        assert got[i] == (d[x] if x in d else 0)
    ex = ht.check_existence(q)
    assert ex.tolist() == [1 if x in d else 0 for x in q.tolist()]
    # all_keys: owned slots in slot order, then conflicts sorted by key
    last = {}
    for i, s in enumerate(slots.tolist()):
        last[s] = i
    owned = [keys[last[s]] for s in sorted(last)]
    conf = sorted(set(range(n)) - set(last.values()), key=lambda i: (keys[i], i))
    exp = np.array(owned + [keys[i] for i in conf], dtype=dt)
    assert np.array_equal(ht.all_keys(), exp)


def test_hashtable_uniqueness_detection():
    # duplicate equal to the slot owner's key
    k = np.array([5, 7, 5], np.int64)
    assert not oracle.UniqueHashtable(k, np.arange(3), check_unique=True).is_unique_ok
    # duplicate that only shows inside the conflict list: 11 slots for n=3;
    # 4 and 15 share slot 4, the last writer is 26 (also slot 4)
    k = np.array([4, 4, 26], np.int64)
    ht = oracle.UniqueHashtable(k, np.arange(3), check_unique=True)
    assert ht.table_size == 11 and ht.conflict_size == 2
    assert not ht.is_unique_ok
    k = np.array([4, 15, 26], np.int64)
    assert oracle.UniqueHashtable(k, np.arange(3), check_unique=True).is_unique_ok


def _brute_pairs(l, li, r, ri):
    out = []
    for i, x in enumerate(l.tolist()):
        for j, y in enumerate(r.tolist()):
            if x == y:
                out.append((int(li[i]), int(ri[j])))
    return sorted(out)


@pytest.mark.parametrize("unique_right", [True, False])
def test_equi_join_brute_force(unique_right):
    rng = np.random.default_rng(4 if unique_right else 6)
    nl, nr = 300, 200
    r = (rng.choice(400, nr, replace=False) if unique_right
         else rng.integers(0, 120, nr)).astype(np.int64)
    l = rng.integers(0, 450, nl).astype(np.int64)
    li = rng.permutation(nl).astype(np.uint64) + np.uint64(1000)
    ri = rng.permutation(nr).astype(np.uint64) + np.uint64(5000)
    lo, ro, used_hash = oracle.equi_join(l, li, r, ri, "ve")
    assert used_hash == unique_right
    assert sorted(zip(lo.tolist(), ro.tolist())) == _brute_pairs(l, li, r, ri)
    lpos = {int(v): p for p, v in enumerate(li.tolist())}
    rpos = {int(v): p for p, v in enumerate(ri.tolist())}
    if unique_right:
        # probe (left) order
        assert [lpos[int(x)] for x in lo] == sorted(lpos[int(x)] for x in lo)
    else:
        # (key, left position, right position) order
        keyseq = [(int(l[lpos[int(a)]]), lpos[int(a)], rpos[int(b)]) for a, b in zip(lo, ro)]
        assert keyseq == sorted(keyseq)
    xo, xr, _ = oracle.equi_join(l, li, r, ri, "x86")
    assert sorted(zip(xo.tolist(), xr.tolist())) == _brute_pairs(l, li, r, ri)


def _load_csv(name):
    with open(os.path.join(GOLD, name)) as f:
        return [row for row in csv.reader(f)]


def test_tutorial_join_known_answer():
    t1 = _load_csv("tut5_5_t1.csv")
    t2 = _load_csv("tut5_5_t2.csv")
    kat = json.load(open(os.path.join(GOLD, "tutorial_known_answers.json")))
    t1c3 = np.array([int(r[2]) for r in t1], np.int32)
    t2c1 = np.array([int(r[0]) for r in t2], np.int32)
    # t1.hash_join(t2, eq(t1c3, t2c1)): equi_join(left=t1c3, right=t2c1)
    for algo in ("ve", "x86"):
        lo, ro, used = oracle.equi_join(t1c3, np.arange(8), t2c1, np.arange(4), algo)
        rows = sorted([int(t1[a][0]), int(t1[a][1]), int(t1[a][2]),
                       int(t2[b][0]), t2[b][1]] for a, b in zip(lo, ro))
        assert rows == sorted(kat["hash_join_t1c3_eq_t2c1_rows"])
    # t2c1 has a duplicate key (2), so the VE body takes the sort-merge path
    assert not oracle.equi_join(t1c3, np.arange(8), t2c1, np.arange(4), "ve")[2]


def test_tutorial_group_by_known_answer():
    t = _load_csv("tut5_6_t.csv")
    kat = json.load(open(os.path.join(GOLD, "tutorial_known_answers.json")))["group_by_c1_agg_c3"]
    c1 = np.array([int(r[0]) for r in t], np.int32)
    c3 = np.array([int(r[2]) for r in t], np.int32)
    gi, split, keys = oracle.group_by(c1)
    assert keys.tolist() == kat["c1"]
    assert oracle.grouped_agg("sum", c3, gi, split).tolist() == kat["sum"]
    assert oracle.grouped_agg("count", c3, gi, split).tolist() == kat["count"]
    assert oracle.grouped_agg("avg", c3, gi, split).tolist() == kat["avg"]
    assert oracle.grouped_agg("max", c3, gi, split).tolist() == kat["max"]
    assert oracle.grouped_agg("min", c3, gi, split).tolist() == kat["min"]


@pytest.mark.parametrize("kt", [np.int64, np.int32, np.float64])
def test_group_by_closed_form(kt):
    rng = np.random.default_rng(9)
    n = 20000
    k = rng.integers(-300, 300, n).astype(kt)
    gi, split, keys = oracle.group_by(k)
    order = np.argsort(k, kind="stable")
    assert np.array_equal(gi, order)
    u, first, cnt = np.unique(k, return_index=True, return_counts=True)
    assert np.array_equal(keys, u)
    assert np.array_equal(np.diff(split), cnt)
    v = rng.standard_normal(n)
    s = oracle.grouped_agg("sum", v, gi, split)
    for g in range(0, len(u), 37):
        rows = order[split[g]:split[g + 1]]
        exact = math.fsum(v[rows].tolist())
        bound = len(rows) * 2.3e-16 * np.abs(v[rows]).sum()
        assert abs(s[g] - exact) <= bound
    iv = rng.integers(-1000, 1000, n).astype(np.int32)
    mx = oracle.grouped_agg("max", iv, gi, split)
    mn = oracle.grouped_agg("min", iv, gi, split)
    starts = split[:-1].astype(np.int64)
    assert np.array_equal(mx, np.maximum.reduceat(iv[order], starts))
    assert np.array_equal(mn, np.minimum.reduceat(iv[order], starts))
    avg = oracle.grouped_agg("avg", v, gi, split)
    assert np.array_equal(avg, s / cnt)


def test_grouped_agg_nulls():
    # nulls read as 0 in sum and are not counted; a group with only nulls
    # yields max() for sum/avg/max (dfcolumn_grouped_agg.cc:105-275,454-510)
    col = np.array([1.0, 2.0, 5.0, 7.0], np.float64)
    gi = np.arange(4, dtype=np.uint64)
    split = np.array([0, 2, 4], np.uint64)
    nulls = np.array([1, 2, 3], np.uint64)
    mx = np.finfo(np.float64).max
    assert oracle.grouped_agg("sum", col, gi, split, nulls).tolist() == [1.0, mx]
    assert oracle.grouped_agg("count", col, gi, split, nulls).tolist() == [1, 0]
    assert oracle.grouped_agg("max", col, gi, split, nulls).tolist() == [1.0, mx]
    assert oracle.grouped_agg("avg", col, gi, split, nulls).tolist() == [1.0, mx]


def test_group_by_with_local_idx_closed_form():
    # the key column is read through local_idx (dfcolumn_groupby.cc:39)
    rng = np.random.default_rng(12)
    col = rng.integers(0, 100, 50000).astype(np.int64)
    idx = np.sort(rng.choice(50000, 20000, replace=False)).astype(np.uint64)
    gi, split, keys = oracle.group_by(col, idx)
    o = np.argsort(col[idx.astype(np.int64)], kind="stable")
    assert np.array_equal(gi, idx[o])
    assert np.array_equal(keys, np.unique(col[idx.astype(np.int64)]))


def _colliding_positive(target_bits, ts, k=1 << 20):
    """A positive double whose bits share the home slot of target_bits."""
    return (np.uint64(target_bits % ts) + np.uint64(ts) * np.uint64(k)).view(np.float64)


@pytest.mark.parametrize("with_negative", [False, True])
def test_hashtable_float_conflict_list_follows_radix_sort(with_negative):
    # The conflict list is radix_sort(conflict_key, conflict_val)
    # (hashtable.hpp:1234), so float keys follow the radix sort's flips: with
    # no negative conflict key the list is in raw-bit order (-0.0 after NaN
    # after the positives); with one, in flipped total order.
    rng = np.random.default_rng(77)
    n = 2000  # about 240 conflicts: the radix path, not the insertion sort
    ts = min(p for p in PRIMES if p >= 3 * n)
    keys = rng.random(n) * 1000.0
    neg0 = np.uint64(1) << np.uint64(63)
    nan = np.uint64(0x7FF8000000000000)
    keys[0] = -0.0
    keys[1] = np.uint64(nan).view(np.float64)
    # later rows own those slots, so rows 0 and 1 are conflicts
    keys[n - 1] = _colliding_positive(int(neg0), ts)
    keys[n - 2] = _colliding_positive(int(nan), ts, k=1 << 21)
    if with_negative:
        keys[2] = -5.0
        keys[n - 3] = _colliding_positive(int(keys[2:3].view(np.uint64)[0]), ts, k=1 << 22)
    vals = np.arange(n, dtype=np.uint64)
    ht = oracle.UniqueHashtable(keys, vals)
    bits = keys.view(np.uint64)
    slots = bits % np.uint64(ts)
    last = {}
    for i, s in enumerate(slots.tolist()):
        last[s] = i
    conf = np.array(sorted(set(range(n)) - set(last.values())), np.int64)
    assert 0 in conf and 1 in conf and len(conf) >= 64
    ck, cv = oracle.radix_sort(keys[conf], conf.astype(np.uint64))
    owned = np.array([keys[last[s]] for s in sorted(last)])
    got = ht.all_keys()
    assert np.array_equal(got.view(np.uint64), np.concatenate([owned, ck]).view(np.uint64))
    tail = ck.view(np.uint64)
    if not with_negative:
        assert tail[-1] == neg0 and tail[-2] == nan
    else:
        assert tail[0] == keys[2:3].view(np.uint64)[0]
        pos_neg0 = list(tail).index(neg0)
        assert pos_neg0 == 1  # right after -5.0, before every positive
    # a NaN query never compares equal: a miss either way
    _, miss = ht.lookup(np.array([np.nan]))
    assert miss.tolist() == [0]


def test_lower_upper_bound_nan_query():
    # lower_bound.hpp:44-48: a NaN query fails `x <= s[mid]` everywhere and
    # ends at m; upper_bound.hpp:44-48: it fails `x >= s[mid]` and ends at 0
    s = np.array([1.0, 2.0, 3.0, 4.0, 5.0])
    q = np.array([np.nan, 2.0, -np.inf, np.inf])
    assert oracle.lower_bound(s, q).tolist() == [5, 1, 0, 5]
    assert oracle.lower_bound(s, q, upper=True).tolist() == [0, 2, 0, 5]

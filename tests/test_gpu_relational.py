"""GPU parity of the column operations around the execution core against
the CPU oracle: global_extract with NULLs and its null list, filter against
a scalar or a column (every type pair, every comparison, null rows), sort /
sort_desc / sort by a further column, multi-column group_by feeding the
grouped aggregates, and the outer equi-join.  Everything is integer or
index work and compared bit-exact; at 2^24 rows size-independent
properties are checked instead (sortedness and permutation, group counts,
join conservation)."""
import itertools

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
ALL_TYPES = [np.int32, np.uint32, np.int64, np.uint64, np.float32, np.float64]
NULL = oracle.NULL_INDEX


@pytest.fixture(scope="module")
def fa():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import frovedis_amd
    frovedis_amd.lib()
    return frovedis_amd


def _col(rng, dt, n, lo=-40, hi=40):
    dt = np.dtype(dt)
    v = rng.integers(lo, hi, n)
    if dt.kind == "u":
        v = np.abs(v)
    if dt.kind == "f":
        return (v * 0.5).astype(dt)
    return v.astype(dt)


def _d(fa, a):
    return fa.DeviceArray.from_numpy(np.ascontiguousarray(a))


# ---- global_extract ---------------------------------------------------------------

@pytest.mark.parametrize("dt", ALL_TYPES)
def test_gather_matches_oracle(fa, dt):
    rng = np.random.default_rng(11)
    for n_src, n in [(1, 1), (1000, 5000), (1 << 20, 1 << 21)]:
        src = _col(rng, dt, n_src, -1000, 1000)
        idx = rng.integers(0, n_src, n).astype(np.uint64)
        idx[::13] = NULL
        got = fa.global_extract(_d(fa, src), _d(fa, idx)).numpy()
        np.testing.assert_array_equal(got, oracle.gather(src, idx))
        nulls = np.unique(rng.integers(0, n_src, max(1, n_src // 50))).astype(np.uint64)
        gn = fa.extract_nulls(_d(fa, idx), _d(fa, nulls), n_src).numpy()
        np.testing.assert_array_equal(gn, oracle.gather_nulls(idx, nulls))
        np.testing.assert_array_equal(fa.extract_nulls(_d(fa, idx)).numpy(),
                                      oracle.gather_nulls(idx))


def test_gather_rejects_out_of_range_and_empty(fa):
    src = _d(fa, np.arange(10, dtype=np.int64))
    with pytest.raises(fa.FhipError):
        fa.global_extract(src, _d(fa, np.array([3, 10], np.uint64)))
    assert len(fa.global_extract(src, _d(fa, np.zeros(0, np.uint64)))) == 0


# ---- filter -----------------------------------------------------------------------

@pytest.mark.parametrize("ct,st", list(itertools.product(ALL_TYPES, ALL_TYPES)))
def test_filter_scalar_matches_oracle(fa, ct, st):
    rng = np.random.default_rng(12)
    n = 20000
    col = _col(rng, ct, n)
    if np.dtype(ct).kind == "i":
        col[::9] = -col[::9] - 3
    nulls = np.unique(rng.integers(0, n, 300)).astype(np.uint64)
    dcol, dnull = _d(fa, col), _d(fa, nulls)
    for raw in [-3, 0, 17]:
        if np.dtype(st).kind == "u" and raw < 0:
            continue
        sc = np.dtype(st).type(raw + 0.5) if np.dtype(st).kind == "f" else np.dtype(st).type(raw)
        for op in fa.CMP_OPS:
            got = fa.filter(dcol, op, sc, nulls=dnull).numpy()
            np.testing.assert_array_equal(got, oracle.filter(col, op, sc, nulls),
                                          err_msg=f"{ct} {st} {op} {sc}")
    got = fa.filter(dcol, "ge", np.dtype(st).type(1)).numpy()
    np.testing.assert_array_equal(got, oracle.filter(col, "ge", np.dtype(st).type(1)))


@pytest.mark.parametrize("at,bt", [(np.int32, np.float64), (np.int64, np.uint64),
                                   (np.uint32, np.int32), (np.float32, np.int64),
                                   (np.uint64, np.uint64), (np.float64, np.float32)])
def test_filter_pair_matches_oracle(fa, at, bt):
    rng = np.random.default_rng(13)
    n = 50000
    a, b = _col(rng, at, n, -5, 5), _col(rng, bt, n, -5, 5)
    an = np.unique(rng.integers(0, n, 500)).astype(np.uint64)
    bn = np.unique(rng.integers(0, n, 500)).astype(np.uint64)
    for op in fa.CMP_OPS:
        got = fa.filter(_d(fa, a), op, _d(fa, b), nulls=_d(fa, an), other_nulls=_d(fa, bn)).numpy()
        np.testing.assert_array_equal(got, oracle.filter_pair(a, b, op, an, bn), err_msg=op)


@pytest.mark.parametrize("shape", [0, 1])
def test_compaction_tile_shapes(fa, shape):
    """Both tiles of the one-pass compaction, on partial last tiles and on
    many tiles (the look-back crosses more than 64 of them)."""
    from frovedis_amd import autotune
    rng = np.random.default_rng(22)
    autotune.tune_set("compact.shape", shape)
    try:
        for n in [1, 4095, 4097, 16385, 3 << 20]:
            col = rng.integers(0, 100, n).astype(np.int32)
            got = fa.filter(_d(fa, col), "lt", np.int32(37)).numpy()
            np.testing.assert_array_equal(got, np.flatnonzero(col < 37))
            none = fa.filter(_d(fa, col), "gt", np.int32(1000)).numpy()
            assert len(none) == 0
            ks = np.sort(rng.integers(0, max(n // 3, 1), n)).astype(np.int64)
            np.testing.assert_array_equal(fa.set_separate(_d(fa, ks)).numpy(),
                                          oracle.set_separate(ks))
    finally:
        autotune.tune_set("compact.shape", None)


def test_filter_tutorial_and_errors(fa):
    c1 = _d(fa, np.arange(1, 9, dtype=np.int32))
    c2 = _d(fa, np.arange(10, 90, 10).astype(np.float64))
    c4 = _d(fa, np.arange(9, 1, -1).astype(np.float64))
    np.testing.assert_array_equal(fa.filter(c1, "eq", c4).numpy(), [4])
    np.testing.assert_array_equal(fa.filter(c2, "eq", np.float64(30.0)).numpy(), [2])
    with pytest.raises(fa.FhipError):
        fa.filter(c1, "like", c4)
    assert len(fa.filter(_d(fa, np.zeros(0, np.int64)), "eq", np.int64(0))) == 0


# ---- sort -------------------------------------------------------------------------

@pytest.mark.parametrize("dt", ALL_TYPES)
def test_sort_matches_oracle(fa, dt):
    rng = np.random.default_rng(14)
    for n in [0, 1, 63, 64, 1000, 300000]:
        col = _col(rng, dt, n, -5000, 5000)
        for desc in [False, True]:
            k, i = (fa.sort_desc if desc else fa.sort)(_d(fa, col))
            rk, ri = oracle.sort_index(col, descending=desc)
            np.testing.assert_array_equal(i.numpy(), ri, err_msg=f"n={n} desc={desc}")
            np.testing.assert_array_equal(k.numpy(), rk)


def test_sort_by_second_column_and_tutorial(fa):
    rng = np.random.default_rng(15)
    n = 200000
    a = _col(rng, np.int64, n, 0, 50)
    b = _col(rng, np.float32, n, -20, 20)
    _, i1 = fa.sort(_d(fa, b))
    k2, i2 = fa.sort_desc(_d(fa, a), i1)
    _, r1 = oracle.sort_index(b)
    rk2, r2 = oracle.sort_index(a, r1, descending=True)
    np.testing.assert_array_equal(i2.numpy(), r2)
    np.testing.assert_array_equal(k2.numpy(), rk2)
    c1 = np.array([1, 3, 3, 2, 2, 2, 1, 1], np.int64)
    c2 = np.array([50, 20, 30, 20, 30, 40, 40, 50], np.int64)
    _, i = fa.sort(_d(fa, c2))
    _, i = fa.sort(_d(fa, c1), i)
    assert "".join("abcdefgh"[j] for j in i.numpy()) == "gahdefbc"


def test_sort_large_properties(fa):
    import torch
    n = 1 << 24
    g = torch.Generator(device="cuda").manual_seed(5)
    t = torch.randint(-(1 << 40), 1 << 40, (n,), device="cuda", dtype=torch.int64, generator=g)
    col = fa.DeviceArray(t, np.int64)
    k, i = fa.sort(col)
    kt, it = k.t, i.t
    assert bool((kt[1:] >= kt[:-1]).all())
    assert bool(torch.equal(t[it], kt))
    assert bool(torch.equal(torch.sort(it).values, torch.arange(n, device="cuda")))
    # stability: equal keys keep increasing row ids
    eq = kt[1:] == kt[:-1]
    assert bool((it[1:][eq] > it[:-1][eq]).all())


# ---- multi-column group_by ---------------------------------------------------------

@pytest.fixture(params=["composite", "column_sorts"])
def mg_path(fa, request):
    """Runs a test with the composite-key path allowed (the default) and
    with it switched off (the column-by-column sorts)."""
    from frovedis_amd import autotune
    autotune.tune_set("multi_group_by.composite", 1 if request.param == "composite" else 0)
    yield request.param
    autotune.tune_set("multi_group_by.composite", None)


@pytest.mark.parametrize("types", [(np.int64, np.int32), (np.uint32, np.float64, np.int64),
                                   (np.float32,), (np.int32, np.int32, np.uint64, np.int64),
                                   (np.uint64, np.int32, np.uint32)])
def test_multi_group_by_matches_oracle(fa, mg_path, types):
    rng = np.random.default_rng(16)
    for n in [0, 1, 40, 5000, 400000]:
        cols = [_col(rng, t, n, -6, 6) for t in types]
        gi, split, keys = fa.multi_group_by([_d(fa, c) for c in cols])
        rgi, rsplit, rkeys = oracle.multi_group_by(cols)
        np.testing.assert_array_equal(gi.numpy(), rgi)
        np.testing.assert_array_equal(split.numpy(), rsplit)
        for k, rk in zip(keys, rkeys):
            np.testing.assert_array_equal(k.numpy(), rk)
        if n:
            v = rng.integers(-1000, 1000, n).astype(np.int64)
            for op in ["sum", "count", "max", "min", "avg"]:
                got = fa.grouped_agg(op, _d(fa, v), gi, split).numpy()
                np.testing.assert_array_equal(got, oracle.grouped_agg(op, v, rgi, rsplit))


@pytest.mark.parametrize("case", ["full64", "int32_uint32", "over64", "constant"])
def test_multi_group_by_key_ranges(fa, mg_path, case):
    """Ranges that fill the 64-bit composite exactly, overflow it (the
    column sorts run instead) or collapse to one value."""
    rng = np.random.default_rng(21)
    n = 100000
    if case == "full64":
        a = rng.integers(np.iinfo(np.int64).min, np.iinfo(np.int64).max, n, dtype=np.int64)
        a[:3] = [np.iinfo(np.int64).min, np.iinfo(np.int64).max, 0]
        a[3:1000] = a[1000:1997]  # repeated keys
        cols = [a]
    elif case == "int32_uint32":
        a = rng.integers(-2**31, 2**31, n).astype(np.int32)
        b = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
        a[:2], b[:2] = [-2**31, 2**31 - 1], [0, 2**32 - 1]
        a[2:5000] = a[5000:9998]
        b[2:5000] = b[5000:9998]
        cols = [a, b]
    elif case == "over64":
        a = rng.integers(0, 1 << 40, n).astype(np.uint64)
        b = rng.integers(-(1 << 40), 1 << 40, n).astype(np.int64)
        a[:4000] = a[4000:8000]
        b[:4000] = b[4000:8000]
        cols = [a, b]
    else:
        cols = [np.full(n, -7, np.int64), np.full(n, 3, np.uint32)]
    gi, split, keys = fa.multi_group_by([_d(fa, c) for c in cols])
    rgi, rsplit, rkeys = oracle.multi_group_by(cols)
    np.testing.assert_array_equal(gi.numpy(), rgi)
    np.testing.assert_array_equal(split.numpy(), rsplit)
    for k, rk in zip(keys, rkeys):
        np.testing.assert_array_equal(k.numpy(), rk)


def test_multi_group_by_subset_and_tutorial(fa):
    rng = np.random.default_rng(17)
    a = _col(rng, np.int64, 10000, 0, 30)
    b = _col(rng, np.int32, 10000, 0, 7)
    idx = np.sort(rng.choice(10000, 3000, replace=False)).astype(np.uint64)
    gi, split, keys = fa.multi_group_by([_d(fa, a), _d(fa, b)], _d(fa, idx))
    rgi, rsplit, rkeys = oracle.multi_group_by([a, b], idx)
    np.testing.assert_array_equal(gi.numpy(), rgi)
    np.testing.assert_array_equal(split.numpy(), rsplit)
    c1 = np.array([1, 3, 3, 2, 2, 2, 1, 1], np.int64)
    c2 = np.array([0, 0, 0, 0, 1, 1, 1, 1], np.int64)
    c3 = np.array([30, 20, 30, 20, 30, 40, 40, 50], np.int64)
    gi, split, (k1, k2) = fa.multi_group_by([_d(fa, c1), _d(fa, c2)])
    s = fa.grouped_agg("sum", _d(fa, c3), gi, split).numpy()
    assert list(zip(k1.numpy().tolist(), k2.numpy().tolist(), s.tolist())) == \
        [(1, 0, 30), (1, 1, 90), (2, 0, 20), (2, 1, 70), (3, 0, 50)]


def test_multi_group_by_large_group_count(fa):
    rng = np.random.default_rng(18)
    n = 1 << 24
    a = rng.integers(0, 3000, n).astype(np.int64)
    b = rng.integers(0, 500, n).astype(np.int32)
    gi, split, (ka, kb) = fa.multi_group_by([_d(fa, a), _d(fa, b)])
    want = np.unique(a * 1000 + b)
    got = ka.numpy() * 1000 + kb.numpy().astype(np.int64)
    np.testing.assert_array_equal(got, want)
    sp = split.numpy()
    assert sp[0] == 0 and sp[-1] == n


# ---- outer join --------------------------------------------------------------------

@pytest.mark.parametrize("dt", [np.int32, np.uint64, np.int64, np.float64])
@pytest.mark.parametrize("unique", [True, False])
@pytest.mark.parametrize("strategy", ["auto", "hash", "sortmerge"])
def test_outer_join_matches_oracle(fa, dt, unique, strategy):
    if strategy == "hash" and not unique:
        pytest.skip("the hash strategy needs unique right keys")
    rng = np.random.default_rng(19)
    for nl, nr in [(0, 5), (5, 0), (1, 1), (3000, 1000), (300000, 200000)]:
        hi = 4 * max(nl, nr, 1)
        l = rng.integers(0, hi, nl).astype(dt)
        r = (rng.choice(hi, nr, replace=False) if unique else rng.integers(0, hi, nr)).astype(dt)
        li = (rng.permutation(nl).astype(np.uint64) + np.uint64(7))
        ri = np.arange(nr, dtype=np.uint64) * np.uint64(3)
        lo, ro, only, used = fa.outer_equi_join(_d(fa, l), _d(fa, li), _d(fa, r), _d(fa, ri),
                                                strategy=strategy)
        algo = "ve_sortmerge" if strategy == "sortmerge" else "ve"
        rlo, rro, ronly, rused = oracle.outer_equi_join(l, li, r, ri, algo)
        if strategy == "auto":
            assert used == rused
        np.testing.assert_array_equal(lo.numpy(), rlo, err_msg=f"{nl} {nr}")
        np.testing.assert_array_equal(ro.numpy(), rro)
        np.testing.assert_array_equal(only.numpy(), ronly)
        # inner part equals equi_join
        ilo, iro, _ = fa.equi_join(_d(fa, l), _d(fa, li), _d(fa, r), _d(fa, ri), strategy=strategy)
        np.testing.assert_array_equal(ilo.numpy(), rlo)


def test_outer_join_tutorial_materialised(fa):
    t1c1 = np.arange(1, 9, dtype=np.int64)
    t1c3 = np.array([3, 2, 1, 2, 3, 2, 4, 1], np.int64)
    t2c1 = np.array([1, 2, 3, 2], np.int64)
    t2c2 = np.array([1, 2, 3, 4], np.int64)  # stands for a, b, c, d
    lo, ro, only, used = fa.outer_equi_join(_d(fa, t1c3), _d(fa, np.arange(8, dtype=np.uint64)),
                                            _d(fa, t2c1), _d(fa, np.arange(4, dtype=np.uint64)))
    assert not used
    left_idx = np.concatenate([lo.numpy(), only.numpy()])
    right_idx = np.concatenate([ro.numpy(), np.full(len(only), NULL, np.uint64)])
    c1 = fa.global_extract(_d(fa, t1c1), _d(fa, left_idx)).numpy()
    r2 = fa.global_extract(_d(fa, t2c2), _d(fa, right_idx)).numpy()
    nulls = fa.extract_nulls(_d(fa, right_idx)).numpy()
    rows = {(int(x), int(y)) for x, y in zip(c1, r2)}
    mx = np.iinfo(np.int64).max
    assert (7, mx) in rows and len(rows) == 11
    np.testing.assert_array_equal(nulls, [len(left_idx) - 1])


def test_outer_join_large_conservation(fa):
    import torch
    rng = np.random.default_rng(20)
    nl, nr = 1 << 24, 1 << 22
    l = rng.integers(0, 1 << 23, nl).astype(np.int64)
    r = rng.choice(1 << 23, nr, replace=False).astype(np.int64)
    dl, dr = _d(fa, l), _d(fa, r)
    li = fa.DeviceArray(torch.arange(nl, device="cuda"), np.uint64)
    ri = fa.DeviceArray(torch.arange(nr, device="cuda"), np.uint64)
    lo, ro, only, used = fa.outer_equi_join(dl, li, dr, ri)
    assert used
    hit = np.isin(l, r)
    assert len(lo) == int(hit.sum()) and len(only) == nl - int(hit.sum())
    np.testing.assert_array_equal(only.numpy(), np.flatnonzero(~hit))
    assert bool(torch.equal(dl.t[lo.t], dr.t[ro.t]))


# ---- star join ---------------------------------------------------------------------

@pytest.mark.parametrize("with_idx", [False, True])
def test_star_join_matches_oracle(fa, with_idx):
    rng = np.random.default_rng(23)
    for nl in [0, 1, 3000, 500000]:
        dims, left = [], []
        for nr, t in [(1000, np.int64), (77, np.int32), (20000, np.uint64), (5, np.uint32)]:
            keys = rng.choice(3 * nr, nr, replace=False).astype(t)
            ids = rng.permutation(nr).astype(np.uint64)
            dims.append((keys, ids))
            left.append(rng.integers(0, 3 * nr, nl).astype(t))
        lidx = (np.sort(rng.choice(nl, nl // 2, replace=False)).astype(np.uint64)
                if with_idx else None)
        rlo, rros = oracle.star_join(left, dims, lidx)
        lo, ros = fa.star_join([_d(fa, c) for c in left],
                               [(_d(fa, k), _d(fa, i)) for k, i in dims],
                               left_idx=_d(fa, lidx) if with_idx else None)
        np.testing.assert_array_equal(lo.numpy(), rlo, err_msg=f"nl={nl}")
        for r, rr in zip(ros, rros):
            np.testing.assert_array_equal(r.numpy(), rr)


def test_star_join_errors_and_empty_dimension(fa):
    left = _d(fa, np.arange(10, dtype=np.int64))
    with pytest.raises(fa.FhipError, match="not unique"):
        fa.star_join([left], [(_d(fa, np.array([1, 1], np.int64)),
                               _d(fa, np.arange(2, dtype=np.uint64)))])
    lo, ros = fa.star_join([left], [(_d(fa, np.zeros(0, np.int64)),
                                     _d(fa, np.zeros(0, np.uint64)))])
    assert len(lo) == 0 and len(ros[0]) == 0


@pytest.mark.parametrize("t", [np.int32, np.uint32, np.int64, np.uint64])
def test_star_join_dense_dimensions(fa, t):
    # dimensions of 2^17 unique keys within twice their rows, with row ids
    # ids[0] + i: probed through a direct array (negative keys for signed
    # types, keys near the top of the range for unsigned ones); a dimension
    # with shuffled ids or a wide key range takes the hash table; the same
    # rows as the oracle either way, and a repeated key still raises
    rng = np.random.default_rng(91)
    nl, nr = 300000, 1 << 17
    signed = np.dtype(t).kind == "i"
    base = -(1 << 16) if signed else np.iinfo(t).max - 3 * nr
    dims, left = [], []
    def shift(a):  # a + base in the key type (two's complement wrap for uint32)
        if signed:
            return (a.astype(np.int64) + base).astype(t)
        return (a.astype(np.uint64) + np.uint64(base)).astype(t)
    for kind in ("dense", "dense", "shuffled_ids", "wide"):
        span = 2 * nr if kind != "wide" else 64 * nr
        keys = shift(rng.choice(span, nr, replace=False))
        ids = (np.arange(nr, dtype=np.uint64) + np.uint64(11) if kind != "shuffled_ids"
               else rng.permutation(nr).astype(np.uint64))
        dims.append((keys, ids))
        left.append(shift(rng.integers(0, span + 100, nl) - 50))  # some below the keys
    rlo, rros = oracle.star_join(left, dims, None)
    lo, ros = fa.star_join([_d(fa, c) for c in left], [(_d(fa, k), _d(fa, i)) for k, i in dims])
    np.testing.assert_array_equal(lo.numpy(), rlo)
    for r, rr in zip(ros, rros):
        np.testing.assert_array_equal(r.numpy(), rr)
    keys = shift(np.arange(nr))
    keys[7] = keys[9]
    with pytest.raises(fa.FhipError, match="not unique"):
        fa.star_join([_d(fa, left[0])], [(_d(fa, keys), _d(fa, np.arange(nr, dtype=np.uint64)))])

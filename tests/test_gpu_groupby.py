"""GPU parity of group_by and grouped aggregation against the CPU oracle.

Bit-exact: local_grouped_idx, local_idx_split, group keys, counts, integer
sums, max/min.  FP64 sums and means: |gpu - ref| <= 1e-12 * max(|ref|,
sum|x|) per group -- the reference's own fold carries rounding error of that
order, so the bound is taken relative to the group's absolute mass, which
stays meaningful for groups whose sum cancels to near zero.
"""
import json
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
FP64_TOL = 1e-12
FP32_TOL = 2e-6


@pytest.fixture(scope="module")
def fa():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import frovedis_amd
    frovedis_amd.lib()
    return frovedis_amd


def _abs_mass(v, gi, split):
    a = np.abs(v[gi.astype(np.int64)])
    c = np.concatenate([[0.0], np.cumsum(a)])
    return c[split[1:].astype(np.int64)] - c[split[:-1].astype(np.int64)]


def _check_fp(got, ref, mass, tol):
    bound = tol * np.maximum(np.abs(ref), mass)
    bad = np.abs(got - ref) > bound
    assert not bad.any(), (np.flatnonzero(bad)[:5], got[bad][:5], ref[bad][:5])


@pytest.mark.parametrize("kt", [np.int32, np.uint32, np.int64, np.uint64, np.float32, np.float64])
def test_group_by_matches_oracle(fa, kt):
    rng = np.random.default_rng(11)
    for n in [0, 1, 63, 64, 5000, 1 << 20]:
        if np.dtype(kt).kind == "f":
            k = rng.integers(-500, 500, n).astype(kt) * kt(0.25)
        elif np.dtype(kt).kind == "u":
            k = rng.integers(0, 3000, n).astype(kt)
        else:
            k = rng.integers(-1500, 1500, n).astype(kt)
        gi, split, keys = oracle.group_by(k)
        dk = fa.DeviceArray.from_numpy(k)
        ggi, gsplit, gkeys = fa.group_by(dk)
        assert np.array_equal(ggi.numpy(), gi), n
        assert np.array_equal(gsplit.numpy(), split), n
        assert np.array_equal(gkeys.numpy(), keys), n


def test_group_by_with_local_idx(fa):
    # local_idx need not be iota (a filtered table): the key column is read
    # through it (dfcolumn_groupby.cc:39)
    rng = np.random.default_rng(12)
    col = rng.integers(0, 100, 50000).astype(np.int64)
    idx = np.sort(rng.choice(50000, 20000, replace=False)).astype(np.uint64)
    gi, split, keys = oracle.group_by(col, idx)
    ggi, gsplit, gkeys = fa.group_by(fa.DeviceArray.from_numpy(col), fa.DeviceArray.from_numpy(idx))
    assert np.array_equal(ggi.numpy(), gi)
    assert np.array_equal(gsplit.numpy(), split)
    assert np.array_equal(gkeys.numpy(), keys)


@pytest.mark.parametrize("vt", [np.int32, np.int64, np.uint64, np.float32, np.float64])
def test_grouped_agg_matches_oracle(fa, vt):
    rng = np.random.default_rng(13)
    n = 300000
    # a skewed key so that some groups exceed the one-wave limit (8192 rows)
    k = np.where(rng.random(n) < 0.3, 7, rng.integers(0, 2000, n)).astype(np.int64)
    if np.dtype(vt).kind == "f":
        v = rng.standard_normal(n).astype(vt)
    elif np.dtype(vt).kind == "u":
        v = rng.integers(0, 1 << 40, n).astype(vt)
    else:
        v = rng.integers(-(1 << 30), 1 << 30, n).astype(vt)
    gi, split, keys = oracle.group_by(k)
    dgi, dsplit = fa.DeviceArray.from_numpy(gi), fa.DeviceArray.from_numpy(split)
    dv = fa.DeviceArray.from_numpy(v)
    for nulls in [None, np.sort(rng.choice(n, 5000, replace=False)).astype(np.uint64)]:
        dn = None if nulls is None else fa.DeviceArray.from_numpy(nulls)
        for op in ["sum", "count", "avg", "max", "min"]:
            ref = oracle.grouped_agg(op, v, gi, split, nulls)
            got = fa.grouped_agg(op, dv, dgi, dsplit, dn).numpy()
            if op in ("count", "max", "min") or np.dtype(vt).kind != "f" and op == "sum":
                assert np.array_equal(got, ref), (op, nulls is None)
            else:
                vv = v.astype(np.float64)
                if nulls is not None:
                    vv = vv.copy()
                    vv[nulls.astype(np.int64)] = 0
                mass = _abs_mass(vv, gi, split)
                if op == "avg":
                    cnt = oracle.grouped_agg("count", v, gi, split, nulls).astype(np.float64)
                    mass = mass / np.maximum(cnt, 1)
                tol = FP32_TOL if vt == np.float32 and op == "sum" else FP64_TOL
                if vt == np.float32 and op == "avg":
                    tol = FP32_TOL
                _check_fp(got, ref, mass, tol)


def test_tutorial_group_by_on_gpu(fa):
    import csv
    rows = list(csv.reader(open(os.path.join(GOLD, "tut5_6_t.csv"))))
    kat = json.load(open(os.path.join(GOLD, "tutorial_known_answers.json")))["group_by_c1_agg_c3"]
    c1 = fa.DeviceArray.from_numpy(np.array([int(r[0]) for r in rows], np.int32))
    c3 = fa.DeviceArray.from_numpy(np.array([int(r[2]) for r in rows], np.int32))
    gi, split, keys = fa.group_by(c1)
    assert keys.numpy().tolist() == kat["c1"]
    for op in ["sum", "count", "avg", "max", "min"]:
        assert fa.grouped_agg(op, c3, gi, split).numpy().tolist() == kat[op], op
    gk, outs = fa.group_by_agg(c1, [("sum", c3), ("count", c3), ("avg", c3), ("max", c3), ("min", c3)])
    assert gk.numpy().tolist() == kat["c1"]
    for op, o in zip(["sum", "count", "avg", "max", "min"], outs):
        assert o.numpy().tolist() == kat[op], op


def _zipf_case(n, seed_shift=0):
    from frovedis_amd import datagen as dg
    k = dg.zipf_keys_np(n, start=seed_shift)
    v = dg.normal_np(n, start=seed_shift)
    iv = dg.int_uniform_np(n, 1000, start=seed_shift)
    return k, v, iv


@pytest.mark.parametrize("n", [1000, 1 << 16, 1 << 21])
def test_group_by_agg_zipf_matches_oracle(fa, n):
    k, v, iv = _zipf_case(n)
    gi, split, keys = oracle.group_by(k)
    ref_sum = oracle.grouped_agg("sum", v, gi, split)
    ref_cnt = oracle.grouped_agg("count", v, gi, split)
    ref_avg = oracle.grouped_agg("avg", v, gi, split)
    ref_max = oracle.grouped_agg("max", iv, gi, split)
    dk, dv, div = (fa.DeviceArray.from_numpy(a) for a in (k, v, iv))
    gk, (s, c, a, m) = fa.group_by_agg(dk, [("sum", dv), ("count", dv), ("avg", dv), ("max", div)])
    assert np.array_equal(gk.numpy(), keys)
    assert np.array_equal(c.numpy(), ref_cnt)
    assert np.array_equal(m.numpy(), ref_max)
    mass = _abs_mass(v, gi, split)
    _check_fp(s.numpy(), ref_sum, mass, FP64_TOL)
    _check_fp(a.numpy(), ref_avg, mass / ref_cnt, FP64_TOL)


def test_group_by_agg_table_growth_and_special_keys(fa):
    # many distinct keys with a tiny table hint: the table is rebuilt larger;
    # INT64_MIN (the empty-slot marker) and INT64_MAX are ordinary keys
    rng = np.random.default_rng(21)
    n = 400000
    k = rng.integers(-(1 << 62), 1 << 62, n).astype(np.int64) // 3
    k[::1000] = np.iinfo(np.int64).min
    k[5::1000] = np.iinfo(np.int64).max
    v = rng.integers(-1000, 1000, n).astype(np.int64)
    gi, split, keys = oracle.group_by(k)
    dk, dv = fa.DeviceArray.from_numpy(k), fa.DeviceArray.from_numpy(v)
    gk, (s, c, mx, mn) = fa.group_by_agg(dk, [("sum", dv), ("count", None), ("max", dv), ("min", dv)],
                                         table_hint=16, capacity=100)
    assert np.array_equal(gk.numpy(), keys)
    assert np.array_equal(s.numpy(), oracle.grouped_agg("sum", v, gi, split))
    assert np.array_equal(c.numpy(), oracle.grouped_agg("count", v, gi, split))
    assert np.array_equal(mx.numpy(), oracle.grouped_agg("max", v, gi, split))
    assert np.array_equal(mn.numpy(), oracle.grouped_agg("min", v, gi, split))


def test_group_by_agg_many_partitions(fa):
    # a table hint of 2^23 groups asks for 4096 partitions, beyond what the
    # wide scatter tile leaves LDS for: the narrow tile takes over
    rng = np.random.default_rng(25)
    n = 1 << 20
    k = rng.integers(0, 1 << 21, n).astype(np.int64)
    v = rng.integers(-1000, 1000, n).astype(np.int64)
    gi, split, keys = oracle.group_by(k)
    dk, dv = fa.DeviceArray.from_numpy(k), fa.DeviceArray.from_numpy(v)
    gk, (s, c, mx) = fa.group_by_agg(dk, [("sum", dv), ("count", None), ("max", dv)],
                                     table_hint=1 << 23)
    assert np.array_equal(gk.numpy(), keys)
    assert np.array_equal(s.numpy(), oracle.grouped_agg("sum", v, gi, split))
    assert np.array_equal(c.numpy(), oracle.grouped_agg("count", v, gi, split))
    assert np.array_equal(mx.numpy(), oracle.grouped_agg("max", v, gi, split))


def test_group_by_agg_full_size_cross_check(fa):
    # 2^26 Zipf(1.1) rows generated on the device: the hash path and the sort
    # path are independent implementations; keys/counts/max must agree
    # exactly and the sums to the FP64 bound; counts add up to n
    import torch
    from frovedis_amd import datagen as dg
    n = 1 << 26
    k = fa.DeviceArray(dg.zipf_keys_dev(n), np.int64)
    v = fa.DeviceArray(dg.normal_dev(n), np.float64)
    iv = fa.DeviceArray(dg.int_uniform_dev(n, 1000), np.int32)
    gk, (s, c, m) = fa.group_by_agg(k, [("sum", v), ("count", v), ("max", iv)])
    gi, split, keys = fa.group_by(k)
    assert torch.equal(gk.t, keys.t)
    assert int(c.t.sum()) == n
    assert torch.equal(c.t, (split.t[1:] - split.t[:-1]))
    assert torch.equal(m.t, fa.grouped_agg("max", iv, gi, split).t)
    s2 = fa.grouped_agg("sum", v, gi, split).t
    absm = fa.grouped_agg("sum", fa.DeviceArray(v.t.abs(), np.float64), gi, split).t
    assert bool(((s.t - s2).abs() <= 2 * FP64_TOL * torch.maximum(s2.abs(), absm)).all())


def test_group_by_agg_min_max_small_and_ragged(fa):
    # chunks whose row count is not a multiple of the wave width, and waves
    # where every row has the same key (the wave-reduction path)
    rng = np.random.default_rng(23)
    for n, nk in [(1, 1), (8, 3), (65, 2), (200, 3), (4099, 5), (100000, 50)]:
        k = rng.integers(0, nk, n).astype(np.int64)
        for dt in [np.int32, np.int64, np.float64]:
            v = rng.integers(-500, 500, n).astype(dt)
            gk, (mn, mx, sm) = fa.group_by_agg(fa.DeviceArray.from_numpy(k),
                                               [("min", fa.DeviceArray.from_numpy(v)),
                                                ("max", fa.DeviceArray.from_numpy(v)),
                                                ("sum", fa.DeviceArray.from_numpy(v))])
            keys = np.unique(k)
            assert np.array_equal(gk.numpy(), keys)
            assert np.array_equal(mn.numpy(), [v[k == x].min() for x in keys]), (n, dt)
            assert np.array_equal(mx.numpy(), [v[k == x].max() for x in keys]), (n, dt)
            assert np.array_equal(sm.numpy(), [v[k == x].sum() for x in keys]), (n, dt)


@pytest.mark.parametrize("kt,base", [(np.int64, -(1 << 40)), (np.int64, -(1 << 31)),
                                     (np.uint64, (1 << 64) - (1 << 33)), (np.int64, (1 << 62))])
def test_group_by_agg_narrow_key_range(fa, kt, base):
    # keys within a 2^32-wide range far from zero (the partitioned copy holds
    # 32-bit offsets from the minimum), and one 2^33-wide range (full keys)
    rng = np.random.default_rng(29)
    n = 300000
    for width in [1 << 20, (1 << 32) - 1, 1 << 33]:
        off = rng.integers(0, width, n, dtype=np.uint64)
        if kt == np.uint64:
            k = (np.uint64(base) + off).astype(np.uint64)
        else:
            k = (np.int64(base) + off.astype(np.int64)).astype(np.int64)
        k[::7] = k[0]
        v = rng.integers(-1000, 1000, n).astype(np.int64)
        gi, split, keys = oracle.group_by(k)
        gk, (s, c, mx) = fa.group_by_agg(fa.DeviceArray.from_numpy(k),
                                         [("sum", fa.DeviceArray.from_numpy(v)), ("count", None),
                                          ("max", fa.DeviceArray.from_numpy(v))])
        assert np.array_equal(gk.numpy(), keys), width
        assert np.array_equal(s.numpy(), oracle.grouped_agg("sum", v, gi, split))
        assert np.array_equal(c.numpy(), oracle.grouped_agg("count", v, gi, split))
        assert np.array_equal(mx.numpy(), oracle.grouped_agg("max", v, gi, split))


def test_group_by_agg_heavy_keys_and_rest(fa):
    # a skewed column (heavy keys aggregated in one pass, the other rows
    # partitioned): Zipf heads plus uniform tails, INT64_MIN among the heavy
    # rows (never a heavy-table key), an all-heavy column, and the skewed
    # column sorted (runs of equal keys fill whole sample groups)
    from frovedis_amd import autotune as at
    rng = np.random.default_rng(31)
    n = 1 << 21
    from frovedis_amd import datagen as dg
    k = dg.zipf_keys_np(n, n_values=1 << 16) * 3 - 50000
    k[::9] = np.iinfo(np.int64).min
    v = dg.normal_np(n)
    iv = rng.integers(-7, 7, n).astype(np.int32)
    for kk in [k, (k % 5).astype(np.int64), np.sort(k)]:
        gi, split, keys = oracle.group_by(kk)
        for pct in [None, 101]:  # default, and the heavy-key pass switched off
            at.tune_set("groupby.hot_min_pct", pct)
            gk, (s, c, a, mx, mn) = fa.group_by_agg(
                fa.DeviceArray.from_numpy(kk),
                [("sum", fa.DeviceArray.from_numpy(v)), ("count", None),
                 ("avg", fa.DeviceArray.from_numpy(v)), ("max", fa.DeviceArray.from_numpy(iv)),
                 ("min", fa.DeviceArray.from_numpy(iv))])
            assert np.array_equal(gk.numpy(), keys)
            assert np.array_equal(c.numpy(), oracle.grouped_agg("count", v, gi, split))
            assert np.array_equal(mx.numpy(), oracle.grouped_agg("max", iv, gi, split))
            assert np.array_equal(mn.numpy(), oracle.grouped_agg("min", iv, gi, split))
            mass = _abs_mass(v, gi, split)
            _check_fp(s.numpy(), oracle.grouped_agg("sum", v, gi, split), mass, FP64_TOL)
            cnt = oracle.grouped_agg("count", v, gi, split).astype(np.float64)
            _check_fp(a.numpy(), oracle.grouped_agg("avg", v, gi, split), mass / cnt, FP64_TOL)
    at.tune_set("groupby.hot_min_pct", None)


@pytest.mark.parametrize("range_path", [1, 0])
@pytest.mark.parametrize("hint", [0, 16, 1 << 16, 1 << 23])
@pytest.mark.parametrize("aggs", [("count",), ("sum",), ("sum", "maxi"),
                                  ("sum", "maxv", "minv", "maxi")])
def test_group_by_agg_heavy_keys_accumulator_counts(fa, aggs, hint, range_path):
    """The heavy-key pass with zero to four distinct accumulators (its
    kernel is specialised on that count).  The keys fill a range of 2^16, so
    the cold rows take the range path unless it is switched off; on the
    hash path a table hint makes the pass also count the cold rows per
    partition (used as they are for 2^16 and for the 4096 partitions of
    2^23, dropped for the recount after the 16-group hint overflows)."""
    from frovedis_amd import autotune as at
    at.tune_set("groupby.range", range_path)
    try:
        _heavy_keys_accumulator_counts(fa, aggs, hint)
    finally:
        at.tune_set("groupby.range", None)


def _heavy_keys_accumulator_counts(fa, aggs, hint):
    from frovedis_amd import datagen as dg
    n = 1 << 21
    k = dg.zipf_keys_np(n, n_values=1 << 16)
    v = dg.normal_np(n)
    iv = np.random.default_rng(32).integers(-1000, 1000, n).astype(np.int32)
    dk, dv, div = (fa.DeviceArray.from_numpy(x) for x in (k, v, iv))
    spec = {"count": ("count", None), "sum": ("sum", dv), "maxi": ("max", div),
            "maxv": ("max", dv), "minv": ("min", dv)}
    gk, outs = fa.group_by_agg(dk, [spec[a] for a in aggs], table_hint=hint)
    gi, split, keys = oracle.group_by(k)
    assert np.array_equal(gk.numpy(), keys)
    mass = _abs_mass(v, gi, split)
    for name, out in zip(aggs, outs):
        got = out.numpy()
        if name == "count":
            assert np.array_equal(got, oracle.grouped_agg("count", v, gi, split))
        elif name == "sum":
            _check_fp(got, oracle.grouped_agg("sum", v, gi, split), mass, FP64_TOL)
        elif name == "maxi":
            assert np.array_equal(got, oracle.grouped_agg("max", iv, gi, split))
        elif name == "maxv":
            assert np.array_equal(got, oracle.grouped_agg("max", v, gi, split))
        else:
            assert np.array_equal(got, oracle.grouped_agg("min", v, gi, split))


def _prof_count(fa, name):
    import ctypes
    t, c = ctypes.c_double(0), ctypes.c_size_t(0)
    fa.check(fa.lib().fhip_prof_query(name.encode(), ctypes.byref(t), ctypes.byref(c)))
    return c.value


@pytest.mark.parametrize("hint", [0, 16])
def test_group_by_agg_low_cardinality_runs_heavy_pass_only(fa, hint):
    # five keys: the heavy-key plan covers every sampled row, so the one
    # heavy-key pass aggregates the whole column and the partitioned
    # aggregation never runs (the assembly of the heavy table alone)
    n = 1 << 21
    rng = np.random.default_rng(55)
    k = (rng.integers(0, 1 << 40, n) % 5).astype(np.int64) - 2
    v = rng.standard_normal(n)
    iv = rng.integers(-100, 100, n).astype(np.int32)
    fa.check(fa.lib().fhip_prof_enable(1))
    gk, (s, c, mx) = fa.group_by_agg(fa.DeviceArray.from_numpy(k),
                                     [("sum", fa.DeviceArray.from_numpy(v)), ("count", None),
                                      ("max", fa.DeviceArray.from_numpy(iv))], table_hint=hint)
    hot, agg = _prof_count(fa, "groupby_hot"), _prof_count(fa, "hashagg")
    fa.check(fa.lib().fhip_prof_enable(0))
    # (with a hint the scatter is queued before the cold-row count is back,
    # and runs over empty regions; no cold row is ever aggregated)
    assert hot == 1 and agg == 0
    gi, split, keys = oracle.group_by(k)
    assert np.array_equal(gk.numpy(), keys)
    assert np.array_equal(c.numpy(), oracle.grouped_agg("count", v, gi, split))
    assert np.array_equal(mx.numpy(), oracle.grouped_agg("max", iv, gi, split))
    _check_fp(s.numpy(), oracle.grouped_agg("sum", v, gi, split), _abs_mass(v, gi, split),
              FP64_TOL)


def test_group_by_agg_skewed_column_runs_heavy_pass(fa):
    # Zipf(1.1) over 2^16 values: the sampled heavy keys cover most rows
    from frovedis_amd import datagen as dg
    n = 1 << 21
    k = dg.zipf_keys_np(n, n_values=1 << 16)
    fa.check(fa.lib().fhip_prof_enable(1))
    gk, (c,) = fa.group_by_agg(fa.DeviceArray.from_numpy(k), [("count", None)])
    hot = _prof_count(fa, "groupby_hot")
    fa.check(fa.lib().fhip_prof_enable(0))
    assert hot == 1
    gi, split, keys = oracle.group_by(k)
    assert np.array_equal(gk.numpy(), keys)
    assert np.array_equal(c.numpy(), oracle.grouped_agg("count", k, gi, split))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_fp_max_min_with_signed_zeros_and_nan(fa, dt):
    # groups holding both -0.0 and +0.0 (and NaN rows, never selected):
    # the sort path reduces in row order, so it keeps the reference's first
    # zero bit for bit (`if (m < v)`, dfcolumn_grouped_agg.cc:366-452); the
    # hash path returns an equal value (either zero)
    n = 1 << 16
    rng = np.random.default_rng(66)
    k = rng.integers(0, 300, n).astype(np.int64)
    v = np.where(rng.random(n) < 0.5, dt(-0.0), dt(0.0)).astype(dt)
    v[rng.random(n) < 0.05] = np.nan
    v[rng.random(n) < 0.3] = -dt(1.5)
    gi, split, keys = oracle.group_by(k)
    for op in ("max", "min"):
        ref = oracle.grouped_agg(op, v, gi, split)
        got = fa.grouped_agg(op, fa.DeviceArray.from_numpy(v), fa.DeviceArray.from_numpy(gi),
                             fa.DeviceArray.from_numpy(split)).numpy()
        assert np.array_equal(got.view(np.uint32 if dt == np.float32 else np.uint64),
                              ref.view(np.uint32 if dt == np.float32 else np.uint64)), op
        gk, (h,) = fa.group_by_agg(fa.DeviceArray.from_numpy(k),
                                   [(op, fa.DeviceArray.from_numpy(v))])
        assert np.array_equal(gk.numpy(), keys)
        assert np.array_equal(h.numpy(), ref), op  # == on values: -0.0 == 0.0


@pytest.mark.parametrize("kt", [np.int32, np.uint32, np.int64, np.uint64])
@pytest.mark.parametrize("dense", [True, False])
def test_group_by_agg_dense_key_ranks(fa, kt, dense):
    # keys filling a narrow range (across zero for signed types, near the
    # top for unsigned ones): the group keys are ordered by ranks from flags
    # and a scan instead of a sort ("groupby.dense_ratio"; 0 switches it
    # off); both give the oracle's keys and aggregates, and so does a sparse
    # column of the same type
    from frovedis_amd import autotune as at
    rng = np.random.default_rng(41)
    n = 1 << 20
    info = np.iinfo(kt)
    dt = np.int64 if info.min < 0 else np.uint64
    lo = -5000 if info.min < 0 else int(info.max) - 300000
    cols = {"dense": rng.integers(lo, lo + 200000, n, dtype=dt),
            "edge": rng.integers(int(info.max) - 60000, int(info.max), n, endpoint=True, dtype=dt),
            "sparse": rng.integers(int(info.min), int(info.max), n, endpoint=True, dtype=dt)}
    at.tune_set("groupby.dense_ratio", None if dense else 0)
    try:
        for name, raw in cols.items():
            k = np.asarray(raw).astype(kt)
            k[::5] = k[3]  # a heavy key
            v = rng.integers(-1000, 1000, n).astype(np.int64)
            gi, split, keys = oracle.group_by(k)
            gk, (s, c, mx) = fa.group_by_agg(fa.DeviceArray.from_numpy(k),
                                             [("sum", fa.DeviceArray.from_numpy(v)),
                                              ("count", None),
                                              ("max", fa.DeviceArray.from_numpy(v))])
            assert np.array_equal(gk.numpy(), keys), name
            assert np.array_equal(s.numpy(), oracle.grouped_agg("sum", v, gi, split)), name
            assert np.array_equal(c.numpy(), oracle.grouped_agg("count", v, gi, split)), name
            assert np.array_equal(mx.numpy(), oracle.grouped_agg("max", v, gi, split)), name
    finally:
        at.tune_set("groupby.dense_ratio", None)

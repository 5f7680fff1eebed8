"""GPU parity of the fused group-by's range path against the CPU oracle.

When the heavy-key pass finds every key within a narrow range (at most the
row count and 2^21 values), the cold rows are partitioned by key offset and
aggregated into arrays indexed by it; the group keys then come out in order
without a sort.  The results must equal the reference's group_by + select
(dataframe/dfcolumn_groupby.cc:16-105, dfcolumn_grouped_agg.cc): keys,
counts, integer aggregates and max/min bit-exact, FP64 sums and means within
1e-12 of each group's absolute mass, and equal to the hash path's.
"""
import ctypes

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
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


@pytest.fixture
def tune(fa):
    from frovedis_amd import autotune
    yield autotune
    for name in ("groupby.range", "groupby.range_max", "groupby.range_chunks"):
        autotune.tune_set(name, None)


def _prof(fa, name):
    t, c = ctypes.c_double(0), ctypes.c_size_t(0)
    fa.check(fa.lib().fhip_prof_query(name.encode(), ctypes.byref(t), ctypes.byref(c)))
    return c.value


def _abs_mass(v, gi, split):
    a = np.abs(v[gi.astype(np.int64)].astype(np.float64))
    c = np.concatenate([[0.0], np.cumsum(a)])
    return c[split[1:].astype(np.int64)] - c[split[:-1].astype(np.int64)]


def _check_fp(got, ref, mass, tol):
    bound = tol * np.maximum(np.abs(ref), mass)
    bad = np.abs(got.astype(np.float64) - ref.astype(np.float64)) > bound
    assert not bad.any(), (np.flatnonzero(bad)[:5], got[bad][:5], ref[bad][:5])


def _keys(kt, n, n_values=1 << 16, stride=1):
    from frovedis_amd import datagen as dg
    z = dg.zipf_keys_np(n, n_values=n_values).astype(np.int64) * stride
    kt = np.dtype(kt)
    if kt.kind == "i":
        base = -(n_values * stride) // 2
    elif kt.itemsize == 4:
        base = (1 << 32) - n_values * stride - 7
    else:
        base = 1 << 45
    return (z + base).astype(kt) if kt.kind == "i" else (z.astype(np.uint64) + np.uint64(base)).astype(kt)


def _run(fa, k, spec, expect_range):
    fa.check(fa.lib().fhip_prof_enable(1))
    gk, outs = fa.group_by_agg(fa.DeviceArray.from_numpy(k), spec)
    ran = _prof(fa, "groupby_range_part")
    fa.check(fa.lib().fhip_prof_enable(0))
    assert ran == (1 if expect_range else 0)
    return gk.numpy(), [o.numpy() for o in outs]


def _check(k, got_keys, outs, names, cols):
    gi, split, keys = oracle.group_by(k)
    assert np.array_equal(got_keys, keys)
    for name, got in zip(names, outs):
        op, col = name
        if op == "count":
            assert np.array_equal(got, oracle.grouped_agg("count", k, gi, split))
            continue
        v = cols[col]
        ref = oracle.grouped_agg(op, v, gi, split)
        if v.dtype.kind == "f" and op in ("sum", "avg"):
            mass = _abs_mass(v, gi, split)
            if op == "avg":
                mass = mass / oracle.grouped_agg("count", v, gi, split)
            _check_fp(got, ref, mass, FP64_TOL if v.dtype == np.float64 else FP32_TOL)
        else:
            assert np.array_equal(got, ref), (op, col)


@pytest.mark.parametrize("kt", [np.int32, np.uint32, np.int64, np.uint64])
@pytest.mark.parametrize("n", [(1 << 21) + 12345, (3 << 20) + 99])
def test_range_path_matches_oracle(fa, kt, n):
    # a last heavy-key region shorter than the others when n is not a
    # multiple of 2^18; signed keys across zero, uint32 keys at the top
    from frovedis_amd import datagen as dg
    k = _keys(kt, n)
    rng = np.random.default_rng(n % 1000)
    cols = {"v": dg.normal_np(n), "i": rng.integers(-1000, 1000, n).astype(np.int32)}
    d = {c: fa.DeviceArray.from_numpy(x) for c, x in cols.items()}
    names = [("sum", "v"), ("count", None), ("avg", "v"), ("max", "i")]
    got_keys, outs = _run(fa, k, [(op, d[c] if c else None) for op, c in names], True)
    _check(k, got_keys, outs, names, cols)


@pytest.mark.parametrize("aggs", [(), ("sum_v",), ("sum_v", "max_i"),
                                  ("sum_v", "max_v", "min_v", "max_i"),
                                  ("sum_l", "min_u", "max_f", "sum_f")])
def test_range_path_accumulator_sets(fa, aggs):
    # 0 to 4 distinct accumulators of every value type the reference sums or
    # compares (int32, uint32, int64, float, double)
    from frovedis_amd import datagen as dg
    n = 1 << 21
    k = _keys(np.int64, n)
    rng = np.random.default_rng(7)
    cols = {"v": dg.normal_np(n), "i": rng.integers(-1000, 1000, n).astype(np.int32),
            "l": rng.integers(-(1 << 40), 1 << 40, n).astype(np.int64),
            "u": rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32),
            "f": rng.standard_normal(n).astype(np.float32)}
    d = {c: fa.DeviceArray.from_numpy(x) for c, x in cols.items()}
    names = [("count", None)] + [tuple(a.split("_")) for a in aggs]
    got_keys, outs = _run(fa, k, [(op, d[c] if c else None) for op, c in names], True)
    _check(k, got_keys, outs, names, cols)


def test_range_path_equals_hash_path(fa, tune):
    from frovedis_amd import datagen as dg
    n = (1 << 22) + 777
    k = _keys(np.int64, n, n_values=1 << 20)
    v = dg.normal_np(n)
    iv = np.random.default_rng(3).integers(0, 1000, n).astype(np.int32)
    dk, dv, div = (fa.DeviceArray.from_numpy(x) for x in (k, v, iv))
    spec = [("sum", dv), ("count", None), ("avg", dv), ("max", div)]
    rk, (rs, rc, ra, rm) = (lambda r: (r[0].numpy(), [o.numpy() for o in r[1]]))(
        fa.group_by_agg(dk, spec))
    tune.tune_set("groupby.range", 0)
    fa.check(fa.lib().fhip_prof_enable(1))
    hk, (hs, hc, ha, hm) = (lambda r: (r[0].numpy(), [o.numpy() for o in r[1]]))(
        fa.group_by_agg(dk, spec))
    assert _prof(fa, "groupby_range_part") == 0
    fa.check(fa.lib().fhip_prof_enable(0))
    assert np.array_equal(rk, hk) and np.array_equal(rc, hc) and np.array_equal(rm, hm)
    gi, split, _ = oracle.group_by(k)
    mass = _abs_mass(v, gi, split)
    _check_fp(rs, hs, 2 * mass, FP64_TOL)
    _check_fp(ra, ha, 2 * mass / rc, FP64_TOL)


@pytest.mark.parametrize("stride,n_values,expect", [(3, 1 << 15, True), (1, 1 << 21, True),
                                                    (64, 1 << 16, False)])
def test_range_path_gaps_and_limits(fa, stride, n_values, expect):
    # keys every `stride` values (empty offsets in every part); a range of
    # 2^21 (the limit); a range of 2^22 (past it: the hash path)
    n = 1 << 21
    k = _keys(np.int64, n, n_values=n_values, stride=stride)
    iv = np.random.default_rng(stride).integers(-5, 5, n).astype(np.int64)
    names = [("count", None), ("sum", "i"), ("min", "i")]
    got_keys, outs = _run(fa, k, [("count", None), ("sum", fa.DeviceArray.from_numpy(iv)),
                                  ("min", fa.DeviceArray.from_numpy(iv))], expect)
    _check(k, got_keys, outs, names, {"i": iv})


def test_range_path_chunks_and_capacity(fa, tune):
    # one chunk per part and many chunks per part give the same groups; too
    # small an output capacity reports the group count
    n = 1 << 21
    k = _keys(np.uint32, n)
    for chunks in (1, 3, 64):
        tune.tune_set("groupby.range_chunks", chunks)
        got_keys, (c,) = _run(fa, k, [("count", None)], True)
        gi, split, keys = oracle.group_by(k)
        assert np.array_equal(got_keys, keys)
        assert np.array_equal(c, oracle.grouped_agg("count", k, gi, split))
    # an output of 4 groups: the count comes back, the wrapper resizes once
    tune.tune_set("groupby.range_chunks", None)
    gk, (c,) = fa.group_by_agg(fa.DeviceArray.from_numpy(k), [("count", None)], capacity=4)
    assert np.array_equal(gk.numpy(), keys)
    assert np.array_equal(c.numpy(), oracle.grouped_agg("count", k, gi, split))


@pytest.mark.parametrize("kt,base", [(np.uint64, (1 << 33) - 5000), (np.int64, -(1 << 32) - 3000),
                                     (np.int64, -20000)])
def test_range_path_low_bits_across_2_32(fa, kt, base):
    # 64-bit keys whose range crosses a multiple of 2^32 (or zero): the
    # heavy-key pass keeps the cold keys' low 32 bits, and the range path
    # rebuilds every key from the range's smallest key
    from frovedis_amd import datagen as dg
    n = (1 << 21) + 5
    z = dg.zipf_keys_np(n, n_values=1 << 16).astype(np.int64)
    k = (z + base).astype(kt) if kt == np.int64 else (z.astype(np.uint64) + np.uint64(base))
    v = dg.normal_np(n)
    fa.check(fa.lib().fhip_prof_enable(1))
    gk, outs = fa.group_by_agg(fa.DeviceArray.from_numpy(k),
                               [("count", None), ("sum", fa.DeviceArray.from_numpy(v))])
    hot, ran = _prof(fa, "groupby_hot"), _prof(fa, "groupby_range_part")
    fa.check(fa.lib().fhip_prof_enable(0))
    assert (hot, ran) == (1, 1)
    _check(k, gk.numpy(), [o.numpy() for o in outs], [("count", None), ("sum", "v")], {"v": v})


@pytest.mark.parametrize("kt", [np.int64, np.uint64])
def test_narrow_heavy_keys_with_far_cold_keys(fa, kt):
    # heavy keys within a few thousand values, a few cold keys 2^40 away:
    # the heavy-key pass keeps the cold keys' low 32 bits for the range path,
    # finds the range too wide, and the group-by runs again with whole keys
    # on the hash path (two heavy-key passes, no range path)
    from frovedis_amd import datagen as dg
    n = 1 << 21
    k = dg.zipf_keys_np(n, n_values=1 << 12).astype(np.int64) + 100
    rng = np.random.default_rng(12)
    far = rng.choice(n, 300, replace=False)
    k[far] = (1 << 40) + rng.integers(0, 1 << 20, 300)
    k = k.astype(kt)
    iv = rng.integers(-50, 50, n).astype(np.int32)
    fa.check(fa.lib().fhip_prof_enable(1))
    gk, outs = fa.group_by_agg(fa.DeviceArray.from_numpy(k),
                               [("count", None), ("max", fa.DeviceArray.from_numpy(iv))])
    hot, ran = _prof(fa, "groupby_hot"), _prof(fa, "groupby_range_part")
    fa.check(fa.lib().fhip_prof_enable(0))
    assert (hot, ran) == (2, 0)
    _check(k, gk.numpy(), [o.numpy() for o in outs], [("count", None), ("max", "i")], {"i": iv})

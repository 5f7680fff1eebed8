"""Batched segments: every segment's output equals the reference's per-segment
radix_sort(key, val) plus the sequential group fold (the oracle), bit-exact
including the FP64 sums (one thread folds each group in row order); graph
replays are checked against the eager run, with eager work in between."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fa():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import frovedis_amd
    frovedis_amd.lib()
    return frovedis_amd


def _case(sizes, seed=1):
    rng = np.random.default_rng(seed)
    keys = np.concatenate([rng.integers(0, max(s // 4, 1), s) for s in sizes]).astype(np.int64)
    vals = rng.random(len(keys))
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    return keys, vals, off


def _check_against_oracle(res, keys, vals, off):
    ok = res["keys"].numpy()
    ov = res["vals"].numpy()
    gseg = res["group_seg_off"].numpy()
    G = res["ngroups"]
    gk = res["group_keys"].numpy()[:G]
    gs = res["group_sum"].numpy()[:G]
    gc = res["group_count"].numpy()[:G]
    assert gseg[0] == 0 and gseg[-1] == G
    for s in range(len(off) - 1):
        a, b = int(off[s]), int(off[s + 1])
        ek, ev = oracle.radix_sort(keys[a:b], vals[a:b].view(np.uint64))
        assert np.array_equal(ok[a:b], ek), s
        assert np.array_equal(ov[a:b].view(np.uint64), ev), s
        gi, split, ukeys = oracle.group_by(keys[a:b])
        g0, g1 = int(gseg[s]), int(gseg[s + 1])
        assert np.array_equal(gk[g0:g1], ukeys), s
        assert np.array_equal(gc[g0:g1], oracle.grouped_agg("count", vals[a:b], gi, split)), s
        # bit-exact: same additions in the same order as the sequential fold
        assert np.array_equal(gs[g0:g1], oracle.grouped_agg("sum", vals[a:b], gi, split)), s


@pytest.mark.parametrize("sizes", [[1], [256, 300, 1000], [5, 0, 70000, 1, 4096, 0, 3],
                                   [256] * 64 + [50000, 123457],
                                   [8191, 8192, 8193, 16384, 7168, 7169]])
def test_batch_matches_per_segment_oracle(fa, sizes):
    keys, vals, off = _case(np.array(sizes, np.int64))
    for use_graph in (False, True):
        res = fa.batch_sort_groupby(fa.DeviceArray.from_numpy(keys),
                                    fa.DeviceArray.from_numpy(vals),
                                    fa.DeviceArray.from_numpy(off), key_bits=18,
                                    use_graph=use_graph)
        _check_against_oracle(res, keys, vals, off)


def test_batch_graph_replays_match_eager(fa):
    import torch
    from frovedis_amd import datagen as dg
    sizes = dg.segment_sizes(nseg=512, total=1 << 22, lo=256, hi=1 << 16)
    keys, vals, off = dg.batch_inputs_dev(sizes)
    k = fa.DeviceArray(keys, np.int64)
    v = fa.DeviceArray(vals, np.float64)
    o = fa.DeviceArray(off, np.uint64)
    eager = fa.batch_sort_groupby(k, v, o, key_bits=16, use_graph=False)
    ref = {name: a.t.clone() for name, a in eager.items() if name != "ngroups"}
    out = None
    for rep in range(4):
        out = fa.batch_sort_groupby(k, v, o, key_bits=16, use_graph=True, out=out)
        torch.cuda.synchronize()
        for name, t in ref.items():
            G = eager["ngroups"]
            a, b = out[name].t, t
            if name.startswith("group_") and name != "group_seg_off":
                a, b = a[:G], b[:G]
            assert torch.equal(a, b), (rep, name)
        # unrelated eager work between replays
        x = torch.randn(1 << 20, device="cuda")
        torch.cuda.synchronize()
        del x
    assert out["ngroups"] == eager["ngroups"]


def test_batch_rejects_out_of_range_keys(fa):
    keys = np.array([1, 2, 1 << 20], np.int64)
    vals = np.zeros(3)
    off = np.array([0, 3], np.uint64)
    with pytest.raises(fa.FhipError):
        fa.batch_sort_groupby(fa.DeviceArray.from_numpy(keys), fa.DeviceArray.from_numpy(vals),
                              fa.DeviceArray.from_numpy(off), key_bits=18, use_graph=False)


def _run_and_check(fa, keys, vals, off, key_bits, use_graph=False):
    res = fa.batch_sort_groupby(fa.DeviceArray.from_numpy(keys), fa.DeviceArray.from_numpy(vals),
                                fa.DeviceArray.from_numpy(off), key_bits=key_bits,
                                use_graph=use_graph)
    _check_against_oracle(res, keys, vals, off)


def test_batch_overflowing_bucket_falls_back(fa):
    # a large segment whose keys fill one fine bin (every key equal) cannot be
    # cut into buckets that fit one LDS job: the composite path takes over
    rng = np.random.default_rng(3)
    sizes = np.array([20000, 300, 50000], np.int64)
    keys = np.concatenate([np.full(20000, 7), rng.integers(0, 75, 300),
                           rng.integers(0, 12500, 50000)]).astype(np.int64)
    vals = rng.random(len(keys))
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    _run_and_check(fa, keys, vals, off, 18)


def test_batch_wide_key_range_in_small_segment_falls_back(fa):
    # a small segment whose keys span more than 2^19 values
    rng = np.random.default_rng(4)
    keys = np.concatenate([[0, 1 << 25, 5, 1 << 25], rng.integers(0, 1000, 3000)]).astype(np.int64)
    vals = rng.random(len(keys))
    off = np.array([0, 4, len(keys)], np.uint64)
    _run_and_check(fa, keys, vals, off, 30)


@pytest.mark.parametrize("digit_bits", [3, 5, 8])
def test_batch_digit_width_and_bucket_rows(fa, digit_bits):
    # the tunables the autotuner sets: LDS sort digit width and bucket rows
    # per size class; every setting gives the reference's result
    from frovedis_amd import autotune as at
    from frovedis_amd import datagen as dg
    sizes = dg.segment_sizes(nseg=64, total=1 << 20, lo=256, hi=1 << 17)
    rng = np.random.default_rng(digit_bits)
    keys, vals, off = _case(sizes, seed=digit_bits)
    at.tune_set("batch.digit_bits", digit_bits)
    at.tune_set("batch.bucket_rows", 2048 if digit_bits == 3 else 6144)
    at.tune_set("batch.bucket_rows.c17", 1024)
    try:
        for use_graph in (False, True):
            _run_and_check(fa, keys, vals, off, 18, use_graph)
    finally:
        for name in ("batch.digit_bits", "batch.bucket_rows", "batch.bucket_rows.c17"):
            at.tune_set(name, None)


def test_batch_composite_path_matches_oracle(fa):
    from frovedis_amd import autotune as at
    keys, vals, off = _case(np.array([256, 9000, 3, 70000], np.int64), seed=9)
    at.tune_set("batch.path", 1)
    try:
        for use_graph in (False, True):
            _run_and_check(fa, keys, vals, off, 18, use_graph)
    finally:
        at.tune_set("batch.path", None)


@pytest.mark.parametrize("rank_max", [None, 0, 5, 8192])
def test_batch_counting_and_lsd_orders(fa, rank_max):
    # jobs whose keys span at most 4096 values are sorted by counting, unless
    # a key holds more than batch.rank_max rows; wider spans take the LSD
    # passes (with and without the early group count): every mix must give
    # the reference's order and folds
    from frovedis_amd import autotune as at
    rng = np.random.default_rng(11 if rank_max is None else rank_max + 12)
    parts = [rng.integers(0, 4096, 8192),          # span 4096: counting, full job
             np.r_[0, 4095, rng.integers(0, 4096, 5000)],   # range 4095: counting
             np.r_[0, 4096, rng.integers(0, 4097, 5000)],   # range 4096: LSD
             rng.integers(0, 10, 3000),            # ~300 rows per key
             np.full(777, 12345),                  # one key
             rng.integers(0, 40000, 6000),         # past the early bitmap's span
             rng.integers(1000, 1200, 900) * 16,   # sparse keys, span 3200
             rng.integers(0, 3000, 30000)]         # a large segment in buckets
    keys = np.concatenate(parts).astype(np.int64)
    vals = rng.standard_normal(len(keys))
    off = np.concatenate([[0], np.cumsum([len(p) for p in parts])]).astype(np.uint64)
    at.tune_set("batch.rank_max", rank_max)
    try:
        for use_graph in (False, True):
            _run_and_check(fa, keys, vals, off, 18, use_graph)
    finally:
        at.tune_set("batch.rank_max", None)

"""GPU parity of the dense-key hash join against the CPU oracle.

A unique integer build column whose keys fill their range (at most
join.dense_ratio x the rows) is joined through the partitioned direct table;
the pairs must equal the reference's hash join (dataframe/join.hpp:304-326:
probe order, the partner's right row id), and a repeated key must send AUTO
to the sort-merge path as the reference's is_unique_ok does.  The size
threshold is lowered so that the path runs at sizes the oracle checks in
seconds.
"""
import ctypes
import zlib

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


@pytest.fixture
def dense(fa):
    from frovedis_amd import autotune
    autotune.tune_set("join.dense_min", 1024)
    yield autotune
    for name in ("join.dense_min", "join.dense_slice_bits", "join.dense_chunks",
                 "join.dense_ratio", "join.dense_lds_build"):
        autotune.tune_set(name, None)


def _prof_count(fa, name):
    t, c = ctypes.c_double(0), ctypes.c_size_t(0)
    fa.check(fa.lib().fhip_prof_query(name.encode(), ctypes.byref(t), ctypes.byref(c)))
    return c.value


def _case(rng, dt, nr, nl, span, dup=False, ids="affine", hot=False):
    dt = np.dtype(dt)
    R = max(nr, int(nr * span))
    if dt.kind == "i":
        base = -(R // 3)
    elif dt.itemsize == 4:
        base = (1 << 32) - R - 5  # near the top of the type
    else:
        base = 1 << 40
    if R > nr:  # both ends present, so the range is exactly R
        off = rng.permutation(np.concatenate(
            [[0, R - 1], 1 + rng.choice(R - 2, nr - 2, replace=False)]))
    else:
        off = rng.permutation(nr)
    r = (np.int64(base) + off.astype(np.int64)).astype(dt) if dt.kind == "i" else \
        (np.uint64(base) + off.astype(np.uint64)).astype(dt)
    if dup:
        r[nr // 2] = r[3]
    if ids == "affine":
        ri = np.arange(nr, dtype=np.uint64) + np.uint64(123456789)
    elif ids == "node":  # global row ids: a node id in the top bits
        ri = (np.uint64(3) << np.uint64(48)) | rng.permutation(nr).astype(np.uint64)
    else:
        ri = rng.permutation(nr).astype(np.uint64) + np.uint64(100000)
    lo_key, hi_key = int(r.min()), int(r.max())
    pick = rng.random(nl)
    l = rng.choice(r, nl).astype(np.int64)
    info = np.iinfo(dt)
    below = rng.integers(max(info.min, lo_key - 1000), lo_key, nl, endpoint=False) \
        if lo_key > info.min else l
    above = rng.integers(hi_key + 1, min(info.max, hi_key + 1000) + 1, nl) \
        if hi_key < info.max else l
    inside = rng.integers(lo_key, hi_key + 1, nl)  # holes when span > 1
    l = np.where(pick < 0.1, below, l)
    l = np.where((pick >= 0.1) & (pick < 0.2), above, l)
    l = np.where((pick >= 0.2) & (pick < 0.3), inside, l).astype(dt)
    if hot:  # one key on a third of the rows: segments far longer than a wave
        l[rng.random(nl) < 0.33] = r[7]
    li = rng.permutation(nl).astype(np.uint64) + np.uint64(7)
    return l, li, r, ri


def _run(fa, l, li, r, ri, strategy="auto", capacity=None):
    d = fa.DeviceArray.from_numpy
    return fa.equi_join(d(l), d(li), d(r), d(ri), strategy=strategy, capacity=capacity)


@pytest.mark.parametrize("dt", [np.int32, np.uint32, np.int64, np.uint64])
@pytest.mark.parametrize("span", [1.0, 1.5, 2.0])
@pytest.mark.parametrize("ids", ["affine", "shuffled", "node"])
def test_dense_join_matches_oracle(fa, dense, dt, span, ids):
    rng = np.random.default_rng(zlib.crc32(f"{np.dtype(dt).str}-{span}-{ids}".encode()))
    for nr, nl in [(1500, 0), (1500, 1), (5000, 70001), (200000, 300000)]:
        l, li, r, ri = _case(rng, dt, nr, nl, span, ids=ids)
        elo, ero, eh = oracle.equi_join(l, li, r, ri, "ve")
        fa.check(fa.lib().fhip_prof_enable(1))
        lo, ro, used = _run(fa, l, li, r, ri)
        ran = _prof_count(fa, "join_dense_build")
        fa.check(fa.lib().fhip_prof_enable(0))
        assert ran == 1, (nr, nl)
        assert used and eh
        assert np.array_equal(lo.numpy(), elo), (nr, nl)
        assert np.array_equal(ro.numpy(), ero), (nr, nl)


def test_dense_join_build_last_wave_with_few_rows(fa, dense):
    # the build's last round of rows leaves fewer valid lanes than tiles, so
    # lanes holding rows of late tiles read segment sums from lanes whose
    # own rows are past the end: every lane must take part in those
    # shuffles (seed 0 makes that happen for 13 build tiles in 4 parts)
    rng = np.random.default_rng(0)
    l, li, r, ri = _case(rng, np.int64, 200000, 300000, 2.0)
    elo, ero, eh = oracle.equi_join(l, li, r, ri, "ve")
    lo, ro, used = _run(fa, l, li, r, ri)
    assert used and eh
    assert np.array_equal(lo.numpy(), elo)
    assert np.array_equal(ro.numpy(), ero)


@pytest.mark.parametrize("lds_build", [1, 0])
@pytest.mark.parametrize("sbits,chunks", [(10, 1), (10, 3), (12, 128), (17, 7)])
def test_dense_join_parts_and_chunks(fa, dense, sbits, chunks, lds_build):
    # many parts (2^10 offsets each) and uneven tile chunks; a hot key on a
    # third of the probe rows; the table built in LDS parts or scattered
    # into the probe's parts and verified
    dense.tune_set("join.dense_slice_bits", sbits)
    dense.tune_set("join.dense_chunks", chunks)
    dense.tune_set("join.dense_lds_build", lds_build)
    rng = np.random.default_rng(sbits * 100 + chunks)
    for nr, nl, hot in [(100000, 250000, False), (70000, 200003, True)]:
        l, li, r, ri = _case(rng, np.int64, nr, nl, 1.3, hot=hot)
        elo, ero, eh = oracle.equi_join(l, li, r, ri, "ve")
        lo, ro, used = _run(fa, l, li, r, ri)
        assert used and eh
        assert np.array_equal(lo.numpy(), elo)
        assert np.array_equal(ro.numpy(), ero)


@pytest.mark.parametrize("lds_build", [1, 0])
@pytest.mark.parametrize("dt", [np.int32, np.uint64])
def test_dense_join_repeated_key_falls_back(fa, dense, dt, lds_build):
    # a repeated build key: AUTO takes the sort-merge path (the reference's
    # is_unique_ok), HASH is an error
    dense.tune_set("join.dense_lds_build", lds_build)
    rng = np.random.default_rng(5)
    l, li, r, ri = _case(rng, dt, 50000, 80000, 1.0, dup=True)
    elo, ero, eh = oracle.equi_join(l, li, r, ri, "ve")
    assert not eh
    lo, ro, used = _run(fa, l, li, r, ri)
    assert not used
    assert np.array_equal(lo.numpy(), elo) and np.array_equal(ro.numpy(), ero)
    # no sampled duplicate test: the dense table's build finds the repeat
    fa.check(fa.lib().fhip_prof_enable(1))
    lo, ro, used = _run(fa, l, li, r, ri, strategy="auto_unique")
    built, probed = _prof_count(fa, "join_dense_build"), _prof_count(fa, "join_dense_probe")
    fa.check(fa.lib().fhip_prof_enable(0))
    assert (built, probed) == (1, 0)
    assert not used
    assert np.array_equal(lo.numpy(), elo) and np.array_equal(ro.numpy(), ero)
    with pytest.raises(fa.FhipError):
        _run(fa, l, li, r, ri, strategy="hash")


def test_dense_join_capacity_and_wide_range(fa, dense):
    rng = np.random.default_rng(9)
    l, li, r, ri = _case(rng, np.int64, 60000, 90000, 1.0)
    elo, ero, _ = oracle.equi_join(l, li, r, ri, "ve")
    # too small an output: the count comes back, the wrapper retries
    lo, ro, used = _run(fa, l, li, r, ri, capacity=10)
    assert used and np.array_equal(lo.numpy(), elo) and np.array_equal(ro.numpy(), ero)
    # a range wider than dense_ratio x rows: the open-addressing table runs
    dense.tune_set("join.dense_ratio", 1)
    l2, li2, r2, ri2 = _case(rng, np.int64, 60000, 90000, 1.5)
    elo, ero, _ = oracle.equi_join(l2, li2, r2, ri2, "ve")
    fa.check(fa.lib().fhip_prof_enable(1))
    lo, ro, used = _run(fa, l2, li2, r2, ri2)
    ran = _prof_count(fa, "join_dense_build")
    fa.check(fa.lib().fhip_prof_enable(0))
    assert ran == 0
    assert used and np.array_equal(lo.numpy(), elo) and np.array_equal(ro.numpy(), ero)


@pytest.mark.parametrize("span", [1.0, 1.7])
def test_dense_outer_join_matches_oracle(fa, dense, span):
    rng = np.random.default_rng(int(span * 10))
    l, li, r, ri = _case(rng, np.int64, 40000, 100000, span, ids="shuffled")
    elo, ero, eol, eh = oracle.outer_equi_join(l, li, r, ri, "ve")
    d = fa.DeviceArray.from_numpy
    fa.check(fa.lib().fhip_prof_enable(1))
    lo, ro, only, used = fa.outer_equi_join(d(l), d(li), d(r), d(ri))
    ran = _prof_count(fa, "join_dense_build")
    fa.check(fa.lib().fhip_prof_enable(0))
    assert ran == 1 and used and eh
    assert np.array_equal(lo.numpy(), elo)
    assert np.array_equal(ro.numpy(), ero)
    assert np.array_equal(only.numpy(), eol)

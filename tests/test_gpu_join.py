"""GPU parity of unique_hashtable and equi_join against the CPU oracle.

The GPU table keeps the reference's layout, so lookups, miss lists,
check_existence, is_unique_ok, the conflict count and the all_keys order
are compared bit-exact.  Join pairs are compared in the reference's own
orders (probe order for the hash join, (key, left, right) for sort-merge),
and canonically against the x86 multimap body.
"""
import csv
import json
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def fa():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import frovedis_amd
    frovedis_amd.lib()
    return frovedis_amd


def _keys(rng, dt, n, unique):
    dt = np.dtype(dt)
    hi = 4 * max(n, 1)
    if unique:
        k = rng.choice(hi, n, replace=False)
    else:
        k = rng.integers(0, max(n // 3, 1), n)
    k = k.astype(np.int64)
    if dt.kind == "i" and n:
        k[::5] = -k[::5] - 1
    if dt.kind == "f":
        return (k.astype(np.float64) * 0.5).astype(dt)
    return k.astype(dt)


@pytest.mark.parametrize("dt", [np.int32, np.uint32, np.int64, np.uint64, np.float32, np.float64])
@pytest.mark.parametrize("unique", [True, False])
def test_hashtable_matches_oracle(fa, dt, unique):
    rng = np.random.default_rng(31)
    for n in [0, 1, 100, 5000, 300000]:
        k = _keys(rng, dt, n, unique)
        v = rng.integers(0, 1 << 62, n).astype(np.uint64)
        ref = oracle.UniqueHashtable(k, v, check_unique=True)
        ht = fa.UniqueHashtable(fa.DeviceArray.from_numpy(k), fa.DeviceArray.from_numpy(v),
                                check_unique=True)
        assert ht.table_size == ref.table_size
        assert ht.conflict_size == ref.conflict_size
        assert ht.is_unique_ok == ref.is_unique_ok
        q = np.concatenate([k, _keys(rng, dt, n // 2 + 3, True)])
        rng.shuffle(q)
        dq = fa.DeviceArray.from_numpy(q)
        gv, gm = ht.lookup(dq)
        ev, em = ref.lookup(q)
        assert np.array_equal(gm.numpy(), em), n
        if ref.is_unique_ok:
            assert np.array_equal(gv.numpy(), ev), n
        assert np.array_equal(ht.check_existence(dq).numpy(), ref.check_existence(q))
        assert np.array_equal(ht.all_keys().numpy().view(np.uint8), ref.all_keys().view(np.uint8))


def test_hashtable_set_constructor(fa):
    rng = np.random.default_rng(32)
    k = rng.integers(0, 1000, 20000).astype(np.int64)
    ref = oracle.UniqueHashtable(k)
    ht = fa.UniqueHashtable(fa.DeviceArray.from_numpy(k))
    assert ht.conflict_size == ref.conflict_size
    q = rng.integers(-10, 1200, 5000).astype(np.int64)
    assert np.array_equal(ht.check_existence(fa.DeviceArray.from_numpy(q)).numpy(),
                          ref.check_existence(q))
    assert np.array_equal(ht.all_keys().numpy(), ref.all_keys())


@pytest.mark.parametrize("dt", [np.int32, np.int64, np.uint64, np.float64])
@pytest.mark.parametrize("unique_right", [True, False])
def test_equi_join_matches_oracle(fa, dt, unique_right):
    rng = np.random.default_rng(33)
    for nl, nr in [(0, 10), (10, 0), (1, 1), (3000, 2000), (200000, 100000)]:
        r = _keys(rng, dt, nr, unique_right)
        l = _keys(rng, dt, nl, False) if nl else np.zeros(0, dt)
        if nl and nr:
            l[: nl // 2] = rng.choice(r, nl // 2)
        li = rng.permutation(nl).astype(np.uint64) + np.uint64(7)
        ri = rng.permutation(nr).astype(np.uint64) + np.uint64(100000)
        elo, ero, eh = oracle.equi_join(l, li, r, ri, "ve")
        dl, dr = fa.DeviceArray.from_numpy(l), fa.DeviceArray.from_numpy(r)
        dli, dri = fa.DeviceArray.from_numpy(li), fa.DeviceArray.from_numpy(ri)
        lo, ro, used = fa.equi_join(dl, dli, dr, dri)
        assert used == eh, (nl, nr)
        assert np.array_equal(lo.numpy(), elo), (nl, nr)
        assert np.array_equal(ro.numpy(), ero), (nl, nr)
        # forced sort-merge gives the sort-merge order on any input
        slo, sro, used = fa.equi_join(dl, dli, dr, dri, strategy="sortmerge")
        assert not used
        got = sorted(zip(slo.numpy().tolist(), sro.numpy().tolist()))
        assert got == sorted(zip(elo.tolist(), ero.tolist()))
        # x86 body (std::unordered_multimap): same pair set
        xlo, xro, _ = oracle.equi_join(l, li, r, ri, "x86")
        assert got == sorted(zip(xlo.tolist(), xro.tolist()))


def test_tutorial_join_on_gpu(fa):
    t1 = list(csv.reader(open(os.path.join(GOLD, "tut5_5_t1.csv"))))
    t2 = list(csv.reader(open(os.path.join(GOLD, "tut5_5_t2.csv"))))
    kat = json.load(open(os.path.join(GOLD, "tutorial_known_answers.json")))
    t1c3 = fa.DeviceArray.from_numpy(np.array([int(r[2]) for r in t1], np.int32))
    t2c1 = fa.DeviceArray.from_numpy(np.array([int(r[0]) for r in t2], np.int32))
    lo, ro, used = fa.equi_join(t1c3, fa.DeviceArray.from_numpy(np.arange(8, dtype=np.uint64)),
                                t2c1, fa.DeviceArray.from_numpy(np.arange(4, dtype=np.uint64)))
    assert not used  # t2c1 has a duplicated key: sort-merge path
    rows = sorted([int(t1[a][0]), int(t1[a][1]), int(t1[a][2]), int(t2[b][0]), t2[b][1]]
                  for a, b in zip(lo.numpy(), ro.numpy()))
    assert rows == sorted(kat["hash_join_t1c3_eq_t2c1_rows"])


def test_equi_join_scaled_join_config_properties(fa):
    # the join configuration at 1/4 scale: unique build side = a permutation
    # of 0..nb-1, probe side 90% hits and 10% out-of-range misses
    import torch
    from frovedis_amd import datagen as dg
    nb, npb = 1 << 25, 1 << 27
    build = dg.perm_keys_dev(nb)
    pk = dg.fk_keys_dev(npb, domain=nb)
    # the device generators give the host generators' values
    assert np.array_equal(build[:4096].cpu().numpy(), dg.perm_keys_np(nb)[:4096])
    assert np.array_equal(pk[-4096:].cpu().numpy(), dg.fk_keys_np(npb, domain=nb)[-4096:])
    bk = fa.DeviceArray(build.to(torch.int64), np.int64)
    pkd = fa.DeviceArray(pk, np.int64)
    bi = fa.DeviceArray(torch.arange(nb, device="cuda", dtype=torch.int64), np.uint64)
    pi = fa.DeviceArray(torch.arange(npb, device="cuda", dtype=torch.int64), np.uint64)
    lo, ro, used = fa.equi_join(pkd, pi, bk, bi)  # probe = left, build = right
    assert used
    hits = int((pk < nb).sum())
    assert len(lo) == hits
    assert torch.equal(pk[lo.t], build[ro.t])
    assert bool((lo.t[1:] > lo.t[:-1]).all())  # probe order
    slo, sro, used = fa.equi_join(bk, bi, pkd, pi, strategy="sortmerge")
    # AUTO on the non-unique side (decided by the sampled duplicate test
    # without a table build) gives the same pairs in the same order
    alo, aro, aused = fa.equi_join(bk, bi, pkd, pi)
    assert not aused
    assert torch.equal(alo.t, slo.t) and torch.equal(aro.t, sro.t)
    del alo, aro
    assert len(slo) == hits
    assert torch.equal(build[slo.t], pk[sro.t])
    kk = build[slo.t]
    assert bool((kk[1:] >= kk[:-1]).all())  # sorted by key


@pytest.mark.parametrize("with_negative", [False, True])
def test_hashtable_float_signed_zero_and_nan_conflicts(fa, with_negative):
    # -0.0 and NaN in the conflict list: the list follows the radix sort's
    # flips (hashtable.hpp:1234), which all_keys exposes, and lookups use
    # the reference's lower_bound comparisons; compared with the oracle
    n = 2000
    rng = np.random.default_rng(88)
    ts = 8192 + 27
    keys = rng.random(n) * 1000.0
    neg0, nan = (1 << 63), 0x7FF8000000000000

    def colliding(bits, k):
        return np.array([bits % ts + ts * k], np.uint64).view(np.float64)[0]
    keys[0] = -0.0
    keys[1] = np.array([nan], np.uint64).view(np.float64)[0]
    keys[n - 1] = colliding(neg0, 1 << 20)
    keys[n - 2] = colliding(nan, 1 << 21)
    if with_negative:
        keys[2] = -5.0
        keys[n - 3] = colliding(int(keys[2:3].view(np.uint64)[0]), 1 << 22)
    vals = np.arange(n, dtype=np.uint64) + np.uint64(10)
    ref = oracle.UniqueHashtable(keys, vals)
    assert ref.table_size == ts
    ht = fa.UniqueHashtable(fa.DeviceArray.from_numpy(keys), fa.DeviceArray.from_numpy(vals))
    assert ht.conflict_size == ref.conflict_size
    assert np.array_equal(ht.all_keys().numpy().view(np.uint64), ref.all_keys().view(np.uint64))
    q = np.concatenate([keys, [0.0, -0.0, np.nan, 7.25]])
    gv, gm = ht.lookup(fa.DeviceArray.from_numpy(q))
    ev, em = ref.lookup(q)
    assert np.array_equal(gm.numpy(), em)
    assert np.array_equal(gv.numpy(), ev)
    assert np.array_equal(ht.check_existence(fa.DeviceArray.from_numpy(q)).numpy(),
                          ref.check_existence(q))


@pytest.mark.parametrize("dt", [np.int32, np.uint32, np.int64, np.uint64])
@pytest.mark.parametrize("case", ["low", "signed_span", "high", "apart"])
def test_sort_merge_narrow_path_order(fa, dt, case):
    # consecutive row ids on both sides: the sort-merge join sorts (low 32
    # key bits, 32-bit row) pairs when each side's keys share their higher
    # bits ("high": a 2^32 stretch far from zero; "apart": the two sides in
    # different stretches, so nothing matches), and the general key/row-id
    # sort otherwise ("signed_span": 64-bit keys on both sides of zero).
    # Pairs in the reference's sort-merge order (join.hpp:304-326), the left
    # rows without a match in stable key order (:352-378), bit for bit.
    rng = np.random.default_rng(71)
    nl, nr = 40000, 90000
    wide = np.dtype(dt).itemsize == 8
    signed = np.dtype(dt).kind == "i"
    base = {"low": 0, "signed_span": -20000 if signed else 0,
            "high": (1 << 40) if wide else (1 << 30), "apart": (1 << 40) if wide else 0}[case]
    l = rng.integers(0, 30000, nl).astype(np.int64) + base
    r = rng.integers(0, 30000, nr).astype(np.int64) + base
    if case == "apart":
        r = r + ((1 << 33) if wide else 50000)
    l, r = l.astype(dt), r.astype(dt)
    li = np.arange(nl, dtype=np.uint64) + np.uint64(5)
    ri = np.arange(nr, dtype=np.uint64) + np.uint64(1000)
    dl, dr = fa.DeviceArray.from_numpy(l), fa.DeviceArray.from_numpy(r)
    dli, dri = fa.DeviceArray.from_numpy(li), fa.DeviceArray.from_numpy(ri)
    elo, ero, _ = oracle.equi_join(l, li, r, ri, "ve_sortmerge")
    lo, ro, used = fa.equi_join(dl, dli, dr, dri, strategy="sortmerge")
    assert not used
    assert np.array_equal(lo.numpy(), elo) and np.array_equal(ro.numpy(), ero)
    olo, oro, only, oused = oracle.outer_equi_join(l, li, r, ri, "ve")
    assert not oused  # the right side repeats keys: the sort-merge body
    glo, gro, gonly, gused = fa.outer_equi_join(dl, dli, dr, dri, strategy="sortmerge")
    assert np.array_equal(glo.numpy(), olo) and np.array_equal(gro.numpy(), oro)
    assert np.array_equal(gonly.numpy(), only)

"""GPU parity on the benchmark configurations' own generators.

- join: perm_keys (unique left) and fk_keys (90 % hits, 10 % misses), at
  2^20 x 2^22 against the oracle's VE bodies bit-exact in the reference's
  orders (hash join: probe order; sort-merge: (key, left, right)), and at
  the full 2^27 x 2^29 through properties (pair count, key equality, probe
  order, sortedness);
- batched segments: segment_sizes(512 segments, 2^22 rows) against the
  per-segment oracle (radix_sort + sequential fold), bit-exact including the
  FP64 sums;
- group-by: the full 2^28-row configuration against the oracle's fold
  (dfcolumn_grouped_agg.cc:153-164), with the worst error reported both
  relative to the group's absolute mass (the bound, 1e-12) and to |sum|.
"""
import sys

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


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def _idx(n):
    import torch
    import frovedis_amd as fa
    return fa.DeviceArray(torch.arange(n, dtype=torch.int64, device="cuda"), np.uint64)


def test_join_config_generators_match_oracle(fa):
    from frovedis_amd import datagen as dg
    nl, nr = 1 << 20, 1 << 22
    lk = dg.perm_keys_np(nl)
    rk = dg.fk_keys_np(nr, domain=nl)
    li = np.arange(nl, dtype=np.uint64)
    ri = np.arange(nr, dtype=np.uint64)
    dlk, drk = fa.DeviceArray.from_numpy(lk), fa.DeviceArray.from_numpy(rk)
    dli, dri = fa.DeviceArray.from_numpy(li), fa.DeviceArray.from_numpy(ri)
    # hash join: right_df.hash_join(left_df, eq(rk, lk)); the unique 2^20
    # side is the build side, pairs in probe order
    _say("join cfg: hash leg vs oracle")
    elo, ero, eh = oracle.equi_join(rk, ri, lk, li, "ve")
    assert eh
    lo, ro, used = fa.equi_join(drk, dri, dlk, dli)
    assert used
    assert np.array_equal(lo.numpy(), elo) and np.array_equal(ro.numpy(), ero)
    assert len(elo) == int((rk < nl).sum())
    # sort-merge: left_df.hash_join(right_df, eq(lk, rk)); the build side
    # (2^22 foreign keys) repeats keys, pairs in (key, left, right) order
    _say("join cfg: sort-merge leg vs oracle")
    elo, ero, eh = oracle.equi_join(lk, li, rk, ri, "ve")
    assert not eh
    lo, ro, used = fa.equi_join(dlk, dli, drk, dri)
    assert not used
    assert np.array_equal(lo.numpy(), elo) and np.array_equal(ro.numpy(), ero)
    # the same pairs as the x86 multimap body, canonically
    xlo, xro, _ = oracle.equi_join(lk, li, rk, ri, "x86")
    a = np.lexsort((ero, elo))
    b = np.lexsort((xro, xlo))
    assert np.array_equal(elo[a], xlo[b]) and np.array_equal(ero[a], xro[b])


def test_join_config_full_size_properties(fa):
    import torch
    from frovedis_amd import datagen as dg
    nl, nr = 1 << 27, 1 << 29
    _say("join cfg full size: generating")
    lk = fa.DeviceArray(dg.perm_keys_dev(nl), np.int64)
    rk = fa.DeviceArray(dg.fk_keys_dev(nr, domain=nl), np.int64)
    li, ri = _idx(nl), _idx(nr)
    hits = int((rk.t < nl).sum())
    _say("join cfg full size: hash leg")
    lo, ro, used = fa.equi_join(rk, ri, lk, li, capacity=nr)
    assert used and len(lo) == hits
    assert torch.equal(rk.t[lo.t], lk.t[ro.t])
    assert bool((lo.t[1:] > lo.t[:-1]).all())  # probe order, each hit once
    # every hit row appears: the probe rows are exactly those with a key < nl
    assert torch.equal(lo.t, torch.nonzero(rk.t < nl).flatten())
    del lo, ro
    torch.cuda.empty_cache()
    _say("join cfg full size: sort-merge leg")
    lo, ro, used = fa.equi_join(lk, li, rk, ri, capacity=nr)
    assert not used and len(lo) == hits
    kk = lk.t[lo.t]
    assert torch.equal(kk, rk.t[ro.t])
    assert bool((kk[1:] >= kk[:-1]).all())
    # within a key (one left row, as the left keys are unique) right rows ascend
    same = kk[1:] == kk[:-1]
    assert bool((ro.t[1:][same] > ro.t[:-1][same]).all())


def test_batch_config_segment_sizes_match_oracle(fa):
    from frovedis_amd import datagen as dg
    sizes = dg.segment_sizes(nseg=512, total=1 << 22)
    assert sizes.min() >= 256 and sizes.max() <= 1 << 20 and sizes.sum() == 1 << 22
    keys_d, vals_d, off_d = dg.batch_inputs_dev(sizes)
    keys, vals, off = keys_d.cpu().numpy(), vals_d.cpu().numpy(), off_d.cpu().numpy()
    k = fa.DeviceArray(keys_d, np.int64)
    v = fa.DeviceArray(vals_d, np.float64)
    o = fa.DeviceArray(off_d, np.uint64)
    for use_graph in (False, True):
        _say(f"batch cfg: 512 segments, graph={use_graph}")
        res = fa.batch_sort_groupby(k, v, o, key_bits=18, use_graph=use_graph)
        ok, ov = res["keys"].numpy(), res["vals"].numpy()
        gseg = res["group_seg_off"].numpy()
        G = res["ngroups"]
        gk = res["group_keys"].numpy()[:G]
        gs = res["group_sum"].numpy()[:G]
        gc = res["group_count"].numpy()[:G]
        assert gseg[0] == 0 and gseg[-1] == G
        for s in range(len(sizes)):
            a, b = int(off[s]), int(off[s + 1])
            ek, ev = oracle.radix_sort(keys[a:b], vals[a:b].view(np.uint64))
            assert np.array_equal(ok[a:b], ek), s
            assert np.array_equal(ov[a:b].view(np.uint64), ev), s
            gi, split, ukeys = oracle.group_by(keys[a:b])
            g0, g1 = int(gseg[s]), int(gseg[s + 1])
            assert np.array_equal(gk[g0:g1], ukeys), s
            assert np.array_equal(gc[g0:g1], oracle.grouped_agg("count", vals[a:b], gi, split)), s
            assert np.array_equal(gs[g0:g1].view(np.uint64),
                                  oracle.grouped_agg("sum", vals[a:b], gi, split).view(np.uint64)), s


def test_groupby_config_full_size_against_oracle_fold(fa):
    import torch
    from frovedis_amd import datagen as dg
    n = 1 << 28
    _say("group-by cfg full size: generating 2^28 rows")
    kd = dg.zipf_keys_dev(n)
    vd = dg.normal_dev(n)
    ivd = dg.int_uniform_dev(n, 1000)
    k = fa.DeviceArray(kd, np.int64)
    v = fa.DeviceArray(vd, np.float64)
    iv = fa.DeviceArray(ivd, np.int32)
    _say("group-by cfg full size: GPU step")
    gk, (s, c, a, mx) = fa.group_by_agg(k, [("sum", v), ("count", v), ("avg", v), ("max", iv)])
    gk, s, c, a, mx = (x.numpy() for x in (gk, s, c, a, mx))
    kh, vh, ivh = kd.cpu().numpy(), vd.cpu().numpy(), ivd.cpu().numpy()
    del k, v, iv, kd, vd, ivd
    torch.cuda.empty_cache()
    _say("group-by cfg full size: oracle group_by (sort of 2^28 keys)")
    gi, split, keys = oracle.group_by(kh)
    assert np.array_equal(gk, keys)
    _say("group-by cfg full size: oracle aggregates")
    assert np.array_equal(c, oracle.grouped_agg("count", vh, gi, split))
    assert np.array_equal(mx, oracle.grouped_agg("max", ivh, gi, split))
    ref = oracle.grouped_agg("sum", vh, gi, split)
    ravg = oracle.grouped_agg("avg", vh, gi, split)
    g = gi.astype(np.int64)
    absv = np.abs(vh[g])
    del g
    cum = np.concatenate([[0.0], np.cumsum(absv)])
    del absv
    mass = cum[split[1:].astype(np.int64)] - cum[split[:-1].astype(np.int64)]
    err = np.abs(s - ref)
    rel_mass = float(np.max(err / np.maximum(np.abs(ref), mass)))
    nz = np.abs(ref) > 0
    rel_ref = float(np.max(err[nz] / np.abs(ref[nz])))
    top = int(np.argmax(c))
    _say(f"group-by cfg full size: {len(keys)} groups, top group {int(c[top])} rows; "
         f"max |gpu-ref|/max(|ref|,mass) = {rel_mass:.3e}, max |gpu-ref|/|ref| = {rel_ref:.3e}, "
         f"top group |gpu-ref|/|ref| = {err[top] / abs(ref[top]):.3e}")
    assert rel_mass <= 1e-12
    cnt = c.astype(np.float64)
    aerr = np.abs(a - ravg) / np.maximum(np.abs(ravg), mass / cnt)
    assert float(np.max(aerr)) <= 1e-12

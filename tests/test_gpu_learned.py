"""GPU checks of the learned partition sort and its network.

- The FP32 forward/backward kernels against the float64 oracle network
  (oracle/mlp.py, itself pinned by finite differences): relative error of
  the gradient <= 1e-3 of its norm, forward within 1e-5.
- The learned sort's output equals the plain radix sort's (bit-exact) on the
  lognormal-mixture keys, with and without pre-sorted runs, and on inputs
  the model never saw (uniform, heavy duplicates).
- What the network gains: with equal-width coarse buckets instead of the
  network the same pipeline sends far more keys to the radix fallback.
"""
import numpy as np
import pytest

import oracle
from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fa():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import frovedis_amd
    frovedis_amd.lib()
    return frovedis_amd


def test_mlp_gradient_and_forward_match_oracle(fa):
    import torch
    rng = np.random.default_rng(1)
    p = (rng.standard_normal(omlp.NPARAM) * 0.2).astype(np.float32)
    p[0:64] = np.abs(p[0:64]) + 0.2
    p[64:128] = -np.linspace(0, 1, 64, dtype=np.float32)
    for S in [1, 63, 64, 1000, 100000]:
        x = rng.random(S).astype(np.float32)
        t = np.sort(rng.random(S)).astype(np.float32)
        dp = torch.from_numpy(p).cuda()
        g, loss = fa.mlp_grad(dp, torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
        ref = omlp.grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
        err = np.linalg.norm(g.cpu().numpy().astype(np.float64) - ref)
        assert err <= 1e-3 * np.linalg.norm(ref) + 1e-7, (S, err, np.linalg.norm(ref))
        assert abs(loss - omlp.loss(p, x, t)) <= 1e-4 * omlp.loss(p, x, t) + 1e-7
        y = fa.mlp_forward(dp, torch.from_numpy(x).cuda()).cpu().numpy()
        assert np.allclose(y, omlp.forward(p, x), rtol=1e-4, atol=1e-5)


def test_adam_step_lowers_loss(fa):
    import torch
    rng = np.random.default_rng(2)
    p = np.zeros(omlp.NPARAM, np.float32)
    p[0:64] = 1.0
    p[64:128] = -np.linspace(0, 1, 64, dtype=np.float32)
    p[128:4224] = rng.standard_normal(4096).astype(np.float32) * 0.1
    p[4288:4352] = rng.standard_normal(64).astype(np.float32) * 0.1
    x = torch.from_numpy(rng.random(20000).astype(np.float32)).cuda()
    t = x ** 2
    dp = torch.from_numpy(p).cuda()
    m = torch.zeros_like(dp)
    v = torch.zeros_like(dp)
    first = None
    for step in range(1, 201):
        g, loss = fa.mlp_grad(dp, x, t)
        first = loss if first is None else first
        fa.mlp_adam(dp, g, m, v, step, 3e-3)
    assert loss < 0.05 * first


def _plain_sorted(fa, keys_t):
    import torch
    k = fa.DeviceArray(keys_t.clone(), np.int64)
    fa.radix_sort(k)
    torch.cuda.synchronize()
    return k.t


def _trained_model(fa, n, seed):
    from frovedis_amd import datagen as dg
    prev = dg.lognormal_mixture_dev(n, seed=seed)
    s = fa.DeviceArray(_plain_sorted(fa, prev), np.int64)
    return fa.CdfModel(s, steps=300)


@pytest.mark.parametrize("n", [1 << 20, 1 << 24])
def test_learned_sort_matches_plain_radix(fa, n):
    import torch
    from frovedis_amd import datagen as dg
    model = _trained_model(fa, n, dg.SEED_LOGN + 100)
    keys = dg.lognormal_mixture_dev(n)
    ref = _plain_sorted(fa, keys)
    k = fa.DeviceArray(keys.clone(), np.int64)
    fb_keys, fb_buckets, _ = model.sort(k)
    assert torch.equal(k.t, ref)
    assert fb_keys < n // 20, (fb_keys, fb_buckets)
    if n == 1 << 20:  # the oracle agrees too
        assert np.array_equal(k.numpy(), oracle.radix_sort(keys.cpu().numpy())[0])


def test_learned_sort_on_unseen_distributions(fa):
    import torch
    from frovedis_amd import datagen as dg
    n = 1 << 22
    model = _trained_model(fa, n, dg.SEED_LOGN + 200)
    cases = [dg.splitmix64_dev(5, 0, n) & ((1 << 62) - 1),       # uniform: misses the model
             torch.randint(0, 50, (n,), device="cuda", dtype=torch.int64),  # duplicates
             torch.arange(n, device="cuda", dtype=torch.int64),  # already sorted
             torch.zeros(1, device="cuda", dtype=torch.int64),
             torch.tensor([5, 3], device="cuda", dtype=torch.int64)]
    for keys in cases:
        ref = _plain_sorted(fa, keys)
        k = fa.DeviceArray(keys.clone(), np.int64)
        model.sort(k)
        assert torch.equal(k.t, ref)


def test_learned_sort_leaf_fallbacks(fa):
    # a model of uniform keys over [0, 2^40) (leaves about 2^24 wide); the
    # input adds dense clusters inside single leaves: 3000 distinct keys in a
    # 2^15-wide range (the interpolation check fails, the leaf goes to the
    # LDS radix sort) and one 20000-key cluster (larger than a leaf's LDS
    # capacity, sorted by the device radix sort)
    import torch
    from frovedis_amd import datagen as dg
    n = 1 << 22
    train = dg.splitmix64_dev(11, 0, n) & ((1 << 40) - 1)
    model = fa.CdfModel(fa.DeviceArray(_plain_sorted(fa, train), np.int64), steps=300)
    base = dg.splitmix64_dev(12, 0, n) & ((1 << 40) - 1)
    parts = [base]
    for c in range(40):
        x0 = (c * 0x9E3779B97F4A7C15 % (1 << 40)) & ~((1 << 24) - 1)
        parts.append(x0 + (1 << 23) + dg.splitmix64_dev(13, c * 4096, 3000) % (1 << 15))
    parts.append((7 << 32) + (1 << 23) + dg.splitmix64_dev(14, 0, 20000) % (1 << 16))
    keys = torch.cat(parts)
    ref = _plain_sorted(fa, keys)
    k = fa.DeviceArray(keys.clone(), np.int64)
    fb_keys, fb_buckets, _ = model.sort(k)
    assert torch.equal(k.t, ref)
    assert fb_buckets >= 41 and fb_keys >= 40 * 3000 + 20000, (fb_keys, fb_buckets)


def test_network_gain_over_equal_width_buckets(fa):
    import torch
    from frovedis_amd import datagen as dg
    n = 1 << 24
    model = _trained_model(fa, n, dg.SEED_LOGN + 300)
    keys = dg.lognormal_mixture_dev(n)
    ref = _plain_sorted(fa, keys)
    k1 = fa.DeviceArray(keys.clone(), np.int64)
    fb_model, _, coarse_model = model.sort(k1)
    k2 = fa.DeviceArray(keys.clone(), np.int64)
    fb_flat, _, coarse_flat = model.sort(k2, equal_width=True)
    assert torch.equal(k1.t, ref) and torch.equal(k2.t, ref)
    # the network spreads the keys over the coarse buckets that feed the
    # split stage (one workgroup each); equal-width buckets over the key
    # range put most keys in a handful of them
    assert fb_model < n // 20
    assert coarse_model * 20 < coarse_flat, (coarse_model, coarse_flat)
    assert coarse_flat > n // 4
    info = model.info()
    assert np.all(np.diff(info["cdf_knots"]) >= 0)
    assert np.all(np.diff(info["coarse_splitters"].astype(np.float64)) >= 0)


def test_model_cardinality_estimate(fa):
    # Zipf(1.1) keys over 2^20 values: the model's distinct-count estimate
    # for a fresh 2^22-row draw is within 2x of the truth
    import torch
    from frovedis_amd import datagen as dg
    n = 1 << 22
    prev = dg.zipf_keys_dev(n, start=n)
    s = fa.DeviceArray(_plain_sorted(fa, prev), np.int64)
    model = fa.CdfModel(s, steps=300)
    cur = dg.zipf_keys_dev(n)
    truth = len(torch.unique(cur))
    est = model.cardinality(n)
    assert truth / 2 <= est <= truth * 2, (est, truth)


def _best_ms(fn, reps=3):
    import time
    import torch
    best = float("inf")
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best * 1e3


def test_network_against_sample_quantile_splitters(fa):
    # The ablation: the same pipeline with the coarse splitters taken as the
    # training sample's quantiles b/2048 instead of the network, and with
    # equal-width ranges.  All three sort exactly.  Measured at 2^24 on the
    # lognormal mixture: network 58K / quantile 56K / equal-width 16.3M keys
    # in the largest coarse bucket, 1.32 / 1.32 / 12.4 ms -- the network
    # matches the sample quantiles (it gains nothing over them on this
    # input) and both beat equal-width buckets by far.  The figures are
    # printed (bench.py reports the times at 2^28).
    import sys
    import torch
    from frovedis_amd import datagen as dg
    n = 1 << 24
    model = _trained_model(fa, n, dg.SEED_LOGN + 300)
    keys = dg.lognormal_mixture_dev(n)
    ref = _plain_sorted(fa, keys)
    out = {}
    for coarse in ("network", "quantile", "equal_width"):
        k = fa.DeviceArray(keys.clone(), np.int64)
        fb, nb, mc = model.sort(k, coarse=coarse)
        assert torch.equal(k.t, ref), coarse

        def run():
            k.t.copy_(keys)
            model.sort(k, coarse=coarse)
        out[coarse] = (mc, fb, _best_ms(run))
    print("coarse splitters: (max coarse bucket, fallback keys, ms)", out, file=sys.stderr)
    assert out["network"][0] <= 2 * out["quantile"][0], out
    assert out["network"][0] * 20 < out["equal_width"][0], out
    assert out["quantile"][0] * 20 < out["equal_width"][0], out


def test_model_sizes_group_by_tables(fa):
    # the model of an earlier input's key column gives the group-by's table
    # hint; the results equal those without it
    import torch
    from frovedis_amd import datagen as dg
    n = 1 << 22
    prev = dg.zipf_keys_dev(n, start=n)
    model = fa.CdfModel(fa.DeviceArray(_plain_sorted(fa, prev), np.int64), steps=300)
    k = fa.DeviceArray(dg.zipf_keys_dev(n), np.int64)
    v = fa.DeviceArray(dg.normal_dev(n), np.float64)
    truth = len(torch.unique(k.t))
    hint = model.table_hint(n)
    assert truth / 2 <= hint <= truth * 2, (hint, truth)
    g0, (s0, c0) = fa.group_by_agg(k, [("sum", v), ("count", None)])
    g1, (s1, c1) = fa.group_by_agg(k, [("sum", v), ("count", None)], model=model)
    assert torch.equal(g0.t, g1.t) and torch.equal(c0.t, c1.t)
    assert torch.allclose(s0.t, s1.t, rtol=1e-12, atol=1e-9)


def test_model_picks_join_strategy(fa):
    # a unique build column (a permutation: the training column's keys are
    # all distinct) -> AUTO without the sampled duplicate test; a foreign-key
    # column with repeats -> AUTO with it; the pairs are those of AUTO
    import torch
    from frovedis_amd import datagen as dg
    nl, nr = 1 << 18, 1 << 20
    lk = fa.DeviceArray(dg.perm_keys_dev(nl), np.int64)
    rk = fa.DeviceArray(dg.fk_keys_dev(nr, domain=nl), np.int64)
    li = fa.DeviceArray(torch.arange(nl, dtype=torch.int64, device="cuda"), np.uint64)
    ri = fa.DeviceArray(torch.arange(nr, dtype=torch.int64, device="cuda"), np.uint64)
    m_unique = fa.CdfModel(fa.DeviceArray(_plain_sorted(fa, dg.perm_keys_dev(nl, seed=77)),
                                          np.int64), steps=100)
    m_fk = fa.CdfModel(fa.DeviceArray(_plain_sorted(fa, dg.fk_keys_dev(nr, domain=nl, seed=78)),
                                      np.int64), steps=100)
    assert m_unique.join_strategy(nl) == "auto_unique"
    assert m_fk.join_strategy(nr) == "auto"
    for (a, ai, b, bi, m) in [(rk, ri, lk, li, m_unique), (lk, li, rk, ri, m_fk)]:
        lo0, ro0, u0 = fa.equi_join(a, ai, b, bi)
        lo1, ro1, u1 = fa.equi_join(a, ai, b, bi, model=m)
        assert u0 == u1 and torch.equal(lo0.t, lo1.t) and torch.equal(ro0.t, ro1.t)


def _prof_count(fa, name):
    import ctypes
    t = ctypes.c_double(0)
    c = ctypes.c_size_t(0)
    fa.check(fa.lib().fhip_prof_query(name.encode(), ctypes.byref(t), ctypes.byref(c)))
    return c.value


def test_model_heavy_keys_plan_the_group_by(fa):
    # the model of an earlier Zipf input keeps its longest runs as heavy
    # keys; the group-by then fills its heavy-key table from them (no sample
    # of the input counted: no "groupby_plan" scope) with the same result,
    # and a model of another distribution changes nothing but the speed
    import torch
    from frovedis_amd import datagen as dg
    n = 1 << 22
    prev = dg.zipf_keys_dev(n, start=n)
    model = fa.CdfModel(fa.DeviceArray(_plain_sorted(fa, prev), np.int64), steps=100)
    nk, frac = model.heavy_keys()
    assert nk == 4096 and 0.6 < frac < 0.95, (nk, frac)
    other = fa.CdfModel(fa.DeviceArray(_plain_sorted(fa, dg.lognormal_mixture_dev(n)), np.int64),
                        steps=100)
    k = fa.DeviceArray(dg.zipf_keys_dev(n), np.int64)
    v = fa.DeviceArray(dg.normal_dev(n), np.float64)
    iv = fa.DeviceArray(dg.int_uniform_dev(n, 1000), np.int32)
    aggs = [("sum", v), ("count", v), ("max", iv)]
    fa.check(fa.lib().fhip_prof_enable(1))
    g0, (s0, c0, m0) = fa.group_by_agg(k, aggs)
    torch.cuda.synchronize()
    assert _prof_count(fa, "groupby_plan") == 1
    fa.check(fa.lib().fhip_prof_enable(0))
    fa.check(fa.lib().fhip_prof_enable(1))
    g1, (s1, c1, m1) = fa.group_by_agg(k, aggs, model=model)
    torch.cuda.synchronize()
    assert _prof_count(fa, "groupby_plan") == 0
    fa.check(fa.lib().fhip_prof_enable(0))
    g2, (s2, c2, m2) = fa.group_by_agg(k, aggs, model=other)
    for g, s_, c, m in ((g1, s1, c1, m1), (g2, s2, c2, m2)):
        assert torch.equal(g0.t, g.t) and torch.equal(c0.t, c.t) and torch.equal(m0.t, m.t)
        assert torch.allclose(s0.t, s_.t, rtol=1e-12, atol=1e-9)
    # against the oracle
    gi, split, ukeys = oracle.group_by(k.numpy())
    assert np.array_equal(g1.numpy(), ukeys)
    assert np.array_equal(c1.numpy(), oracle.grouped_agg("count", v.numpy(), gi, split))
    assert np.array_equal(m1.numpy(), oracle.grouped_agg("max", iv.numpy(), gi, split))

"""GPU parity of the core layer against the CPU oracle: radix_sort (every key
type, payload width, positive_only, descending, the < 64 insertion path),
prefix_sum, set_separate, set_is_unique and lower/upper_bound.  Integer and
index results are compared bit-exact."""
import ctypes
import zlib

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

KEY_TYPES = [np.int32, np.uint32, np.int64, np.uint64, np.float32, np.float64]
SIZES = [0, 1, 2, 63, 64, 65, 4095, 4096, 4097, 100003, 1 << 20]


@pytest.fixture(scope="module")
def fa():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import frovedis_amd
    frovedis_amd.lib()
    return frovedis_amd


def _gen(rng, dt, n, kind):
    dt = np.dtype(dt)
    if dt.kind == "f":
        if kind == "full":
            a = rng.standard_normal(n) * 1e6
        elif kind == "pos":
            a = rng.random(n) * 1000
        else:
            a = rng.integers(-20, 20, n).astype(np.float64) * 0.5
        return a.astype(dt)
    if kind == "full":
        b = rng.integers(0, np.iinfo(np.uint64).max, n, dtype=np.uint64, endpoint=True)
        if dt.itemsize == 4:
            b = (b >> np.uint64(32)).astype(np.uint32)
            return b.view(np.int32) if dt.kind == "i" else b
        return b.view(np.int64) if dt.kind == "i" else b
    if kind == "pos":
        return rng.integers(0, 1 << 20, n).astype(dt)
    lo = -40 if dt.kind == "i" else 0
    return rng.integers(lo, 40, n).astype(dt)


@pytest.mark.parametrize("dt", KEY_TYPES)
@pytest.mark.parametrize("kind", ["full", "pos", "dup"])
@pytest.mark.parametrize("vbytes", [0, 4, 8])
def test_radix_sort_matches_oracle(fa, dt, kind, vbytes):
    rng = np.random.default_rng(zlib.crc32(f"{np.dtype(dt).str}{kind}{vbytes}".encode()))
    for n in SIZES:
        k = _gen(rng, dt, n, kind)
        v = None
        if vbytes:
            v = np.arange(n, dtype=np.uint32 if vbytes == 4 else np.uint64)
        ek, ev = oracle.radix_sort(k, v)
        dk = fa.DeviceArray.from_numpy(k)
        dv = fa.DeviceArray.from_numpy(v) if v is not None else None
        fa.radix_sort(dk, dv)
        gk = dk.numpy()
        assert np.array_equal(gk.view(np.uint8), ek.view(np.uint8)), (dt, kind, n)
        if v is not None:
            assert np.array_equal(dv.numpy(), ev), (dt, kind, n, vbytes)


@pytest.mark.parametrize("dt", KEY_TYPES)
def test_radix_sort_desc_and_positive_only(fa, dt):
    rng = np.random.default_rng(17)
    for n in [10, 64, 5000, 200001]:
        for kind in ["full", "pos", "dup"]:
            k = _gen(rng, dt, n, kind)
            v = np.arange(n, dtype=np.uint64)
            for po in [False, True]:
                for desc in [False, True]:
                    ek, ev = oracle.radix_sort(k, v, positive_only=po, descending=desc)
                    dk = fa.DeviceArray.from_numpy(k)
                    dv = fa.DeviceArray.from_numpy(v)
                    fa.radix_sort(dk, dv, positive_only=po, descending=desc)
                    assert np.array_equal(dk.numpy().view(np.uint8), ek.view(np.uint8)), (n, kind, po, desc)
                    assert np.array_equal(dv.numpy(), ev), (n, kind, po, desc)


def test_radix_sort_float_special_values(fa):
    # -0.0 after positives when nothing is negative; NaN placed by bit pattern
    for ft in [np.float32, np.float64]:
        rng = np.random.default_rng(3)
        k = (rng.random(3000) * 10).astype(ft)
        k[5] = -0.0
        k[6] = np.nan
        k[7] = np.inf
        v = np.arange(3000, dtype=np.uint64)
        for kneg in [False, True]:
            kk = k.copy()
            if kneg:
                kk[8] = -1.5
                kk[9] = -np.inf
            ek, ev = oracle.radix_sort(kk, v)
            dk, dv = fa.DeviceArray.from_numpy(kk), fa.DeviceArray.from_numpy(v)
            fa.radix_sort(dk, dv)
            assert np.array_equal(dv.numpy(), ev)
            assert np.array_equal(dk.numpy().view(np.uint8), ek.view(np.uint8))


def test_radix_sort_full_size_properties(fa):
    # 2^24 uint64 keys over the full range + size_t index payload (the
    # reference's radix kv configuration): sorted, and the payload is the
    # stable permutation of the input
    import torch
    n = 1 << 24
    g = torch.Generator(device="cuda").manual_seed(1)
    kt = torch.randint(-(2 ** 63), 2 ** 63 - 1, (n,), device="cuda", dtype=torch.int64, generator=g)
    keys = fa.DeviceArray(kt.clone(), np.uint64)
    vals = fa.DeviceArray(torch.arange(n, device="cuda", dtype=torch.int64), np.uint64)
    fa.radix_sort(keys, vals)
    sk = keys.t
    # unsigned order on the int64 carrier: compare with the sign bit flipped
    flipped = sk ^ torch.tensor(-(2 ** 63), dtype=torch.int64, device="cuda")
    assert bool((flipped[1:] >= flipped[:-1]).all())
    perm = vals.t
    assert bool(torch.equal(kt[perm], sk))
    assert bool(torch.equal(torch.sort(perm).values, torch.arange(n, device="cuda")))
    # within equal keys the payload keeps input order (stability)
    eq = sk[1:] == sk[:-1]
    assert bool((perm[1:][eq] > perm[:-1][eq]).all())


@pytest.mark.parametrize("dt", [np.int32, np.uint32, np.int64, np.uint64, np.float64])
def test_prefix_sum(fa, dt):
    rng = np.random.default_rng(5)
    for n in [0, 1, 100, 4096, 4097, 1000003]:
        a = rng.integers(0, 1000, n).astype(dt)
        e = oracle.prefix_sum(a)
        got = fa.prefix_sum(fa.DeviceArray.from_numpy(a)).numpy()
        if np.dtype(dt).kind == "f":
            # integer-valued doubles below 2^53 add exactly in any order
            assert np.array_equal(got, e)
        else:
            assert np.array_equal(got, e)
    a = rng.standard_normal(100000)
    got = fa.prefix_sum(fa.DeviceArray.from_numpy(a)).numpy()
    e = oracle.prefix_sum(a)
    scale = np.cumsum(np.abs(a))
    assert np.all(np.abs(got - e) <= 1e-12 * np.maximum(scale, 1.0))


@pytest.mark.parametrize("dt", KEY_TYPES)
def test_set_separate_and_unique(fa, dt):
    rng = np.random.default_rng(6)
    for n in [0, 1, 2, 2047, 2048, 2049, 300001]:
        k = np.sort(_gen(rng, dt, n, "dup"))
        e = oracle.set_separate(k)
        got = fa.set_separate(fa.DeviceArray.from_numpy(k)).numpy()
        assert np.array_equal(got, e), n
        assert fa.set_is_unique(fa.DeviceArray.from_numpy(k)) == oracle.set_is_unique(k)
        u = np.unique(k)
        assert fa.set_is_unique(fa.DeviceArray.from_numpy(u)) == oracle.set_is_unique(u)


@pytest.mark.parametrize("dt", [np.int64, np.uint32, np.float64])
def test_lower_upper_bound(fa, dt):
    rng = np.random.default_rng(8)
    s = np.sort(_gen(rng, dt, 50000, "pos"))
    q = _gen(rng, dt, 70001, "pos")
    ds, dq = fa.DeviceArray.from_numpy(s), fa.DeviceArray.from_numpy(q)
    assert np.array_equal(fa.lower_bound(ds, dq).numpy(), oracle.lower_bound(s, q))
    assert np.array_equal(fa.upper_bound(ds, dq).numpy(), oracle.lower_bound(s, q, upper=True))
    empty = fa.DeviceArray.from_numpy(np.zeros(0, dt))
    assert np.array_equal(fa.lower_bound(empty, dq).numpy(), np.zeros(len(q), np.uint64))


def test_back_to_back_sorts_on_two_streams(fa):
    # workspace blocks go back to the cache while a call's kernels are still
    # queued; a call on another stream is ordered after them, so two sorts
    # issued on two streams without a host wait both stay exact
    import torch
    rng = np.random.default_rng(404)
    n = 1 << 22
    ka = rng.integers(0, 1 << 63, n, dtype=np.uint64)
    kb = rng.integers(-(1 << 40), 1 << 40, n).astype(np.int64)
    va = np.arange(n, dtype=np.uint64)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    outs = []
    for rep in range(3):
        da, dva = fa.DeviceArray.from_numpy(ka), fa.DeviceArray.from_numpy(va)
        db, dvb = fa.DeviceArray.from_numpy(kb), fa.DeviceArray.from_numpy(va)
        torch.cuda.synchronize()
        fa.radix_sort(da, dva, stream=ctypes.c_void_p(s1.cuda_stream))
        fa.radix_sort(db, dvb, stream=ctypes.c_void_p(s2.cuda_stream))
        fa.radix_sort(da, dva, descending=True, stream=ctypes.c_void_p(s2.cuda_stream))
        torch.cuda.synchronize()
        outs.append((da.numpy(), dva.numpy(), db.numpy(), dvb.numpy()))
    ea, eva = oracle.radix_sort(ka, va, descending=True)
    eb, evb = oracle.radix_sort(kb, va)
    for a, av, b, bv in outs:
        assert np.array_equal(a, ea) and np.array_equal(av, eva)
        assert np.array_equal(b, eb) and np.array_equal(bv, evb)

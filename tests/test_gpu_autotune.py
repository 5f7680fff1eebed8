"""The autotuner on the device: every candidate it measures gives the
default setting's output, and it leaves the fastest one set."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fa():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import frovedis_amd
    frovedis_amd.lib()
    return frovedis_amd


def test_tune_radix_every_shape_sorts_the_same(fa):
    import torch
    from frovedis_amd import autotune as at
    from frovedis_amd import datagen as dg
    n = (1 << 20) + 12345
    keys = fa.DeviceArray(dg.radix_keys_dev(n), np.uint64)
    vals = fa.DeviceArray(torch.arange(n, dtype=torch.int64, device="cuda"), np.uint64)
    best_s, shape, results = at.tune_radix(keys, vals, reps=1)
    assert len(results) == len(at.RADIX_SHAPES)
    assert all(ok for _, _, ok in results), results
    assert at.tune_get("radix.shape_kv") == shape
    ref = np.sort(keys.numpy(), kind="stable")
    fa.radix_sort(keys, vals)
    assert np.array_equal(keys.numpy(), ref)
    at.tune_set("radix.shape_kv", None)


def test_tune_groupby_small_space(fa, tmp_path):
    from frovedis_amd import autotune as at
    from frovedis_amd import datagen as dg
    n = 1 << 22
    k = fa.DeviceArray(dg.zipf_keys_dev(n), np.int64)
    v = fa.DeviceArray(dg.normal_dev(n), np.float64)
    space = {"groupby.span": (4, 16), "groupby.chunk": (1 << 16, 1 << 17)}
    best_s, setting, results = at.tune_groupby(k, [("sum", v), ("count", v)], space=space, reps=1)
    assert len(results) == 1 + 4
    for nm, val in setting.items():
        assert at.tune_get(nm, -1) == (-1 if val is None else val)
    p = tmp_path / "tune.json"
    at.save(p)
    for nm in space:
        at.tune_set(nm, None)
    assert at.load(p)
    for nm, val in setting.items():
        assert at.tune_get(nm, -1) == (-1 if val is None else val)
    for nm in space:
        at.tune_set(nm, None)


def test_tune_batch_per_size_class(fa):
    # the batched segments' tuner: digit width, then the bucket rows of each
    # size class; every kept setting reproduces the default's output and the
    # batch path then runs with it
    import oracle
    from frovedis_amd import autotune as at
    from frovedis_amd import datagen as dg
    sizes = dg.segment_sizes(nseg=64, total=1 << 20, lo=256, hi=1 << 16)
    keys, vals, off = dg.batch_inputs_dev(sizes)
    k = fa.DeviceArray(keys, np.int64)
    v = fa.DeviceArray(vals, np.float64)
    o = fa.DeviceArray(off, np.uint64)
    classes = sorted({at.batch_size_class(int(m)) for m in sizes if m > 4096}, reverse=True)[:2]
    best_s, setting, results = at.tune_batch(k, v, o, 16, reps=1, classes=classes)
    assert len(results) == 1 + len(at.BATCH_DIGIT_BITS) + len(classes) * len(at.BATCH_BUCKET_ROWS)
    assert all(ok for _, _, ok in results), [r for r in results if not r[2]]
    for nm, val in setting.items():
        assert at.tune_get(nm, -1) == (-1 if val is None else val)
    res = fa.batch_sort_groupby(k, v, o, key_bits=16, use_graph=True)
    kh, oh = keys.cpu().numpy(), off.cpu().numpy()
    s = int(np.argmax(sizes))
    a, b = int(oh[s]), int(oh[s + 1])
    ek, _ = oracle.radix_sort(kh[a:b])
    assert np.array_equal(res["keys"].numpy()[a:b], ek)
    for nm in setting:
        at.tune_set(nm, None)

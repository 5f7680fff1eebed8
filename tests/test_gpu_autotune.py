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

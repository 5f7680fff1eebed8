"""The no-allocation radix sort entry point (fhip_dev_radix_sort_noalloc):
through the C ABI, two calls in a row (the second smaller, on the same
scratch array) against fhip_dev_radix_sort on the same inputs, bit-exact;
sizes beyond the scratch limit are rejected with an error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCRATCH = 64 << 20


@pytest.fixture(scope="module")
def fa():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import frovedis_amd
    frovedis_amd.lib()
    return frovedis_amd


def _sort(fa, fn, keys, vals, positive_only=0, descending=0):
    k = fa.DeviceArray.from_numpy(keys)
    v = fa.DeviceArray.from_numpy(vals) if vals is not None else None
    rc = fn(k.code, k.ptr, 0 if v is None else v.dtype.itemsize, None if v is None else v.ptr,
            ctypes.c_size_t(len(keys)), positive_only, descending, fa.current_stream())
    fa.check(rc)
    return k.numpy(), None if v is None else v.numpy()


@pytest.mark.parametrize("kt,vt,pos,desc", [(np.uint64, np.uint64, 0, 0),
                                            (np.int64, np.uint64, 0, 1),
                                            (np.int32, np.uint32, 1, 0),
                                            (np.float64, None, 0, 0),
                                            (np.int64, None, 1, 0)])
def test_noalloc_matches_allocating_entry_point(fa, kt, vt, pos, desc):
    L = fa.lib()
    rng = np.random.default_rng(7)
    # a large call, then a smaller one on the scratch array it left behind
    for n in [1 << 21, 100_003, 50]:
        if np.dtype(kt).kind == "f":
            keys = rng.standard_normal(n).astype(kt)
        elif np.dtype(kt).kind == "u":
            keys = rng.integers(0, 1 << 63, n, dtype=np.uint64).astype(kt)
        else:
            keys = rng.integers(-(1 << 30), 1 << 30, n).astype(kt)
            if pos:
                keys = np.abs(keys)
        vals = None if vt is None else np.arange(n).astype(vt)
        ek, ev = _sort(fa, L.fhip_dev_radix_sort, keys, vals, pos, desc)
        gk, gv = _sort(fa, L.fhip_dev_radix_sort_noalloc, keys, vals, pos, desc)
        assert np.array_equal(gk.view(np.uint8), ek.view(np.uint8)), n
        if vt is not None:
            assert np.array_equal(gv, ev), n


def test_noalloc_rejects_inputs_beyond_the_scratch(fa):
    import torch
    n = SCRATCH // 16 + 1  # 8-byte keys + 8-byte values cannot fit
    k = fa.DeviceArray(torch.zeros(n, dtype=torch.int64, device="cuda"), np.uint64)
    v = fa.DeviceArray(torch.zeros(n, dtype=torch.int64, device="cuda"), np.uint64)
    rc = fa.lib().fhip_dev_radix_sort_noalloc(k.code, k.ptr, 8, v.ptr, ctypes.c_size_t(n), 0, 0,
                                              fa.current_stream())
    assert rc != 0
    assert b"scratch" in fa.lib().fhip_last_error()

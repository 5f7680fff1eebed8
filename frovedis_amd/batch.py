"""Batched segments: many independent (sort + group sum/count) jobs in one
HIP graph launch (fhip_dev_batch_sort_groupby)."""
import ctypes

import numpy as np

from . import check, lib
from .core import DeviceArray, _sz, current_stream

__all__ = ["batch_sort_groupby"]


def batch_sort_groupby(keys, vals, seg_off, key_bits, capacity=None, use_graph=True,
                       out=None, stream=None):
    """keys: int64 device array, vals: float64, seg_off: uint64 device array of
    nseg + 1 offsets.  Returns a dict of device arrays: keys / vals sorted per
    segment (stable), group keys, sums, counts and the per-segment group
    offsets.  `out` may pass a dict from an earlier call to reuse its buffers
    (the captured graph is replayed when shapes and buffers repeat)."""
    n = len(keys)
    nseg = len(seg_off) - 1
    dev = keys.t.device
    cap = int(capacity) if capacity is not None else max(n, 1)
    if out is None:
        out = {"keys": DeviceArray.empty(n, np.int64, dev),
               "vals": DeviceArray.empty(n, np.float64, dev),
               "group_keys": DeviceArray.empty(cap, np.int64, dev),
               "group_sum": DeviceArray.empty(cap, np.float64, dev),
               "group_count": DeviceArray.empty(cap, np.uint64, dev),
               "group_seg_off": DeviceArray.empty(nseg + 1, np.uint64, dev)}
    ng = ctypes.c_size_t(0)
    check(lib().fhip_dev_batch_sort_groupby(
        keys.ptr, vals.ptr, _sz(n), seg_off.ptr, _sz(nseg), int(key_bits), out["keys"].ptr,
        out["vals"].ptr, out["group_keys"].ptr, out["group_sum"].ptr, out["group_count"].ptr,
        out["group_seg_off"].ptr, _sz(len(out["group_keys"])), ctypes.byref(ng),
        int(bool(use_graph)), stream or current_stream()))
    out["ngroups"] = ng.value
    return out

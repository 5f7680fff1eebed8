"""Local vector algorithms (the reference's src/frovedis/core layer) on
device arrays: radix_sort, prefix_sum, set_separate / set_is_unique and
lower_bound / upper_bound.  Each function calls one C-ABI entry point."""
import ctypes

import numpy as np

from . import CODE_DTYPE, TYPE_CODE, FhipError, check, lib

__all__ = ["DeviceArray", "radix_sort", "radix_sort_desc", "prefix_sum",
           "set_separate", "set_is_unique", "lower_bound", "upper_bound",
           "current_stream"]

_TORCH_CARRIER = {4: "int32", 8: "int64"}


def _torch():
    import torch
    return torch


def current_stream():
    """The torch current stream as the void* the C ABI takes."""
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class DeviceArray:
    """A 1-D device array with a logical element type (numpy dtype).

    The storage is a torch tensor of the same element width (uint32/uint64
    travel as int32/int64 bit patterns, since torch's unsigned support is
    partial); the logical type picks the C-ABI type code."""

    def __init__(self, tensor, dtype):
        self.t = tensor
        self.dtype = np.dtype(dtype)
        if tensor.element_size() != self.dtype.itemsize:
            raise ValueError("carrier width does not match dtype")

    @classmethod
    def from_numpy(cls, a, device="cuda"):
        torch = _torch()
        a = np.ascontiguousarray(a)
        dt = a.dtype
        carrier = a.view(getattr(np, _TORCH_CARRIER[dt.itemsize])) if dt.kind == "u" else a
        return cls(torch.from_numpy(carrier.copy()).to(device), dt)

    @classmethod
    def empty(cls, n, dtype, device="cuda"):
        torch = _torch()
        dt = np.dtype(dtype)
        tdt = {("i", 4): torch.int32, ("u", 4): torch.int32, ("i", 8): torch.int64,
               ("u", 8): torch.int64, ("f", 4): torch.float32,
               ("f", 8): torch.float64}[(dt.kind, dt.itemsize)]
        return cls(torch.empty(int(n), dtype=tdt, device=device), dt)

    @classmethod
    def empty_many(cls, n, dtypes, device="cuda"):
        """Arrays of n elements of each dtype, carved from one device
        allocation (256-byte aligned parts): one allocator call for the
        several outputs of one operation."""
        torch = _torch()
        dts = [np.dtype(d) for d in dtypes]
        offs, total = [], 0
        for dt in dts:
            offs.append(total)
            total += (int(n) * dt.itemsize + 255) // 256 * 256
        buf = torch.empty(max(total, 1), dtype=torch.uint8, device=device)
        out = []
        for dt, o in zip(dts, offs):
            tdt = {("i", 4): torch.int32, ("u", 4): torch.int32, ("i", 8): torch.int64,
                   ("u", 8): torch.int64, ("f", 4): torch.float32,
                   ("f", 8): torch.float64}[(dt.kind, dt.itemsize)]
            out.append(cls(buf[o:o + int(n) * dt.itemsize].view(tdt), dt))
        return out

    @property
    def code(self):
        return TYPE_CODE[self.dtype]

    @property
    def ptr(self):
        return ctypes.c_void_p(self.t.data_ptr()) if self.t.numel() else ctypes.c_void_p(0)

    def __len__(self):
        return int(self.t.numel())

    def numpy(self):
        a = self.t.detach().cpu().numpy()
        return a.view(self.dtype) if self.dtype.kind == "u" else a

    def __getitem__(self, sl):
        return DeviceArray(self.t[sl], self.dtype)


def _sz(n):
    return ctypes.c_size_t(int(n))


def radix_sort(keys, vals=None, positive_only=False, descending=False, stream=None):
    """In-place stable sort, radix_sort(K*, V*, size_t, bool positive_only)
    (core/radix_sort.hpp:197-219); key-only when vals is None."""
    if vals is not None and len(vals) != len(keys):
        raise FhipError("radix_sort: different size of arrays")
    vb = 0 if vals is None else vals.dtype.itemsize
    check(lib().fhip_dev_radix_sort(
        keys.code, keys.ptr, vb, vals.ptr if vals is not None else None,
        _sz(len(keys)), int(bool(positive_only)), int(bool(descending)),
        stream or current_stream()))
    return keys, vals


def radix_sort_desc(keys, vals=None, positive_only=False, stream=None):
    """radix_sort_desc (core/radix_sort.hpp:342-487): stable descending."""
    return radix_sort(keys, vals, positive_only, True, stream)


def prefix_sum(a, stream=None):
    """Inclusive scan (core/prefix_sum.hpp:67-88)."""
    out = DeviceArray.empty(len(a), a.dtype, a.t.device)
    check(lib().fhip_dev_prefix_sum(a.code, a.ptr, out.ptr, _sz(len(a)),
                                    stream or current_stream()))
    return out


def set_separate(sorted_keys, stream=None):
    """Run boundaries [0, i1, ..., n] of a sorted array; {0} when empty
    (core/set_operations.hpp:146-163)."""
    n = len(sorted_keys)
    out = DeviceArray.empty(n + 1, np.uint64, sorted_keys.t.device)
    got = ctypes.c_size_t(0)
    check(lib().fhip_dev_set_separate(sorted_keys.code, sorted_keys.ptr, _sz(n), out.ptr,
                                      ctypes.byref(got), stream or current_stream()))
    return out[: got.value]


def set_is_unique(sorted_keys, stream=None):
    r = ctypes.c_int(1)
    check(lib().fhip_dev_set_is_unique(sorted_keys.code, sorted_keys.ptr,
                                       _sz(len(sorted_keys)), ctypes.byref(r),
                                       stream or current_stream()))
    return bool(r.value)


def lower_bound(sorted_keys, queries, upper=False, stream=None):
    """Batched lower_bound (core/lower_bound.hpp:13-101)."""
    if queries.dtype != sorted_keys.dtype:
        raise FhipError("lower_bound: key types differ")
    out = DeviceArray.empty(len(queries), np.uint64, queries.t.device)
    check(lib().fhip_dev_lower_bound(sorted_keys.code, sorted_keys.ptr, _sz(len(sorted_keys)),
                                     queries.ptr, _sz(len(queries)), out.ptr, int(bool(upper)),
                                     stream or current_stream()))
    return out


def upper_bound(sorted_keys, queries, stream=None):
    return lower_bound(sorted_keys, queries, True, stream)


assert set(CODE_DTYPE) == set(range(6))

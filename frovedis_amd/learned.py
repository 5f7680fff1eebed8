"""Learned CDF model and learned partition sort (device arrays).

CdfModel trains the FP32 MLP on the GPU (fhip_learned_train) from (key,
rank) samples of an already sorted key column -- normally the output of the
plain radix sort on an earlier input -- and keeps the distilled splitters on
the host side of the library.  CdfModel.sort(keys) sorts a 64-bit integer
key column in place with the learned partition sort; the result equals
radix_sort(K*, size_t) (core/radix_sort.hpp:221-242).
"""
import ctypes

import numpy as np

from . import FhipError, check, lib
from .core import DeviceArray, _sz, current_stream

__all__ = ["CdfModel", "mlp_nparam", "mlp_grad", "mlp_forward", "mlp_adam"]

KNOTS = 4096
COARSE = 1024
COARSE_MODE = {"network": 0, "equal_width": 1, "quantile": 2}


class CdfModel:
    def __init__(self, sorted_keys, samples=None, steps=300, lr=3e-3, stream=None):
        if sorted_keys.dtype not in (np.dtype(np.int64), np.dtype(np.uint64)):
            raise FhipError("learned sort: 64-bit integer keys only")
        n = len(sorted_keys)
        samples = max(2, n // 100) if samples is None else int(samples)
        self.dtype = sorted_keys.dtype
        self._m = ctypes.c_void_p(0)
        check(lib().fhip_learned_train(sorted_keys.code, sorted_keys.ptr, _sz(n), _sz(samples),
                                       int(steps), ctypes.c_float(lr), ctypes.byref(self._m),
                                       stream or current_stream()))

    def info(self):
        loss = ctypes.c_double(0)
        knots = np.zeros(KNOTS, np.float64)
        coarse = np.zeros(COARSE, np.uint64)
        check(lib().fhip_learned_model_info(self._m, ctypes.byref(loss),
                                            knots.ctypes.data_as(ctypes.c_void_p),
                                            coarse.ctypes.data_as(ctypes.c_void_p)))
        return {"loss": loss.value, "cdf_knots": knots, "coarse_splitters": coarse}

    def params(self):
        p = np.zeros(mlp_nparam(), np.float32)
        check(lib().fhip_learned_model_params(self._m, p.ctypes.data_as(ctypes.c_void_p)))
        return p

    def sort(self, keys, equal_width=False, coarse=None, stream=None):
        """Sorts keys in place; returns (fallback_keys, fallback_buckets,
        max_coarse_bucket).  coarse: "network" (default), or the same
        pipeline without the network: "equal_width" (also equal_width=True)
        or "quantile" (the training sample's quantiles)."""
        if keys.dtype != self.dtype:
            raise FhipError("learned sort: key type differs from the model's")
        coarse = coarse or ("equal_width" if equal_width else "network")
        mode = COARSE_MODE[coarse]
        fk, fb, mc = ctypes.c_size_t(0), ctypes.c_size_t(0), ctypes.c_size_t(0)
        check(lib().fhip_dev_learned_sort(self._m, keys.ptr, _sz(len(keys)), mode,
                                          ctypes.byref(fk), ctypes.byref(fb), ctypes.byref(mc),
                                          stream or current_stream()))
        return fk.value, fb.value, mc.value

    def cardinality(self, n):
        d = ctypes.c_double(0)
        check(lib().fhip_learned_cardinality(self._m, _sz(n), ctypes.byref(d)))
        return d.value

    def table_hint(self, n):
        """The group-by table hint for n rows of this column."""
        h = ctypes.c_size_t(0)
        check(lib().fhip_learned_table_hint(self._m, _sz(n), ctypes.byref(h)))
        return h.value

    def heavy_keys(self):
        """(number of heavy keys kept, fraction of the training rows they
        hold): the longest runs of the sorted training column."""
        k = ctypes.c_size_t(0)
        f = ctypes.c_double(0)
        check(lib().fhip_learned_heavy_keys(self._m, ctypes.byref(k), ctypes.byref(f)))
        return k.value, f.value

    def join_strategy(self, n_build):
        """"auto_unique" when the model predicts n_build distinct keys on the
        build side, else "auto" (the join's result is the same)."""
        s = ctypes.c_int(0)
        check(lib().fhip_learned_join_strategy(self._m, _sz(n_build), ctypes.byref(s)))
        return {0: "auto", 3: "auto_unique"}[s.value]

    def __del__(self):
        try:
            if self._m:
                lib().fhip_learned_model_destroy(self._m)
                self._m = ctypes.c_void_p(0)
        except Exception:
            pass


def mlp_nparam():
    return int(lib().fhip_mlp_nparam())


def mlp_grad(params, x, t, stream=None):
    """Gradient of the mean squared error of the MLP on (x, t); device
    float32 tensors in, (grad tensor, loss) out."""
    import torch
    g = torch.zeros(mlp_nparam(), dtype=torch.float32, device=params.device)
    loss = torch.zeros(1, dtype=torch.float32, device=params.device)
    check(lib().fhip_dev_mlp_grad(ctypes.c_void_p(params.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                  ctypes.c_void_p(t.data_ptr()), _sz(x.numel()),
                                  ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(loss.data_ptr()),
                                  stream or current_stream()))
    return g, float(loss.item())


def mlp_forward(params, x, stream=None):
    import torch
    y = torch.empty_like(x)
    check(lib().fhip_dev_mlp_forward(ctypes.c_void_p(params.data_ptr()),
                                     ctypes.c_void_p(x.data_ptr()), _sz(x.numel()),
                                     ctypes.c_void_p(y.data_ptr()), stream or current_stream()))
    return y


def mlp_adam(params, grad, m, v, step, lr, stream=None):
    check(lib().fhip_dev_mlp_adam(ctypes.c_void_p(params.data_ptr()),
                                  ctypes.c_void_p(grad.data_ptr()), ctypes.c_void_p(m.data_ptr()),
                                  ctypes.c_void_p(v.data_ptr()), int(step), ctypes.c_float(lr),
                                  stream or current_stream()))

"""Autotuner for the library's tunable parameters.

The tunables live in the library (fhip_tune_set / fhip_tune_get) and are
saved and loaded as one flat JSON object (fhip_autotune_save / _load):

- ``groupby.slots``: slots of the LDS hash table of the fused group-by
  (power of two);
- ``groupby.fill_pct``: expected groups per partition, as a percentage of
  the slots (the hash load factor);
- ``groupby.chunk``: rows aggregated by one workgroup;
- ``groupby.span``: 4096-row tiles partitioned by one workgroup;
- ``groupby.hot``: slots of the heavy-key table (0 turns the heavy-key pass
  off); ``groupby.hot_min_pct``: the sample share the heavy keys must cover;
  ``groupby.hot_sample``: log2 of the rows sampled to find them (12 to 20);
- ``radix.shape_k`` / ``radix.shape_kv``: the radix sort pass's tile shape
  (threads x keys per thread, an index into RADIX_SHAPES) for key-only and
  key/value sorts;
- ``compact.shape``: the tile of the one-pass stream compaction (filter,
  set_separate, join outputs), an index into COMPACT_SHAPES;
- ``multi_group_by.composite``: 0 switches off the composite-key path of
  the multi-column group_by;
- ``join.load_pct``: load factor (10 to 50 %) of the joins' hash table for
  integer keys;
- ``batch.digit_bits``: digit width (2 to 8 bits) of the LDS sort passes of
  the batched segments' jobs; ``batch.bucket_rows``: rows per bucket job of a
  large segment, and ``batch.bucket_rows.c<k>`` the same for the size class
  of segments with 2^(k-1) < rows <= 2^k (these decide how many bucket jobs,
  the batch path's tiles, a segment is cut into); ``batch.path``: 1 selects
  the composite-key sort instead of the work queue.

``tune_groupby``, ``tune_radix`` and ``tune_batch`` measure candidate settings on the device for a given key
column and aggregates, keeps the fastest whose output equals the default
setting's, and sets it.  ``load`` / ``save`` keep the result between runs,
for example in ``~/.cache/frovedis_amd/tune.json``.
"""
import ctypes
import itertools
import os
import time

from . import check, lib

__all__ = ["tune_set", "tune_get", "save", "load", "tune_groupby", "tune_radix", "tune_batch",
           "DEFAULT_CACHE", "RADIX_SHAPES", "COMPACT_SHAPES", "BATCH_DIGIT_BITS",
           "BATCH_BUCKET_ROWS", "batch_size_class"]

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "frovedis_amd", "tune.json")

GROUPBY_SPACE = {
    "groupby.slots": (2048, 4096),
    "groupby.fill_pct": (40, 50, 70),
    "groupby.chunk": (1 << 16, 1 << 17, 1 << 18),
    "groupby.span": (4, 16),
    "groupby.hot": (0, 4096),
}

RADIX_SHAPES = ((256, 16), (256, 32), (512, 16), (512, 24), (512, 12), (1024, 8), (1024, 16),
                (768, 16))

COMPACT_SHAPES = ((256, 16), (1024, 16))

BATCH_DIGIT_BITS = (4, 6, 8)
BATCH_BUCKET_ROWS = (4096, 6144, 7168)


def batch_size_class(rows):
    """The size class k of a segment of `rows` rows: 2^(k-1) < rows <= 2^k
    (the suffix of the per-class tunable ``batch.bucket_rows.c<k>``)."""
    k = 0
    while (1 << k) < rows:
        k += 1
    return k


def tune_set(name, value):
    """Sets a tunable; None (or a negative value) restores the default."""
    v = -1 if value is None else int(value)
    check(lib().fhip_tune_set(name.encode(), ctypes.c_longlong(v)))


def tune_get(name, default=-1):
    out = ctypes.c_longlong(0)
    check(lib().fhip_tune_get(name.encode(), ctypes.c_longlong(int(default)), ctypes.byref(out)))
    return out.value


def save(path=DEFAULT_CACHE):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    check(lib().fhip_autotune_save(os.fspath(path).encode()))


def load(path=DEFAULT_CACHE):
    """Loads a saved setting; returns False when the file does not exist."""
    if not os.path.exists(path):
        return False
    check(lib().fhip_autotune_load(os.fspath(path).encode()))
    return True


def _timed(fn, reps):
    import torch
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps, out


def tune_groupby(keys, aggs, space=None, reps=3, table_hint=None):
    """Times group_by_agg(keys, aggs) for every setting in `space` (a dict
    name -> candidate values; GROUPBY_SPACE by default), sets the fastest
    setting whose group keys equal the default's, and returns
    (best_seconds, best_setting, all_results)."""
    import torch

    from .dataframe import group_by_agg
    space = dict(GROUPBY_SPACE if space is None else space)
    names = list(space)
    for nm in names:
        tune_set(nm, None)

    def run():
        if table_hint is None:
            return group_by_agg(keys, aggs)
        return group_by_agg(keys, aggs, table_hint=table_hint)

    base_s, (ref_keys, _) = _timed(run, reps)
    results = [(base_s, {nm: None for nm in names})]
    best = results[0]
    for combo in itertools.product(*(space[nm] for nm in names)):
        setting = dict(zip(names, combo))
        for nm, v in setting.items():
            tune_set(nm, v)
        s, (gk, _) = _timed(run, reps)
        ok = bool(torch.equal(gk.t, ref_keys.t))
        results.append((s, setting))
        if ok and s < best[0]:
            best = (s, setting)
    for nm, v in best[1].items():
        tune_set(nm, v)
    return best[0], best[1], results


def tune_radix(keys, vals=None, reps=3):
    """Times radix_sort(keys[, vals]) on copies of the inputs for every tile
    shape, checks that every shape gives the default's output, sets the
    fastest for this kind of sort (key-only or key/value) and returns
    (best_seconds, best_shape_index, all_results)."""
    import torch

    from .core import radix_sort
    name = "radix.shape_kv" if vals is not None else "radix.shape_k"
    tune_set(name, None)
    k0 = keys.t.clone()
    v0 = vals.t.clone() if vals is not None else None

    def run():
        keys.t.copy_(k0)
        if vals is not None:
            vals.t.copy_(v0)
            radix_sort(keys, vals)
        else:
            radix_sort(keys)

    run()
    ref_k = keys.t.clone()
    ref_v = vals.t.clone() if vals is not None else None
    results = []
    for shape in range(len(RADIX_SHAPES)):
        tune_set(name, shape)
        best = float("inf")
        for _ in range(reps + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            run()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        ok = bool(torch.equal(keys.t, ref_k)) and (vals is None or bool(torch.equal(vals.t, ref_v)))
        results.append((best, shape, ok))
    good = [r for r in results if r[2]]
    best = min(good) if good else (float("inf"), None, False)
    tune_set(name, best[1])
    keys.t.copy_(k0)
    if vals is not None:
        vals.t.copy_(v0)
    return best[0], best[1], results


def tune_batch(keys, vals, seg_off, key_bits, reps=2, classes=None):
    """Autotunes the batched segments (fhip_dev_batch_sort_groupby) on the
    given batch: first the LDS sort's digit width (BATCH_DIGIT_BITS), then,
    one size class at a time (the classes present among the segments with
    more rows than one job holds, largest rows first, or `classes`), the
    bucket rows of that class (BATCH_BUCKET_ROWS).  A candidate is kept only
    if it is faster and its output equals the default setting's (sorted
    rows, group keys, sums and counts).  Returns (best_seconds, setting,
    all_results)."""
    import numpy as np
    import torch

    from .batch import batch_sort_groupby
    sizes = np.diff(seg_off.numpy().astype(np.int64))
    if classes is None:
        classes = sorted({batch_size_class(int(m)) for m in sizes if m > 4096}, reverse=True)
    names = ["batch.digit_bits"] + [f"batch.bucket_rows.c{c}" for c in classes]
    for nm in names:
        tune_set(nm, None)
    out = {}

    def run():
        return batch_sort_groupby(keys, vals, seg_off, key_bits, use_graph=False, out=out.get("o"))

    ref = run()
    out["o"] = None
    ref = {k: (v.t.clone() if hasattr(v, "t") else v) for k, v in ref.items()}

    def same(r):
        G = ref["ngroups"]
        if r["ngroups"] != G:
            return False
        for k in ("keys", "vals"):
            if not torch.equal(r[k].t, ref[k]):
                return False
        return all(torch.equal(r[k].t[:G], ref[k][:G]) for k in ("group_keys", "group_sum",
                                                                   "group_count"))

    def measure():
        best = float("inf")
        r = None
        for _ in range(reps + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = run()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best, same(r)

    best_s, _ = measure()
    setting = {nm: None for nm in names}
    results = [(best_s, dict(setting), True)]
    for nm, cands in [("batch.digit_bits", BATCH_DIGIT_BITS)] + \
            [(f"batch.bucket_rows.c{c}", BATCH_BUCKET_ROWS) for c in classes]:
        for v in cands:
            tune_set(nm, v)
            t, ok = measure()
            cand = dict(setting, **{nm: v})
            results.append((t, cand, ok))
            if ok and t < best_s:
                best_s, setting = t, cand
        tune_set(nm, setting[nm])
    return best_s, setting, results

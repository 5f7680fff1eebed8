"""Benchmark: dftable group_by + sum/count/mean/max on one MI355X.

Workload (one "step"): 2^28 rows already resident in HBM -- int64 key
Zipf(s=1.1) over 2^20 values, double value N(0,1), int32 column U[0,1000) --
grouped by key with sum/count/avg of the double column and max of the int32
column, groups ascending by key (dftable.cc:346-366,3724-3772 on one rank).
Prints one JSON line.  Side measurements (the plain radix sort key/value
configuration) go under "extra".

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--rows LOG2] [--no-cpu]
                  [--no-extra] [--leg groupby|radix|batch|join|learned|copy|relational]
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
GROUPBY_BYTES_PER_ROW = 20  # read key 8 + value 8 + int32 4, once


def _prof_ms(fa, name):
    t = ctypes.c_double(0)
    c = ctypes.c_size_t(0)
    fa.check(fa.lib().fhip_prof_query(name.encode(), ctypes.byref(t), ctypes.byref(c)))
    return t.value, c.value


def _traffic_from_profiles(name):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", f"{name}_pmc.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def gen_groupby(n, device="cuda"):
    import numpy as np

    import frovedis_amd as fa
    from frovedis_amd import datagen as dg
    k = fa.DeviceArray(dg.zipf_keys_dev(n, device=device), np.int64)
    v = fa.DeviceArray(dg.normal_dev(n, device=device), np.float64)
    iv = fa.DeviceArray(dg.int_uniform_dev(n, 1000, device=device), np.int32)
    return k, v, iv


def groupby_model(fa, n, sample_log2=24):
    """A CdfModel of the key column trained on an earlier input of the same
    generator (another seed, 2^sample_log2 rows, sorted): its cardinality
    estimate for n rows sizes the group-by's tables.  Returns the model and
    the seconds spent (outside the timed region, as a model is trained once
    per column, not per query)."""
    import numpy as np
    import torch

    from frovedis_amd import datagen as dg
    t0 = time.perf_counter()
    prev = fa.DeviceArray(dg.zipf_keys_dev(1 << sample_log2, seed=dg.SEED_ZIPF + 1), np.int64)
    fa.radix_sort(prev)
    model = fa.CdfModel(prev, steps=300)
    torch.cuda.synchronize()
    return model, time.perf_counter() - t0


def step_groupby(fa, k, v, iv, model):
    return fa.group_by_agg(k, [("sum", v), ("count", v), ("avg", v), ("max", iv)],
                           model=model)


def _host_info():
    """CPU model, logical CPUs usable by this process and RAM of the host."""
    model, sockets, per_core = "unknown", set(), None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name") and model == "unknown":
                model = line.split(":", 1)[1].strip()
            elif line.startswith("physical id"):
                sockets.add(line.split(":", 1)[1].strip())
    except OSError:
        pass
    ram = None
    try:
        for line in open("/proc/meminfo"):
            if line.startswith("MemTotal"):
                ram = round(int(line.split()[1]) / 2**20, 1)
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"cpu_model": model, "sockets": len(sockets) or None, "logical_cpus": os.cpu_count(),
            "usable_cpus": usable, "ram_gib": ram}


def _cpu_reps(states, prep, run, reps=5):
    """One warm-up and `reps` timed repetitions of run(state) for every state
    at once (one thread per state: the oracle's C calls release the GIL, so
    the copies run concurrently, one per core, like one single-threaded rank
    per core).  prep(state) (fresh inputs) runs before each repetition,
    untimed.  Returns (best, median) wall seconds of a repetition."""
    import statistics
    import threading
    times = []
    for rep in range(1 + reps):
        for st in states:
            prep(st)
        go = threading.Barrier(len(states) + 1)
        errs = []

        def work(st):
            go.wait()
            try:
                run(st)
            except Exception as e:  # reported after the join
                errs.append(e)
        th = [threading.Thread(target=work, args=(st,)) for st in states]
        for t in th:
            t.start()
        go.wait()
        t0 = time.perf_counter()
        for t in th:
            t.join()
        dt = time.perf_counter() - t0
        if errs:
            raise errs[0]
        if rep:
            times.append(dt)
    return min(times), statistics.median(times)


def cpu_baselines(ncores=None, groupby_log2=24, join_log2=(20, 22), reps=5):
    """The CPU baseline of record, timed on this host: the oracle's C
    restatement of the reference's x86 bodies (liboracle.so, built with the
    reference's -O3 -pthread), one warm-up then `reps` repetitions, best and
    median, on 1 core (one rank, latency) and on N cores (N copies at once,
    aggregate throughput), N = the CPUs this process may use, at most 16.
      radix:   radix_sort(uint64*, size_t*, 2^24) (radix_sort.hpp:197-219),
               the full 2^24 configuration;
      groupby: group_by + sum/count/avg/max (dfcolumn_groupby.cc:16-105,
               dfcolumn_grouped_agg.cc), the first 2^groupby_log2 rows of the
               2^28-row generators (a bounded sample);
      join:    the x86 equi_join body (std::unordered_multimap, join.hpp:25-45)
               on 2^20 unique left keys x 2^22 foreign keys, the join
               generators at 1/128 scale."""
    import numpy as np

    import oracle
    from frovedis_amd import datagen as dg
    info = _host_info()
    N = ncores or max(1, min(16, info["usable_cpus"] or 1))
    L = oracle.lib()
    out = {"host": info, "cores_n": N, "code": "oracle restatement (liboracle.so, -O3 -pthread)",
           "timing": f"1 warm-up + {reps} repetitions, best and median"}

    def report(name, unit_items, sample, states1, statesN, prep, run):
        print(f"# cpu baseline: {name}, 1 core", file=sys.stderr, flush=True)
        b1, m1 = _cpu_reps(states1, prep, run, reps)
        print(f"# cpu baseline: {name}, {N} cores", file=sys.stderr, flush=True)
        bn, mn = _cpu_reps(statesN, prep, run, reps)
        out[name] = {"sample": sample,
                     "one_core": {"best_per_s": unit_items / b1, "median_per_s": unit_items / m1,
                                  "best_s": b1, "median_s": m1},
                     "n_cores": {"cores": N, "best_per_s": N * unit_items / bn,
                                 "median_per_s": N * unit_items / mn, "best_s": bn, "median_s": mn}}

    # radix: full-range uint64 keys + size_t index, sorted in place
    n = 1 << 24
    k0 = dg.splitmix64_np(dg.SEED_RADIX, 0, n)  # the keys of radix_keys_dev
    v0 = np.arange(n, dtype=np.uint64)

    def mk_radix(_):
        return {"k": np.empty_like(k0), "v": np.empty_like(v0)}

    def prep_radix(st):
        np.copyto(st["k"], k0)
        np.copyto(st["v"], v0)

    def run_radix(st):
        rc = L.orc_radix_sort(3, oracle._p(st["k"]), 8, oracle._p(st["v"]), n, 0, 0)
        assert rc == 0
    report("radix_kv_u64_2^24", n, "2^24 full-range uint64 keys + size_t index (the whole configuration)",
           [mk_radix(0)], [mk_radix(i) for i in range(N)], prep_radix, run_radix)
    del k0, v0

    # group-by: the first 2^groupby_log2 rows of the Zipf / normal / int generators
    g = 1 << groupby_log2
    gk = dg.zipf_keys_np(g)
    gv = dg.normal_np(g)
    gi = dg.int_uniform_np(g, 1000)

    def run_groupby(_):
        idx, split, _keys = oracle.group_by(gk)
        oracle.grouped_agg("sum", gv, idx, split)
        oracle.grouped_agg("count", gv, idx, split)
        oracle.grouped_agg("avg", gv, idx, split)
        oracle.grouped_agg("max", gi, idx, split)
    report("groupby_2^28_sample", g,
           f"first 2^{groupby_log2} of the 2^28 rows (Zipf(1.1) int64 key, N(0,1) double, U[0,1000) int32)",
           [0], list(range(N)), lambda st: None, run_groupby)
    del gk, gv, gi

    # join: x86 unordered_multimap body, unique left keys as the build side
    nl, nr = 1 << join_log2[0], 1 << join_log2[1]
    lk = dg.perm_keys_np(nl)
    rk = dg.fk_keys_np(nr, domain=nl)
    li = np.arange(nl, dtype=np.uint64)
    ri = np.arange(nr, dtype=np.uint64)

    def run_join(_):
        a, b, _used = oracle.equi_join(lk, li, rk, ri, "x86")
        assert len(a) == len(b)
    report("join_x86_2^20_x_2^22", nl + nr,
           "2^20 permutation keys x 2^22 foreign keys (90% hits), the join generators at 1/128 scale; "
           "unit: input rows (nl + nr)", [0], list(range(N)), lambda st: None, run_join)
    return out


def cpu_baseline_groupby(base):
    """The headline's cpu_baseline object: the group-by figure on N cores."""
    gb = base["groupby_2^28_sample"]
    return {"value": gb["n_cores"]["best_per_s"], "unit": "rows/s", "cores": base["cores_n"],
            "kind": "port", "sample": gb["sample"] + f"; {base['cores_n']} concurrent copies, "
            f"{base['timing']} (median {gb['n_cores']['median_per_s']:.4g} rows/s); one core: "
            f"best {gb['one_core']['best_per_s']:.4g}, median {gb['one_core']['median_per_s']:.4g} rows/s",
            "host": base["host"]}


def bench_radix(fa, log2n, steps=3):
    """Plain key/value radix sort: uint64 keys over the full 64-bit range,
    size_t index payload (the reference's radix_sort(K*, V*, n) case)."""
    import numpy as np
    import torch

    from frovedis_amd import datagen as dg
    n = 1 << log2n
    k0 = dg.radix_keys_dev(n)
    v0 = torch.arange(n, dtype=torch.int64, device="cuda")
    keys = fa.DeviceArray(k0.clone(), np.uint64)
    vals = fa.DeviceArray(v0.clone(), np.uint64)
    # the tile shape the autotuner picks for this sort on this machine
    # (radix.shape_kv; every shape's output is checked against the default's)
    from frovedis_amd import autotune as at
    _, shape, tried = at.tune_radix(keys, vals, reps=2)
    default_ms = next((r[0] * 1e3 for r in tried if r[1] == at.RADIX_SHAPES.index((1024, 16))), None)
    times = []
    fa.check(fa.lib().fhip_prof_enable(1))
    for _ in range(steps):
        keys.t.copy_(k0)
        vals.t.copy_(v0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fa.radix_sort(keys, vals)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    pass_ms, passes = _prof_ms(fa, "radix_pass")
    hist_ms, _ = _prof_ms(fa, "radix_hist")
    fa.check(fa.lib().fhip_prof_enable(0))
    t = min(times)
    P = passes // steps if steps else 0
    algo = n * (8 + 32 * P)
    kern_s = (pass_ms + hist_ms) / 1e3 / steps
    return {"keys": n, "keys_per_s": n / t, "ms": t * 1e3, "passes": P, "digit_bits": 8,
            "tile_shape": list(at.RADIX_SHAPES[shape]) if shape is not None else None,
            "default_shape_ms": default_ms,
            "algorithmic_bytes": algo, "kernel_ms": kern_s * 1e3,
            "hist_ms": hist_ms / steps, "pass_ms": pass_ms / max(passes, 1),
            "hbm_frac": algo / kern_s / HBM_PEAK if kern_s else None}


def bench_radix_host(fa, log2n=24, steps=3):
    """The host-vector drop-in: radix_sort(std::vector<uint64_t>&,
    std::vector<size_t>&) (radix_sort.hpp:489-496) through fhip_radix_sort on
    host memory -- copies in and out through pinned staging included, so the
    rate is PCIe-inclusive (not the device-resident value of the other legs).
    Same keys as the key/value leg; the output is checked sorted and a
    permutation of the input rows."""
    import ctypes

    import numpy as np

    from frovedis_amd import datagen as dg
    n = 1 << log2n
    k0 = dg.radix_keys_dev(n).cpu().numpy().view(np.uint64)
    keys = np.empty_like(k0)
    vals = np.empty(n, dtype=np.uint64)
    best = float("inf")
    L = fa.lib()
    for _ in range(steps + 1):
        keys[:] = k0
        vals[:] = np.arange(n, dtype=np.uint64)
        t0 = time.perf_counter()
        fa.check(L.fhip_radix_sort(fa.TYPE_CODE[np.dtype(np.uint64)], keys.ctypes.data_as(ctypes.c_void_p),
                                   8, vals.ctypes.data_as(ctypes.c_void_p), n, 0, 0))
        best = min(best, time.perf_counter() - t0)
    ok = bool(np.all(keys[1:] >= keys[:-1])) and bool(np.array_equal(k0[vals], keys))
    return {"keys": n, "ms": best * 1e3, "keys_per_s": n / best, "checked": ok,
            "note": "host pointers: pinned-staged PCIe copies in and out included"}


def bench_learned(fa, log2n, steps=3):
    """Learned-CDF partition sort vs the plain radix sort on the same keys:
    2^log2n int64 keys from the 8-component lognormal mixture with 5%
    pre-sorted runs; the model is trained on a 1% sample of the plain sort
    of an earlier input (another seed).  Outputs are compared bit-exact."""
    import numpy as np
    import torch

    from frovedis_amd import datagen as dg
    n = 1 << log2n
    prev = fa.DeviceArray(dg.lognormal_mixture_dev(n, seed=dg.SEED_LOGN + 1), np.int64)
    fa.radix_sort(prev)
    t0 = time.perf_counter()
    model = fa.CdfModel(prev, steps=300)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    del prev
    k0 = dg.lognormal_mixture_dev(n)
    keys = fa.DeviceArray(k0.clone(), np.int64)

    def timed(fn):
        best = float("inf")
        for _ in range(steps):
            keys.t.copy_(k0)
            torch.cuda.synchronize()
            t = time.perf_counter()
            r = fn()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t)
        return best, r

    # the plain sort with the tile shape the autotuner picks on these keys
    # (radix.shape_k; outside the timed runs)
    from frovedis_amd import autotune as at
    keys.t.copy_(k0)
    _, shape_k, _ = at.tune_radix(keys, reps=2)
    plain_s, _ = timed(lambda: fa.radix_sort(keys))
    ref = keys.t.clone()
    fa.check(fa.lib().fhip_prof_enable(1))
    learned_s, fb = timed(lambda: model.sort(keys))
    stages = {}
    for name in ("learned_count", "learned_scatter", "learned_split", "learned_leaf_sort"):
        ms, cnt = _prof_ms(fa, name)
        stages[name] = {"ms_per_step": ms / steps, "launches_per_step": cnt / steps}
    fa.check(fa.lib().fhip_prof_enable(0))
    exact = bool(torch.equal(keys.t, ref))
    # the same pipeline without the network: equal-width coarse ranges, and
    # the training sample's quantiles as coarse splitters
    flat_s, fb_flat = timed(lambda: model.sort(keys, coarse="equal_width"))
    exact = exact and bool(torch.equal(keys.t, ref))
    quant_s, fb_quant = timed(lambda: model.sort(keys, coarse="quantile"))
    exact = exact and bool(torch.equal(keys.t, ref))
    del ref
    return {"keys": n, "plain_ms": plain_s * 1e3, "learned_ms": learned_s * 1e3,
            "speedup": plain_s / learned_s, "bit_exact": exact,
            "fallback_keys": fb[0], "fallback_buckets": fb[1], "max_coarse_bucket": fb[2],
            "equal_width_ms": flat_s * 1e3, "equal_width_fallback_keys": fb_flat[0],
            "equal_width_max_coarse_bucket": fb_flat[2],
            "quantile_ms": quant_s * 1e3, "quantile_fallback_keys": fb_quant[0],
            "quantile_max_coarse_bucket": fb_quant[2],
            "network_gain_over_quantile": quant_s / learned_s,
            "plain_keys_per_s": n / plain_s, "learned_keys_per_s": n / learned_s,
            "train_s": train_s, "model_loss": model.info()["loss"], "stages": stages,
            "plain_tile_shape": list(at.RADIX_SHAPES[shape_k]) if shape_k is not None else None}


def bench_batch(fa, steps=3):
    """4096 independent sort + group sum/count jobs, row counts log-uniform
    in [256, 2^20] rescaled to 2^28 rows, int64 keys U[0, rows/4), double
    values U[0,1); one HIP graph launch per step (replayed)."""
    import math

    import numpy as np
    import torch

    from frovedis_amd import datagen as dg
    sizes = dg.segment_sizes()
    keys, vals, off = dg.batch_inputs_dev(sizes)
    k = fa.DeviceArray(keys, np.int64)
    v = fa.DeviceArray(vals, np.float64)
    o = fa.DeviceArray(off, np.uint64)
    out = fa.batch_sort_groupby(k, v, o, key_bits=18, use_graph=True)  # capture
    best = float("inf")
    for _ in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fa.batch_sort_groupby(k, v, o, key_bits=18, use_graph=True, out=out)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    n = int(sizes.sum())
    G = out["ngroups"]
    # algorithmic bytes: per segment 8 + 32 * ceil(b_s / 8) + 16 per row
    # (b_s = bits of rows_s / 4), plus 24 per group
    bits = np.array([max(1, math.ceil(math.log2(max(s // 4, 2)))) for s in sizes])
    algo = int(np.sum(sizes * (8 + 32 * np.ceil(bits / 8) + 16))) + 24 * G
    return {"rows": n, "segments": len(sizes), "groups": G, "ms": best * 1e3,
            "rows_per_s": n / best, "algorithmic_bytes": algo,
            "hbm_frac": algo / best / HBM_PEAK, "graph": True}


def bench_join(fa, steps=3, log2_left=27, log2_right=29):
    """equi_join of a unique 2^27 key column (a random permutation of
    [0, 2^27)) with a 2^29 column (90% keys in [0, 2^27), 10% misses in
    [2^27, 2^28)).  "hash": the unique side is the build side
    (right_df.hash_join(left_df, eq(rk, lk)) -> unique_hashtable, output in
    probe order); "sortmerge": left_df.hash_join(right_df, eq(lk, rk)) whose
    build side is not unique (sort both, merge; (key, left, right) order).
    Algorithmic bytes: 8 per input row + 16 per output pair."""
    import numpy as np
    import torch

    from frovedis_amd import datagen as dg
    nl, nr = 1 << log2_left, 1 << log2_right
    lk = fa.DeviceArray(dg.perm_keys_dev(nl), np.int64)
    rk = fa.DeviceArray(dg.fk_keys_dev(nr, domain=nl), np.int64)
    li = fa.DeviceArray(torch.arange(nl, dtype=torch.int64, device="cuda"), np.uint64)
    ri = fa.DeviceArray(torch.arange(nr, dtype=torch.int64, device="cuda"), np.uint64)
    hits = int((rk.t < nl).sum())
    out = {"left_rows": nl, "right_rows": nr, "pairs": hits}
    legs = {"hash": (rk, ri, lk, li, "auto", ("join_dense_stats", "join_dense_build_part",
                                              "join_dense_build", "join_dense_scatter",
                                              "join_dense_verify",
                                              "join_dense_part", "join_dense_probe",
                                              "join_dense_emit", "join_build", "join_probe",
                                              "join_compact")),
            "sortmerge": (lk, li, rk, ri, "auto", ("radix_hist", "radix_pass", "join_merge",
                                                    "join_write"))}
    for name, (a, ai, b, bi, strat, kernels) in legs.items():
        lo, ro, used = fa.equi_join(a, ai, b, bi, strategy=strat, capacity=nr)
        ok = len(lo) == hits and bool(torch.equal(a.t[lo.t], b.t[ro.t]))
        ok = ok and used == (name == "hash")
        del lo, ro
        torch.cuda.empty_cache()
        best = float("inf")
        fa.check(fa.lib().fhip_prof_enable(1))
        for _ in range(steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            lo, ro, used = fa.equi_join(a, ai, b, bi, strategy=strat, capacity=nr)
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
            del lo, ro
        kern = {}
        for kname in kernels:
            ms, cnt = _prof_ms(fa, kname)
            kern[kname] = {"ms_per_step": ms / steps, "launches_per_step": cnt / steps}
        fa.check(fa.lib().fhip_prof_enable(0))
        algo = 8 * (nl + nr) + 16 * hits
        out[name] = {"ms": best * 1e3, "rows_per_s": (nl + nr) / best,
                     "algorithmic_bytes": algo, "hbm_frac": algo / best / HBM_PEAK,
                     "pairs_checked": ok, "kernels": kern}
    return out


def _timed_min(fn, steps):
    import torch
    best, out = float("inf"), None
    for _ in range(steps):
        out = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best, out


def bench_relational(fa, log2n=28, steps=3):
    """The column operations around the core on the group-by table (2^log2n
    rows: int64 key Zipf(1.1) over 2^20, double N(0,1), int32 U[0,1000)) and
    on the join tables (2^29 foreign keys, 10% without a partner, outer-joined
    to 2^27 unique keys):
      filter      iv < 500 (filter_lt_immed): reads 4 B/row, writes 8 B per hit;
      extract     v at the filter's rows (global_extract): 8 B index + 8 B
                  value read + 8 B written per hit;
      sort        dftable::sort on the key column: key/value radix sort of
                  (key, row id), 3 significant key bytes -> 3 passes of 32 B/row
                  plus the 8 B/row histogram read;
      group_by2   group_by({key, iv}): the key ranges (16 + 12 B/row: row id
                  and column), the 30-bit composite key (8 + 8 + 4 read, 4
                  written), its key/value radix sort (4 B/row histogram, 4
                  passes of 24 B/row) and the run boundaries (2 x 4 B/row):
                  160 B/row;
      star_join   star join of the table on (key, iv) to a dimension of 2^20
                  unique keys and one of 900 int32 keys: 12 B/row read, 24 B
                  per surviving row written, 16 B per dimension row built;
      outer_join  outer_hash_join of the 2^29 side to the unique 2^27 side:
                  8 B per input row + 16 B per pair + 8 B per left-only row."""
    import numpy as np
    import torch

    from frovedis_amd import datagen as dg
    n = 1 << log2n
    k, v, iv = gen_groupby(n)
    out = {"rows": n}

    def stage(name, fn, algo_bytes, kernels, check=None):
        fn()  # warm-up: the workspace pool reaches its size
        torch.cuda.synchronize()
        fa.check(fa.lib().fhip_prof_enable(1))
        t, res = _timed_min(fn, steps)
        kern = {}
        for kname in kernels:
            ms, cnt = _prof_ms(fa, kname)
            kern[kname] = {"ms_per_step": ms / steps, "launches_per_step": cnt / steps}
        fa.check(fa.lib().fhip_prof_enable(0))
        ab = algo_bytes(res)
        out[name] = {"ms": t * 1e3, "rows_per_s": n / t, "algorithmic_bytes": ab,
                     "hbm_frac": ab / t / HBM_PEAK, "kernels": kern}
        if check is not None:
            out[name]["checked"] = bool(check(res))
        return res

    sel = stage("filter", lambda: fa.filter(iv, "lt", np.int32(500)),
                lambda r: 4 * n + 8 * len(r), ["filter"],
                lambda r: len(r) == int((iv.t < 500).sum()))
    stage("extract", lambda: fa.global_extract(v, sel), lambda r: 24 * len(sel), ["gather"],
          lambda r: bool(torch.equal(r.t, v.t[sel.t])))
    del sel
    stage("sort", lambda: fa.sort(k), lambda r: 8 * n + 3 * 32 * n,
          ["radix_hist", "radix_pass"],
          lambda r: bool((r[0].t[1:] >= r[0].t[:-1]).all()) and
          bool(torch.equal(k.t[r[1].t], r[0].t)))
    torch.cuda.empty_cache()
    g = stage("group_by2", lambda: fa.multi_group_by([k, iv]),
              lambda r: 160 * n,
              ["multi_group_range", "multi_group_composite", "radix_hist", "radix_pass",
               "multi_group_split"])
    out["group_by2"]["groups"] = len(g[1]) - 1
    del g
    # star join of the table to two dimensions on (key, iv): 2^20 unique
    # keys, and 900 of the 1000 iv values
    d1 = fa.DeviceArray(dg.perm_keys_dev(1 << 20), np.int64)
    d1i = fa.DeviceArray(torch.arange(1 << 20, dtype=torch.int64, device="cuda"), np.uint64)
    d2 = fa.DeviceArray(torch.arange(900, dtype=torch.int32, device="cuda").flip(0), np.int32)
    d2i = fa.DeviceArray(torch.arange(900, dtype=torch.int64, device="cuda"), np.uint64)
    want = int(((k.t < (1 << 20)) & (iv.t < 900)).sum())
    stage("star_join", lambda: fa.star_join([k, iv], [(d1, d1i), (d2, d2i)]),
          lambda r: 12 * n + 24 * len(r[0]) + 16 * ((1 << 20) + 900),
          ["join_build", "join_probe", "join_compact"],
          lambda r: len(r[0]) == want and
          bool(torch.equal(d1.t[r[1][0].t], k.t[r[0].t])) and
          bool(torch.equal(d2.t[r[1][1].t], iv.t[r[0].t])))
    del k, v, iv, d1, d1i, d2, d2i
    torch.cuda.empty_cache()

    nl, nr = 1 << 27, 1 << 29
    lk = fa.DeviceArray(dg.perm_keys_dev(nl), np.int64)
    rk = fa.DeviceArray(dg.fk_keys_dev(nr, domain=nl), np.int64)
    li = fa.DeviceArray(torch.arange(nl, dtype=torch.int64, device="cuda"), np.uint64)
    ri = fa.DeviceArray(torch.arange(nr, dtype=torch.int64, device="cuda"), np.uint64)
    hits = int((rk.t < nl).sum())

    def outer():
        return fa.outer_equi_join(rk, ri, lk, li, capacity=nr)

    res = stage("outer_join", outer, lambda r: 8 * (nl + nr) + 16 * hits + 8 * (nr - hits),
                ["join_build", "join_probe", "join_compact"],
                lambda r: len(r[0]) == hits and len(r[2]) == nr - hits and r[3] and
                bool(torch.equal(rk.t[r[0].t], lk.t[r[1].t])) and
                bool((rk.t[r[2].t] >= nl).all()))
    out["outer_join"].update({"left_rows": nr, "right_rows": nl, "pairs": hits,
                              "left_only": nr - hits})
    out["outer_join"]["rows_per_s"] = (nl + nr) / (out["outer_join"]["ms"] / 1e3)
    del res
    return out


def bench_copy(log2n, steps=3):
    """A plain device copy of 2^log2n int64 (reads and writes 8 B each): the
    reference point for the HBM byte counters and the achievable bandwidth."""
    import torch
    n = 1 << log2n
    src = torch.arange(n, dtype=torch.int64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    best = float("inf")
    for _ in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dst.copy_(src)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return {"bytes_read": 8 * n, "bytes_written": 8 * n, "ms": best * 1e3,
            "gb_per_s": 16 * n / best / 1e9, "hbm_frac": 16 * n / best / HBM_PEAK}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rows", type=int, default=28, help="log2 of the row count")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true", help="skip the side measurements")
    ap.add_argument("--leg", default="all",
                    choices=["all", "groupby", "radix", "batch", "join", "learned", "copy",
                             "relational", "radix_host"],
                    help="run one measurement alone (for profiling); 'copy' is a plain "
                         "device copy of a known size, the counter calibration")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU's share only (--gpus {args.gpus} requested); "
              "sharding the rows over several GPUs is not part of this build.", file=sys.stderr)
        sys.exit(2)

    import torch

    import frovedis_amd as fa
    torch.cuda.set_device(0)
    fa.lib()
    if args.leg not in ("all", "groupby"):
        legs = {"radix": lambda: bench_radix(fa, args.rows),
                "batch": lambda: bench_batch(fa),
                "join": lambda: bench_join(fa),
                "learned": lambda: bench_learned(fa, args.rows),
                "relational": lambda: bench_relational(fa, args.rows),
                "radix_host": lambda: bench_radix_host(fa, min(args.rows, 26)),
                "copy": lambda: bench_copy(args.rows)}
        print(f"# leg {args.leg}", file=sys.stderr, flush=True)
        print(json.dumps({"leg": args.leg, **legs[args.leg]()}))
        return
    if args.leg == "groupby":
        args.no_extra = args.no_cpu = True
    n = 1 << args.rows
    print(f"# generating {n} rows", file=sys.stderr, flush=True)
    k, v, iv = gen_groupby(n)
    model, train_s = groupby_model(fa, n)
    hint = model.table_hint(n)
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        step_groupby(fa, k, v, iv, model)
    torch.cuda.synchronize()
    # per-kernel breakdown of the step (every named kernel timed), outside
    # the timed region: it finds the dominant kernel
    fa.check(fa.lib().fhip_prof_enable(1))
    for _ in range(args.steps):
        step_groupby(fa, k, v, iv, model)
    torch.cuda.synchronize()
    breakdown = {}
    for name in ("groupby_hot", "groupby_hot_merge", "groupby_range_part", "groupby_range_agg",
                 "groupby_range_emit", "groupby_part_hist", "groupby_part_scatter", "hashagg",
                 "groupby_merge", "radix_hist", "radix_pass"):
        ms, cnt = _prof_ms(fa, name)
        breakdown[name] = {"ms_per_step": ms / args.steps, "launches_per_step": cnt / args.steps}
    fa.check(fa.lib().fhip_prof_enable(0))
    # the roofline is taken on the step's dominant kernel (the longest one)
    kernel_of = {"groupby_hot": "hot_agg_kernel", "groupby_part_hist": "pa_count_kernel",
                 "groupby_part_scatter": "pa_scatter_kernel", "hashagg": "pa_agg_kernel",
                 "groupby_merge": "pa_merge_kernel"}
    dom = max(kernel_of, key=lambda nm: breakdown[nm]["ms_per_step"])

    # the timed region: HIP events around the dominant kernel only
    fa.check(fa.lib().fhip_prof_filter(dom.encode()))
    fa.check(fa.lib().fhip_prof_enable(1))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gk, outs = step_groupby(fa, k, v, iv, model)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    dom_ms, dom_cnt = _prof_ms(fa, dom)
    fa.check(fa.lib().fhip_prof_enable(0))
    fa.check(fa.lib().fhip_prof_filter(None))
    ms_per_step = dt * 1e3 / args.steps
    groups = len(gk)
    print(f"# {groups} groups, {ms_per_step:.3f} ms/step", file=sys.stderr, flush=True)
    kern_s = dom_ms / 1e3 / max(dom_cnt, 1)
    algo_bytes = GROUPBY_BYTES_PER_ROW * n
    achieved = algo_bytes / kern_s if kern_s else 0.0
    traffic = _traffic_from_profiles(kernel_of[dom])
    result = {
        "metric": "group-by rows/s (sum/count/mean/max, Zipf keys)",
        "value": n / (ms_per_step / 1e3),
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 on device)",
        "config": {"workload": f"dftable group_by + sum/count/avg/max, 2^{args.rows} rows, "
                               "int64 key Zipf(1.1) over 2^20, double N(0,1), int32 U[0,1000)",
                   "rows": n, "groups": groups, "path": "fused hash group-by"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": traffic, "kernel": kernel_of[dom],
                     "kernel_ms": kern_s * 1e3,
                     "algorithmic_bytes_per_launch": algo_bytes},
        "cpu_baseline": None,
        "extra": {"groupby_kernels": breakdown,
                  "table_hint": {"from": "CdfModel of an earlier 2^24-row input (another seed)",
                                 "estimated_groups": hint, "groups": groups,
                                 "train_s": train_s}},
    }
    del k, v, iv, gk, outs
    torch.cuda.empty_cache()
    if not args.no_extra:
        print("# radix sort side measurements", file=sys.stderr, flush=True)
        result["extra"]["radix_kv_u64_2^24"] = bench_radix(fa, 24)
        result["extra"]["radix_kv_u64_2^28"] = bench_radix(fa, 28)
        result["extra"]["radix_kv_u64_2^24_host_vectors"] = bench_radix_host(fa, 24)
        torch.cuda.empty_cache()
        print("# batched segments", file=sys.stderr, flush=True)
        result["extra"]["batched_segments_2^28"] = bench_batch(fa)
        print("# equi_join", file=sys.stderr, flush=True)
        result["extra"]["join_2^27_x_2^29"] = bench_join(fa)
        torch.cuda.empty_cache()
        print("# learned vs plain sort", file=sys.stderr, flush=True)
        result["extra"]["learned_sort_2^28"] = bench_learned(fa, 28)
        torch.cuda.empty_cache()
        print("# filter / extract / sort / group_by2 / outer join", file=sys.stderr, flush=True)
        result["extra"]["relational_2^28"] = bench_relational(fa, 28)
        torch.cuda.empty_cache()
    if not args.no_cpu:
        print("# cpu baseline (oracle)", file=sys.stderr, flush=True)
        base = cpu_baselines()
        result["cpu_baseline"] = cpu_baseline_groupby(base)
        result["extra"]["cpu_baselines"] = base
    print(json.dumps(result))


if __name__ == "__main__":
    main()

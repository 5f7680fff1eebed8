"""Seeded synthetic inputs for the benchmark configurations.

SplitMix64 is counter based: element i of stream `seed` is
splitmix64(seed + (i + 1) * golden), so the host (numpy) and device (torch)
generators produce the same bits and any slice can be generated alone.
Integer-valued outputs (keys, ranks, int32 columns) are bit-identical on host
and device; FP outputs that go through log/cos (normal, lognormal) may
differ in the last ulp between the two.
"""
import math

import numpy as np

GOLDEN = 0x9E3779B97F4A7C15
M1 = 0xBF58476D1CE4E5B9
M2 = 0x94D049BB133111EB

SEED_RADIX = 0x5EED0001
SEED_ZIPF = 0x5EED0002
SEED_NORMAL = 0x5EED0003
SEED_INT = 0x5EED0004
SEED_PERM = 0x5EED0005
SEED_FK = 0x5EED0006
SEED_SEGS = 0x5EED0007
SEED_LOGN = 0x5EED0008


def _s64(x):
    """uint64 constant as the int64 with the same bits (torch carrier)."""
    x &= (1 << 64) - 1
    return x - (1 << 64) if x >= (1 << 63) else x


# ---------------------------------------------------------------- host (numpy)
def splitmix64_np(seed, start, n):
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * np.uint64(GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(M1)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(M2)
        return z ^ (z >> np.uint64(31))


def uniform01_np(seed, start, n):
    return (splitmix64_np(seed, start, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def zipf_cdf(n_values=1 << 20, s=1.1):
    k = np.arange(1, n_values + 1, dtype=np.float64)
    w = k ** (-s)
    c = np.cumsum(w)
    return c / c[-1]


def zipf_keys_np(n, start=0, n_values=1 << 20, s=1.1, seed=SEED_ZIPF):
    u = uniform01_np(seed, start, n)
    return (np.searchsorted(zipf_cdf(n_values, s), u, side="right") + 1).astype(np.int64)


def int_uniform_np(n, hi, start=0, seed=SEED_INT):
    return ((splitmix64_np(seed, start, n) >> np.uint64(32)) % np.uint64(hi)).astype(np.int32)


def normal_np(n, start=0, seed=SEED_NORMAL):
    # Box-Muller on two interleaved uniforms of the same stream
    u = uniform01_np(seed, 2 * start, 2 * n)
    u1 = 1.0 - u[0::2]
    u2 = u[1::2]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


# --------------------------------------------------------------- device (torch)
def _torch():
    import torch
    return torch


def _srl(torch, z, s):
    # logical right shift on the int64 carrier
    return (z >> s) & ((1 << (64 - s)) - 1)


def splitmix64_dev(seed, start, n, device="cuda"):
    torch = _torch()
    i = torch.arange(start + 1, start + n + 1, dtype=torch.int64, device=device)
    z = i * _s64(GOLDEN) + _s64(seed)
    z = (z ^ _srl(torch, z, 30)) * _s64(M1)
    z = (z ^ _srl(torch, z, 27)) * _s64(M2)
    return z ^ _srl(torch, z, 31)


def uniform01_dev(seed, start, n, device="cuda"):
    torch = _torch()
    z = splitmix64_dev(seed, start, n, device)
    return _srl(torch, z, 11).to(torch.float64) * (2.0 ** -53)


def zipf_keys_dev(n, start=0, n_values=1 << 20, s=1.1, seed=SEED_ZIPF, device="cuda",
                  chunk=1 << 26):
    torch = _torch()
    cdf = torch.from_numpy(zipf_cdf(n_values, s)).to(device)
    out = torch.empty(n, dtype=torch.int64, device=device)
    for b in range(0, n, chunk):
        m = min(chunk, n - b)
        u = uniform01_dev(seed, start + b, m, device)
        out[b:b + m] = torch.searchsorted(cdf, u, right=True) + 1
        del u
    return out


def int_uniform_dev(n, hi, start=0, seed=SEED_INT, device="cuda", chunk=1 << 26):
    torch = _torch()
    out = torch.empty(n, dtype=torch.int32, device=device)
    for b in range(0, n, chunk):
        m = min(chunk, n - b)
        z = splitmix64_dev(seed, start + b, m, device)
        out[b:b + m] = (_srl(torch, z, 32) % hi).to(torch.int32)
    return out


def normal_dev(n, start=0, seed=SEED_NORMAL, device="cuda", chunk=1 << 25):
    torch = _torch()
    out = torch.empty(n, dtype=torch.float64, device=device)
    for b in range(0, n, chunk):
        m = min(chunk, n - b)
        u = uniform01_dev(seed, 2 * (start + b), 2 * m, device)
        u1 = 1.0 - u[0::2]
        u2 = u[1::2]
        out[b:b + m] = torch.sqrt(-2.0 * torch.log(u1)) * torch.cos(2.0 * math.pi * u2)
    return out


def radix_keys_dev(n, seed=SEED_RADIX, device="cuda"):
    """Full-range uint64 keys (int64 carrier) for the radix kv config."""
    return splitmix64_dev(seed, 0, n, device)


# ------------------------------------------------ lognormal mixture (learned sort)
# 8 components: weight, mean and sigma of ln(key)
LOGN_WEIGHTS = (0.20, 0.15, 0.15, 0.10, 0.10, 0.10, 0.10, 0.10)
LOGN_MU = (4.0, 8.0, 12.0, 16.0, 20.0, 24.0, 28.0, 34.0)
LOGN_SIGMA = (0.5, 1.0, 1.5, 0.8, 2.0, 1.2, 0.6, 1.5)
RUN_LEN = 4096
RUN_FRACTION = 0.05


def lognormal_mixture_dev(n, seed=SEED_LOGN, device="cuda", chunk=1 << 25, runs=True):
    """int64 keys: 95% floor(exp(N(mu, sigma))) from the 8-component mixture,
    clipped to [0, 2^62); 5% of the RUN_LEN-aligned runs are pre-sorted."""
    torch = _torch()
    cw = torch.tensor(np.cumsum(LOGN_WEIGHTS), dtype=torch.float64, device=device)
    cw[-1] = 1.0
    mu = torch.tensor(LOGN_MU, dtype=torch.float64, device=device)
    sg = torch.tensor(LOGN_SIGMA, dtype=torch.float64, device=device)
    out = torch.empty(n, dtype=torch.int64, device=device)
    lim = float(2 ** 62 - 1024)
    for b in range(0, n, chunk):
        m = min(chunk, n - b)
        u = uniform01_dev(seed, 3 * b, 3 * m, device)
        comp = torch.searchsorted(cw, u[0::3], right=True).clamp_(max=7)
        z = torch.sqrt(-2.0 * torch.log(1.0 - u[1::3])) * torch.cos(2.0 * math.pi * u[2::3])
        v = torch.exp(mu[comp] + sg[comp] * z).clamp_(0.0, lim)
        out[b:b + m] = torch.floor(v).to(torch.int64)
        del u, comp, z, v
    if runs and n >= RUN_LEN:
        nruns = n // RUN_LEN
        pick = splitmix64_dev(seed + 1, 0, nruns, device)
        sel = torch.nonzero(_srl(torch, pick, 11).to(torch.float64) * (2.0 ** -53)
                            < RUN_FRACTION).flatten()
        view = out[: nruns * RUN_LEN].view(nruns, RUN_LEN)
        for c in range(0, len(sel), 1024):
            rows = sel[c:c + 1024]
            view[rows] = torch.sort(view[rows], dim=1).values
    return out


# ---------------------------------------------------------- batched segments
def segment_sizes(nseg=4096, total=1 << 28, lo=256, hi=1 << 20, seed=SEED_SEGS):
    """Row counts log-uniform on [lo, hi], rescaled by one common factor and
    clamped to [lo, hi] so that they add up to `total` (the last rounding
    remainder goes to the segments with room, largest first)."""
    u = uniform01_np(seed, 0, nseg)
    raw = np.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))

    def sized(f):
        return np.clip(np.rint(raw * f), lo, hi).astype(np.int64)

    a, b = 1e-6, 1e6
    for _ in range(200):
        mid = math.sqrt(a * b)
        if sized(mid).sum() < total:
            a = mid
        else:
            b = mid
    s = sized(a)
    rem = total - int(s.sum())
    order = np.argsort(-s, kind="stable")
    i = 0
    while rem != 0 and i < 10 * nseg:
        j = order[i % nseg]
        step = 1 if rem > 0 else -1
        if lo <= s[j] + step <= hi:
            s[j] += step
            rem -= step
        i += 1
    return s


def batch_inputs_dev(sizes, seed=SEED_SEGS + 1, device="cuda"):
    """keys int64 U[0, rows_s / 4) per segment, values double U[0, 1)."""
    torch = _torch()
    sizes_t = torch.from_numpy(np.asarray(sizes, np.int64)).to(device)
    n = int(sizes_t.sum())
    seg = torch.repeat_interleave(torch.arange(len(sizes), device=device), sizes_t)
    span = torch.clamp(sizes_t // 4, min=1)[seg]
    u = uniform01_dev(seed, 0, n, device)
    keys = torch.floor(u * span.to(torch.float64)).to(torch.int64)
    del u, seg, span
    vals = uniform01_dev(seed + 1, 0, n, device)
    off = torch.zeros(len(sizes) + 1, dtype=torch.int64, device=device)
    off[1:] = torch.cumsum(sizes_t, 0)
    return keys, vals, off


# ------------------------------------------------------------- join inputs
def perm_keys_np(nl, seed=SEED_PERM):
    """A uniformly random permutation of [0, nl): the stable argsort of the
    stream's values (top 63 bits), identical on host and device."""
    z = splitmix64_np(seed, 0, nl) >> np.uint64(1)
    return np.argsort(z.astype(np.int64), kind="stable").astype(np.int64)


def perm_keys_dev(nl, seed=SEED_PERM, device="cuda"):
    torch = _torch()
    z = _srl(torch, splitmix64_dev(seed, 0, nl, device), 1)
    return torch.sort(z, stable=True).indices.to(torch.int64)


def fk_keys_np(nr, domain=1 << 27, hit=0.9, seed=SEED_FK):
    """Probe keys: with probability `hit` U[0, domain), else
    U[domain, 2 * domain) (a non-negative miss domain).  Element i uses
    draws 2i (hit or miss) and 2i+1 (the key); domain is a power of two."""
    bits = domain.bit_length() - 1
    z = splitmix64_np(seed, 0, 2 * nr)
    u = (z[0::2] >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    k = (z[1::2] >> np.uint64(64 - bits)).astype(np.int64)
    return np.where(u < hit, k, k + domain)


def fk_keys_dev(nr, domain=1 << 27, hit=0.9, seed=SEED_FK, device="cuda", chunk=1 << 26):
    torch = _torch()
    bits = domain.bit_length() - 1
    out = torch.empty(nr, dtype=torch.int64, device=device)
    for b in range(0, nr, chunk):
        m = min(chunk, nr - b)
        z = splitmix64_dev(seed, 2 * b, 2 * m, device)
        u = _srl(torch, z[0::2], 11).to(torch.float64) * (2.0 ** -53)
        k = _srl(torch, z[1::2], 64 - bits)
        out[b:b + m] = torch.where(u < hit, k, k + domain)
        del z, u, k
    return out

"""CPU checks of the synthetic generators' host side."""
import numpy as np

from frovedis_amd import datagen as dg


def test_splitmix64_reference_values():
    # SplitMix64 from state 0: the first outputs of Vigna's reference
    # generator (seed 0, one golden-ratio increment per output)
    out = dg.splitmix64_np(0, 0, 3)
    assert [int(x) for x in out] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_zipf_head_share():
    k = dg.zipf_keys_np(1 << 20)
    assert k.min() >= 1 and k.max() <= 1 << 20
    share = np.mean(k == 1)
    assert abs(share - 0.124) < 0.01  # 1 / sum_{k<=2^20} k^-1.1


def test_segment_sizes_add_up():
    s = dg.segment_sizes()
    assert s.sum() == 1 << 28 and s.min() >= 256 and s.max() <= 1 << 20 and len(s) == 4096


def test_int_uniform_range():
    a = dg.int_uniform_np(100000, 1000)
    assert a.min() >= 0 and a.max() < 1000 and len(np.unique(a)) == 1000


def test_join_generators():
    from frovedis_amd import datagen as dg
    p = dg.perm_keys_np(1 << 12)
    assert np.array_equal(np.sort(p), np.arange(1 << 12))
    k = dg.fk_keys_np(1 << 16, domain=1 << 10)
    assert k.min() >= 0 and k.max() < 1 << 11
    hit = (k < 1 << 10).mean()
    assert abs(hit - 0.9) < 0.01
    # element i depends only on draws 2i, 2i+1 (chunked generation agrees)
    assert np.array_equal(dg.fk_keys_np(100, domain=1 << 10), k[:100])

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

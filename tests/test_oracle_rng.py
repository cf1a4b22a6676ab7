"""Pins of the oracle's counter RNGs (CPU).

Philox4x32-10 has no implementation in the reference (its generator is SplitMix64), so it is
pinned by the published Random123 known-answer vectors (Salmon et al., SC'11) and by the key=42
vectors of the workload.  SplitMix64 + u01 restate cuPythia/common/rng.cuh:47-57 and are pinned
by the reference's integer KAT: 1,680,778 of 2^24 candidates accepted
(cuPythia/10_cub_compaction/README.md:12, parameters cub_compaction.cu:60-63)."""
import numpy as np


def test_philox_random123_kats(O):
    assert O.philox_block([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert O.philox_block([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert O.philox_block([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_philox_key42_vectors(O):
    assert O.philox_block([0, 0, 0, 0], [42, 0]) == [0x9ceaf053, 0x77f5493b, 0x12bf50ad, 0x5742b3d7]
    assert O.philox_block([1, 0, 0, 0], [42, 0]) == [0xfcdb2127, 0x53ba6cfd, 0x838f5a6e, 0x744e06fb]


def test_philox_fill_layout(O):
    # event e of stream s -> counter (e_lo, e_hi, s, 0); key (key_lo, key_hi)
    w = O.philox_fill(42, (1 << 32) + 5, 3, stream=7)
    for i in range(3):
        e = (1 << 32) + 5 + i
        assert list(w[4 * i:4 * i + 4]) == O.philox_block([e & 0xffffffff, e >> 32, 7, 0], [42, 0])
    # a 64-bit key splits into (lo, hi)
    key = 0x0123456789abcdef
    w = O.philox_fill(key, 0, 1)
    assert list(w) == O.philox_block([0, 0, 0, 0], [key & 0xffffffff, key >> 32])


def test_uniform_definitions(O):
    L = O.lib()
    assert L.orc_u53(0, 0) == 0.0
    assert L.orc_u53(0xffffffff, 0xffffffff) == 1.0 - 2.0 ** -53
    assert L.orc_u53(0x800, 0) == 2.0 ** -53
    assert L.orc_u24(0xffffffff) == np.float32(1.0 - 2.0 ** -24)
    assert L.orc_u24(0x100) == np.float32(2.0 ** -24)


def test_splitmix64_reference_kat(O):
    L = O.lib()
    # SplitMix64 of 0 (the published SplitMix64 first output for state 0 after one increment)
    assert L.orc_splitmix64(0) == 0xe220a8397b1dcdaf
    wmax = L.orc_c1_wmax_scan(2000000, 1e4, 0.118, 0.9)
    assert L.orc_c1_unweight_compat(0xC0B0, 1 << 24, 1e4, 0.118, 0.9, wmax, None, None) == 1680778

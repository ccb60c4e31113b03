"""Philox4x32-10 known-answer vectors (Random123 kat_vectors, philox4x32 10 rounds)."""
import numpy as np

from oracle.rng import philox4x32, sweep_draws


def _h(x):
    return [int(v) for v in x]


def test_philox_known_answers():
    z = philox4x32(np.zeros(4, np.uint32), np.zeros(2, np.uint32))
    assert _h(z) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    f = philox4x32(np.full(4, 0xFFFFFFFF, np.uint32), np.full(2, 0xFFFFFFFF, np.uint32))
    assert _h(f) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    pi = philox4x32(np.array([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], np.uint32),
                    np.array([0xA4093822, 0x299F31D0], np.uint32))
    assert _h(pi) == [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_uniforms_in_unit_interval_and_distinct_streams():
    u, w = sweep_draws(7, 3, 5, np.arange(4096))
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.02
    u2, _ = sweep_draws(7, 3, 6, np.arange(4096))
    assert not np.array_equal(u, u2)

"""Pins the Philox4x32-10 stream shared by the oracle and the GPU kernels."""
import numpy as np


def test_philox_known_answers(oracle_mod):
    # Random123 known-answer vectors for philox4x32_10
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in kat:
        assert list(oracle_mod.philox(ctr, key)) == want


def test_gauss_pair_moments(oracle_mod):
    z = np.array([oracle_mod.gauss_pair(11, 0, 1, c, i) for c in range(40) for i in range(250)]).ravel()
    assert abs(z.mean()) < 0.03
    assert abs(z.var() - 1.0) < 0.04
    # distinct tags, systems and counters give distinct draws
    a = oracle_mod.gauss_pair(11, 0, 1, 5, 3)
    assert a != oracle_mod.gauss_pair(11, 0, 2, 5, 3)
    assert a != oracle_mod.gauss_pair(11, 1, 1, 5, 3)
    assert a != oracle_mod.gauss_pair(11, 0, 1, 6, 3)
    assert a == oracle_mod.gauss_pair(11, 0, 1, 5, 3)

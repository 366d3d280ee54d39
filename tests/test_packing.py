"""Synthetic packings: sizing rules and the spatial renumbering of rings."""
import numpy as np

from cudaring_amd import packing


def test_spatial_order_is_a_ring_permutation(oracle_mod):
    st = packing.lattice_system(50, 0.7, 9)
    s2, order = packing.spatial_order(st)
    assert sorted(order) == list(range(50))
    f1 = np.concatenate([[0], np.cumsum(st["nvp"])])
    f2 = np.concatenate([[0], np.cumsum(s2["nvp"])])
    for p, q in enumerate(order):
        np.testing.assert_array_equal(s2["pos"].reshape(-1, 2)[f2[p]:f2[p + 1]],
                                      st["pos"].reshape(-1, 2)[f1[q]:f1[q + 1]])
        assert s2["a0"][p] == st["a0"][q]
    # same physics: the total energy does not depend on the numbering
    from tests.helpers import oracle_from_state
    us = []
    for s in (st, s2):
        o = oracle_from_state(oracle_mod, s, interaction="all")
        o.calc_force_energy()
        us.append(o.potential_energy())
    assert abs(us[0] - us[1]) <= 1e-12 * abs(us[0])
    # neighbouring rings in memory are close in space
    c = np.array([s2["pos"].reshape(-1, 2)[f2[p]] for p in range(50)])
    step = np.linalg.norm(np.diff(c, axis=0), axis=1)
    c0 = np.array([st["pos"].reshape(-1, 2)[f1[p]] for p in range(50)])
    assert np.median(step) < 0.5 * np.median(np.linalg.norm(np.diff(c0, axis=0), axis=1))


def test_size_bidisperse_system():
    st = packing.size_bidisperse_system(40, 0.6, 2, n=32, ratio=1.4)
    assert np.all(st["nvp"] == 32)
    a = np.unique(np.round(st["a0"], 12))
    np.testing.assert_allclose(a, [1.0, 1.96])
    l0 = np.unique(np.round(st["l0"], 12))
    np.testing.assert_allclose(l0[1] / l0[0], 1.4, rtol=1e-10)

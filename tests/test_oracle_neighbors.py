"""Pins the oracle's Verlet list (H10) against brute force and the reference's row semantics."""
import numpy as np
import pytest

from cudaring_amd import packing
from tests.helpers import oracle_from_state


def brute_sets(st, cut):
    pos = st["pos"].reshape(-1, 2)
    L = st["box"]
    ring = np.repeat(np.arange(len(st["nvp"])), st["nvp"])
    d = pos[:, None, :] - pos[None, :, :]
    d -= L * np.round(d / L)          # numpy rounds half to even; no pair sits at exactly L/2 here
    dist = np.sqrt((d ** 2).sum(-1))
    rs = st["rad"][:, None] + st["rad"][None, :]
    nb = (-(1 - dist / rs) < cut) & (ring[:, None] != ring[None, :])
    return [list(np.nonzero(nb[v])[0]) for v in range(len(pos))]


@pytest.mark.parametrize("pot,skin", [("harmonic", 0.3), ("wca", 0.5), ("adhesive", 0.2), ("lj", 0.1)])
def test_list_matches_brute_force(oracle_mod, pot, skin):
    st = packing.random_system(14, 0.8, 21)
    o = oracle_from_state(oracle_mod, st, potential=pot)
    cut = o.set_displacement_cutoff(skin, 2 * st["rad"].mean())
    o.calc_neighbor_list(cut)
    rows, W, cnt = o.neighbors()
    ref = brute_sets(st, cut)
    assert rows == ref                         # same sets, ascending order
    assert W == 2 ** int(np.ceil(np.log2(cnt.max())))  # grow-only power-of-two width


def test_rebuild_on_displacement(oracle_mod):
    st = packing.random_system(8, 0.6, 2)
    o = oracle_from_state(oracle_mod, st)
    cut = o.set_displacement_cutoff(0.3, 2 * st["rad"].mean())
    o.calc_neighbor_list(cut)
    o.L.orc_reset_last_positions(o.h)
    assert o.check_neighbors() == 0
    pos = st["pos"].copy()
    pos[0] += 0.3 * 2 * st["rad"].mean()  # 3 |dr| > cutoff
    o.set("pos", pos)
    assert o.check_neighbors() == 1
    assert o.update_count() == 1
    st2 = dict(st); st2["pos"] = pos
    assert o.neighbors()[0] == brute_sets(st2, cut)


@pytest.mark.parametrize("pot,skin", [("harmonic", 0.3), ("wca", 0.5)])
def test_cell_list_rows_equal_all_pairs(oracle_mod, pot, skin):
    # a box of many cells, with rings translated by whole box lengths (positions are never
    # wrapped): the cell-list build gives the reference's all-pairs rows, entry for entry
    st = packing.random_system(300, 0.8, 31)
    pos = st["pos"].reshape(-1, 2).copy()
    ring = np.repeat(np.arange(300), st["nvp"])
    pos += np.stack([(ring % 3 - 1), (ring % 5 - 2)], 1) * st["box"]
    st["pos"] = pos.ravel()
    rows = []
    for cells in (True, False):
        o = oracle_from_state(oracle_mod, st, potential=pot)
        o.set_list_method(cells)
        cut = o.set_displacement_cutoff(skin, 2 * st["rad"].mean())
        o.calc_neighbor_list(cut)
        r, W, cnt = o.neighbors()
        rows.append((r, W, list(cnt)))
    assert rows[0] == rows[1]
    assert rows[0][0] == brute_sets(st, cut)
    assert max(rows[0][2]) > 0

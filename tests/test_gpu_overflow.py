"""A neighbour row that outgrows the list width inside a loop: the loop is replayed on a wider
list from a copy of its starting state, so no force uses a truncated row and the trajectory
equals the oracle's (whose rows have no width; the reference grows its width in
calcNeighborList, DPM2D.cu:3479-3507).

The rows are built on the rings shrunk to 0.3 of their size about their centres (few neighbours:
a width of 8 or 16), then the full-size rings are set without a rebuild.
The loop's first displacement test flags every system, and the in-loop rebuild finds rows of
tens of entries, past that width, so the loop is replayed on wider lists."""
import numpy as np
import pytest

from tests.conftest import rel_err
from tests.helpers import gpu_from_state, oracle_from_state

pytestmark = pytest.mark.gpu

SKIN = 2.0  # a wide list (radius 4 vertex diameters for harmonic) so that rows pass width 8


def states(nrings=16, phi=0.8, seed=5, shrink=0.3, potential="harmonic", interaction="vertexVertex"):
    """A packing relaxed with the test's interaction (no deep WCA overlaps), and its rings shrunk
    about their centres."""
    from tests.helpers import relaxed_state
    st = relaxed_state(nrings, phi, seed, final=(potential, interaction))
    pos = st["pos"].reshape(-1, 2)
    small = pos.copy()
    start = 0
    for n in st["nvp"]:
        ring = slice(start, start + n)
        c = pos[ring].mean(0)
        small[ring] = c + shrink * (pos[ring] - c)
        start += n
    return st, small.ravel()


def setup(st, small, potential, interaction, oracle_mod=None):
    size = 2 * float(np.mean(st["rad"]))
    if oracle_mod is None:
        s = gpu_from_state(st, potential=potential, interaction=interaction)
    else:
        s = oracle_from_state(oracle_mod, st, potential=potential, interaction=interaction)
    cut = s.set_displacement_cutoff(SKIN, size)
    s.set("pos", small)
    s.calc_neighbor_list(cut)
    return s, cut


@pytest.mark.parametrize("potential,interaction", [("harmonic", "vertexVertex"), ("wca", "vertexSmooth")])
def test_langevin_loop_widens_list_and_matches_oracle(oracle_mod, potential, interaction):
    st, small = states(potential=potential, interaction=interaction)
    g, cut = setup(st, small, potential, interaction)
    o, _ = setup(st, small, potential, interaction, oracle_mod)
    w0 = g.neighbors()[1]
    assert w0 <= 16
    for s in (g, o):
        s.init_langevin(1e-3, 1.0)
        s.set("pos", st["pos"])  # full-size rings, rows of the shrunk ones
        s.calc_force_energy()
    g.langevin(40)
    o.langevin(40)
    rows, w1, cnt = g.neighbors()
    assert w1 > w0 and cnt.max() > w0, (w0, w1, cnt.max())
    assert rel_err(g.get("pos"), o.get("pos"), 1.0) < 1e-11
    assert rel_err(g.get("force"), o.get("force")) < 1e-10
    assert g.update_count()[0] == o.update_count()
    # the rows after the loop are the oracle's
    orows = o.neighbors()[0]
    assert all(sorted(a) == sorted(b) for a, b in zip(rows, orows))


def test_fire_minimize_widens_list_and_matches_oracle(oracle_mod):
    st, small = states(seed=6)
    p7 = (0.2, 0.5, 1.1, 0.99, 1e-3, 1e-2, 0.2)
    runs = []
    for persistent in (1, 0):
        g, cut = setup(st, small, "harmonic", "vertexVertex")
        w0 = g.neighbors()[1]
        g.set("pos", st["pos"])
        g.calc_force_energy()
        g.init_fire(p7, 20, 0)
        g.set_fire_kernel(persistent, 0)
        g.fire_minimize(0.0, 300)
        assert g.neighbors()[1] > w0
        runs.append(g.get("pos"))
        g.close()
    o, _ = setup(st, small, "harmonic", "vertexVertex", oracle_mod)
    o.set("pos", st["pos"])
    o.calc_force_energy()
    o.init_fire(p7, 20, 0)
    o.fire(300)
    for pos in runs:
        assert rel_err(pos, o.get("pos"), 1.0) < 1e-10

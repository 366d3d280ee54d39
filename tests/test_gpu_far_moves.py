"""Step loops whose drift moves vertices farther than the cells' margin (cutoff / 3) in one step:
the next rebuild searches all vertices of the system (the reference's all-pairs list build,
cudaKernel.cuh:1735-1763) instead of the previous positions' cells, and the trajectory still
matches the oracle, which rebuilds from scratch."""
import numpy as np
import pytest

from tests.conftest import rel_err
from tests.helpers import gpu_from_state, oracle_from_state
from tests.test_oracle_rigid import ring_block

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("inter", ["vertexVertex", "vertexSmooth"])
def test_fast_rings_nve(oracle_mod, inter):
    print(f"case far_moves nve {inter}", flush=True)
    st = ring_block(gap=1.9, seed=3)
    dt = 2e-3
    o = oracle_from_state(oracle_mod, st, potential="wca", interaction=inter, costs=(1, 1, 0.1, 1), dt=dt)
    g = gpu_from_state(st, potential="wca", interaction=inter, costs=(1, 1, 0.1, 1), dt=dt)
    size = 2 * st["rad"].mean()
    cut_o = o.set_displacement_cutoff(0.5, size)
    cut_g = g.set_displacement_cutoff(0.5, size)
    o.calc_neighbor_list(cut_o)
    g.calc_neighbor_list(cut_g)
    o.calc_force_energy()
    g.calc_force_energy()
    o.init_nve(0.0)
    g.init_nve(0.0)
    # every ring translates at a speed that covers 0.6 * cutoff per step (cutoff = 0.5 size)
    rng = np.random.default_rng(1)
    speed = 0.6 * 0.5 * size / dt
    ang = rng.uniform(0, 2 * np.pi, len(st["nvp"]))
    v = np.repeat(np.c_[np.cos(ang), np.sin(ang)] * speed, st["nvp"], axis=0).ravel()
    o.set("vel", v)
    g.set("vel", v)
    for chunk in (1, 2):  # three steps: the rings meet hard soon after
        o.nve(chunk)
        g.nve(chunk)
        assert rel_err(g.get("pos"), o.get("pos"), scale=1.0) < 1e-11, chunk
        assert rel_err(g.get("force"), o.get("force")) < 1e-9, chunk
    assert g.update_count()[0] >= 2  # the list was rebuilt (all pairs) in these steps
    g.close()


@pytest.mark.parametrize("inter", ["vertexVertex", "vertexSmooth"])
def test_loop_start_far_from_list_positions(oracle_mod, inter):
    """Positions moved between loops, without a list rebuild, to 0.8 cutoff from the list positions
    (those of the last rebuild; here a translation of the whole packing, which keeps every pair
    distance): the step loop bins the list positions, whose cells allow for vertices at most
    2 cutoff / 3 away, so its start marks the system and the first rebuild searches all pairs;
    the trajectory matches the oracle, which rebuilds from scratch."""
    print(f"case far_moves loop start {inter}", flush=True)
    st = ring_block(gap=1.9, seed=5)
    dt = 2e-3
    o = oracle_from_state(oracle_mod, st, potential="wca", interaction=inter, costs=(1, 1, 0.1, 1), dt=dt)
    g = gpu_from_state(st, potential="wca", interaction=inter, costs=(1, 1, 0.1, 1), dt=dt)
    size = 2 * st["rad"].mean()
    for obj in (o, g):
        obj.calc_neighbor_list(obj.set_displacement_cutoff(0.5, size))
        obj.calc_force_energy()
        obj.init_nve(0.0)
    rng = np.random.default_rng(2)
    # rings drifting at 0.02 cutoff per step, the whole packing moved by 0.8 cutoff
    ang = rng.uniform(0, 2 * np.pi, len(st["nvp"]))
    v = np.repeat(np.c_[np.cos(ang), np.sin(ang)] * 0.02 * 0.5 * size / dt, st["nvp"], axis=0).ravel()
    shift = np.tile(0.8 * 0.5 * size * np.array([np.cos(0.7), np.sin(0.7)]), len(st["rad"]))
    for obj in (o, g):
        obj.set("pos", st["pos"] + shift)
        obj.set("vel", v)
        obj.calc_force_energy()
    u0 = g.update_count()[0]
    for chunk in (1, 4):
        o.nve(chunk)
        g.nve(chunk)
        assert rel_err(g.get("pos"), o.get("pos"), scale=1.0) < 1e-11, chunk
        assert rel_err(g.get("force"), o.get("force")) < 1e-9, chunk
    assert g.update_count()[0] > u0  # the first step rebuilt the rows (all pairs)
    g.close()


def test_fast_rigid_rings_langevin(oracle_mod):
    """Rigid Langevin (whose kick adds unit draws whatever the time step) with ring velocities that
    cover 0.6 cutoff per step: each drift flags the system as a far move, so the next rebuild
    searches all pairs; the trajectory matches the oracle, which rebuilds from scratch.  The rings
    start apart (rigid rings cannot give way: a start in contact at these speeds overlaps deeply
    within two steps and the forces blow up) and first touch in the third step, through pairs the
    far-move rebuilds found."""
    print("case far_moves rigid langevin", flush=True)
    st = ring_block(gap=4.0, seed=7)
    dt = 2e-3
    o = oracle_from_state(oracle_mod, st, potential="wca", interaction="vertexVertex", costs=(0, 0, 0, 1), dt=dt)
    g = gpu_from_state(st, potential="wca", interaction="vertexVertex", costs=(0, 0, 0, 1), dt=dt)
    size = 2 * st["rad"].mean()
    for obj in (o, g):
        obj.calc_neighbor_list(obj.set_displacement_cutoff(0.5, size))
        obj.calc_rigid_force_energy()
        obj.init_rigid_langevin(1e-2, 1.0)
    rng = np.random.default_rng(2)
    speed = 0.6 * 0.5 * size / dt
    ang = rng.uniform(0, 2 * np.pi, len(st["nvp"]))
    pv = (np.c_[np.cos(ang), np.sin(ang)] * speed).ravel()
    o.set("pvel", pv)
    g.set("pvel", pv)
    for chunk in (1, 2):
        o.rigid_loop(chunk)
        g.rigid_langevin(chunk)
        assert rel_err(g.get("pos"), o.get("pos"), scale=1.0) < 1e-11, chunk
        assert rel_err(g.get("pvel"), o.get("pvel")) < 1e-10, chunk
    assert g.update_count()[0] >= 2  # the list was rebuilt (all pairs) in these steps
    g.close()

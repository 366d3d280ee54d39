"""The integrator stages callable one by one (SimInterface, Simulator.h:47-52): driven in
Langevin::integrate's order (Simulator.cu:23-30) they reproduce the fused step, the noise of a
stage-wise step is the fused step's draw, injectKineticEnergy changes the velocities only, and
conserveMomentum removes the mean vertex velocity."""
import numpy as np
import pytest

from tests.conftest import rel_err
from tests.helpers import gpu_from_state, relaxed_state

pytestmark = pytest.mark.gpu


def two_sets(interaction):
    st = relaxed_state(16, 0.82, 17, final=("wca", interaction))  # no deep WCA overlaps
    out = []
    for _ in range(2):
        g = gpu_from_state(st, potential="wca", interaction=interaction, dt=2e-3, seed=5)
        cut = g.set_displacement_cutoff(0.5, 2 * float(np.mean(st["rad"])))
        g.calc_neighbor_list(cut)
        g.calc_force_energy()
        g.init_langevin(5e-3, 1.0)
        out.append(g)
    return out


@pytest.mark.parametrize("interaction", ["vertexVertex", "vertexSmooth"])
def test_langevin_stages_equal_fused_step(interaction):
    print(f"case langevin stages {interaction}", flush=True)
    fused, staged = two_sets(interaction)
    dt = 2e-3
    for step in range(30):
        fused.langevin(1)
        staged.update_velocity(0.5 * dt)
        staged.update_position(dt)
        staged.check_neighbors()
        staged.calc_force_energy()
        staged.update_thermal_vel()
        staged.update_velocity(0.5 * dt)
    for f in ("pos", "vel", "force"):
        assert rel_err(staged.get(f), fused.get(f)) < 1e-13, f
    assert int(staged.update_count()[0]) == int(fused.update_count()[0])
    # the fused loop continues from the stage-wise state with the same draws
    fused.langevin(10)
    staged.langevin(10)
    assert rel_err(staged.get("pos"), fused.get("pos")) < 1e-13


def test_inject_kinetic_energy_changes_velocities_only():
    g, _ = two_sets("vertexVertex")
    g.langevin(50)
    g.set_initial_positions()
    g.langevin(20)
    last, msd, pos = g.get("lastpos"), g.vertex_msd(), g.get("pos")
    v_before = g.get("vel")
    g.inject_kinetic_energy(2e-2)
    v = g.get("vel").reshape(-1, 2)
    assert not np.allclose(v, v_before.reshape(-1, 2))
    np.testing.assert_array_equal(g.get("lastpos"), last)
    np.testing.assert_array_equal(g.get("pos"), pos)
    assert g.vertex_msd()[0] == msd[0]
    # one draw per ring: every vertex of a ring has the ring's velocity
    off = np.concatenate([[0], np.cumsum(g.nvp)])
    for r in range(g.np):
        assert np.all(v[off[r]:off[r + 1]] == v[off[r]])
    g.conserve_momentum()
    assert np.abs(g.get("vel").reshape(-1, 2).mean(0)).max() < 1e-15

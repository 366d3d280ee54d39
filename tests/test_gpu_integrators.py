"""GPU against the oracle for the BAOAB, Brownian and driven-Brownian integrators and the
activeShape rest-length dynamics (Simulator.cu:112-203, DPM2D.cu:2037-2056), on the same
seeded Philox streams: trajectories agree to rounding."""
import numpy as np
import pytest

from tests.conftest import rel_err
from tests.helpers import gpu_from_state, oracle_from_state, relaxed_state

pytestmark = pytest.mark.gpu

CASES = [("harmonic", "vertexVertex"), ("wca", "vertexSmooth")]


def prepare(oracle_mod, st, pot, inter, noise, particle="passive", dt=0.002, skin=0.5):
    o = oracle_from_state(oracle_mod, st, potential=pot, interaction=inter, dt=dt, seed=17, particle=particle)
    g = gpu_from_state(st, potential=pot, interaction=inter, dt=dt, seed=17, particle=particle)
    o.set_noise_type(noise)
    g.set_noise_type(noise)
    size = 2 * st["rad"].mean()
    c1 = o.set_displacement_cutoff(skin, size)
    c2 = g.set_displacement_cutoff(skin, size)
    o.calc_neighbor_list(c1)
    g.calc_neighbor_list(c2)
    return o, g


@pytest.mark.parametrize("pot,inter", CASES)
@pytest.mark.parametrize("noise", ["baoab", "brownian", "drivenBrownian"])
def test_noise_type_trajectory(oracle_mod, noise, pot, inter):
    print(f"case noise_type {noise} {pot} {inter}", flush=True)
    st = dict(relaxed_state(16, 0.85, 23, final=(pot, inter)))
    # off the minimum, so that the noiseless overdamped scheme moves too
    st["pos"] = st["pos"] + np.random.default_rng(5).normal(0, 0.05 * st["rad"].mean(), st["pos"].shape)
    o, g = prepare(oracle_mod, st, pot, inter, noise)
    o.calc_force_energy()
    g.calc_force_energy()
    o.init_langevin(1e-2, 2.0)
    g.init_langevin(1e-2, 2.0)
    x0 = o.get("pos").copy()
    for chunk in (1, 3, 40):  # the 40-step chunk runs as a captured graph
        o.langevin(chunk)
        g.langevin(chunk)
        assert rel_err(g.get("pos"), o.get("pos"), scale=1.0) < 1e-11
        assert rel_err(g.get("vel"), o.get("vel")) < 1e-9
        assert rel_err(g.get("force"), o.get("force")) < 1e-9
    moved = np.abs(o.get("pos") - x0).max()
    assert moved > 1e-3  # the comparison is over a trajectory that went somewhere
    assert int(g.update_count()[0]) == int(o.update_count())
    g.close()


def test_active_shape_rest_lengths(oracle_mod):
    print("case active_shape", flush=True)
    st = relaxed_state(16, 0.85, 29, final=("wca", "vertexSmooth"))
    o, g = prepare(oracle_mod, st, "wca", "vertexSmooth", "langevin", particle="activeShape")
    for s in (o, g):
        s.set_active_tension(0.5, 1e-4)
    l00 = o.get("l0").copy()
    o.calc_force_energy()
    g.calc_force_energy()
    assert rel_err(g.get("l0"), o.get("l0"), scale=1.0) < 1e-14
    # relative to the net force of a relaxed packing, far below the size of its terms
    assert rel_err(g.get("force"), o.get("force")) < 1e-10
    o.init_langevin(1e-3, 1.0)
    g.init_langevin(1e-3, 1.0)
    for chunk in (2, 30):
        o.langevin(chunk)
        g.langevin(chunk)
        assert rel_err(g.get("l0"), o.get("l0"), scale=1.0) < 1e-12
        assert rel_err(g.get("pos"), o.get("pos"), scale=1.0) < 1e-11
    assert np.abs(o.get("l0") - l00).max() > 1e-5
    g.close()


def test_active_shape_needs_tension():
    print("case active_shape_state", flush=True)
    from cudaring_amd.dpm import DPMError
    st = relaxed_state(9, 0.8, 31, final=("harmonic", "vertexVertex"))
    g = gpu_from_state(st, particle="activeShape")
    g.calc_neighbor_list(g.set_displacement_cutoff(0.5, 2 * st["rad"].mean()))
    with pytest.raises(DPMError):
        g.calc_force_energy()
    g.init_langevin(1e-3, 1.0)
    with pytest.raises(DPMError):
        g.langevin(5)
    g.close()

"""GPU against the oracle for rigid rings (calcRigidForceEnergy DPM2D.cu:2128-2142; RigidNVE /
RigidLangevin Simulator.cu:241-371; rigid FIRE FIRE.cu:173-298): ring forces, torques and
energies, NVE and Langevin trajectories on the same Philox stream, FIRE iterations, and two
systems batched in one set."""
import numpy as np
import pytest

from tests.conftest import rel_err
from tests.helpers import gpu_from_state, oracle_from_state
from tests.test_oracle_rigid import ring_block

pytestmark = pytest.mark.gpu


def pair(oracle_mod, st, pot="harmonic", inter="vertexVertex", dt=2e-3, seed=5):
    o = oracle_from_state(oracle_mod, st, potential=pot, interaction=inter, costs=(0, 0, 0, 1), dt=dt, seed=seed)
    g = gpu_from_state(st, potential=pot, interaction=inter, costs=(0, 0, 0, 1), dt=dt, seed=seed)
    size = 2 * st["rad"].mean()
    o.calc_neighbor_list(o.set_displacement_cutoff(0.5, size))
    g.calc_neighbor_list(g.set_displacement_cutoff(0.5, size))
    o.calc_rigid_force_energy()
    g.calc_rigid_force_energy()
    return o, g


def compare_rings(o, g, tol=1e-12):
    assert rel_err(g.get("pforce"), o.get("pforce")) < tol
    assert rel_err(g.get("ptorque"), o.get("ptorque")) < tol
    assert rel_err(g.get("force"), o.get("force")) < tol
    assert rel_err(g.get("vtorque"), o.get("vtorque")) < tol
    assert rel_err(g.get("ppos"), o.get("ppos"), scale=1.0) < tol


@pytest.mark.parametrize("pot,inter", [("harmonic", "vertexVertex"), ("wca", "vertexVertex"),
                                       ("harmonic", "vertexSmooth"), ("wca", "vertexSmooth"),
                                       ("lj", "vertexVertex"), ("adhesive", "vertexVertex"),
                                       ("lj", "vertexSmooth"), ("adhesive", "vertexSmooth")])
def test_rigid_force_parity(oracle_mod, pot, inter):
    print(f"case rigid_force {pot} {inter}", flush=True)
    st = ring_block(gap=0.6 if pot == "harmonic" else 1.9)
    o, g = pair(oracle_mod, st, pot, inter)
    assert np.abs(o.get("pforce")).max() > 1e-3
    compare_rings(o, g)
    assert abs(g.potential_energy()[0] - o.potential_energy()) < 1e-12 * abs(o.potential_energy())
    assert rel_err(g.get("penergy"), o.get("penergy")) < 1e-12
    g.close()


@pytest.mark.parametrize("inter", ["vertexVertex", "vertexSmooth"])
def test_rigid_nve_trajectory(oracle_mod, inter):
    print(f"case rigid_nve {inter}", flush=True)
    st = ring_block(gap=0.8)
    o, g = pair(oracle_mod, st, "harmonic", inter)
    o.init_rigid_nve(1e-3)
    g.init_rigid_nve(1e-3)
    assert rel_err(g.get("pvel"), o.get("pvel")) < 1e-13
    w = np.linspace(-0.2, 0.2, len(st["nvp"]))
    for s in (o, g):
        s.set("pangvel", w)
        s.calc_rigid_force_energy()
    for chunk in (1, 3, 40):
        o.rigid_loop(chunk)
        g.rigid_nve(chunk)
        assert rel_err(g.get("pos"), o.get("pos"), scale=1.0) < 1e-11, chunk
        assert rel_err(g.get("pvel"), o.get("pvel")) < 1e-9, chunk
        assert rel_err(g.get("pangvel"), o.get("pangvel")) < 1e-9, chunk
        assert rel_err(g.get("pangle"), o.get("pangle"), scale=1.0) < 1e-11, chunk
    compare_rings(o, g, tol=1e-9)
    g.close()


def test_rigid_langevin_trajectory(oracle_mod):
    print("case rigid_langevin", flush=True)
    st = ring_block(gap=0.8)
    o, g = pair(oracle_mod, st, "harmonic", "vertexVertex", dt=1e-3)
    o.init_rigid_langevin(1e-2, 1.0)
    g.init_rigid_langevin(1e-2, 1.0)
    o.calc_rigid_force_energy()
    g.calc_rigid_force_energy()
    for chunk in (2, 40):
        o.rigid_loop(chunk)
        g.rigid_langevin(chunk)
        assert rel_err(g.get("pos"), o.get("pos"), scale=1.0) < 1e-11, chunk
        assert rel_err(g.get("pvel"), o.get("pvel")) < 1e-10, chunk
        assert rel_err(g.get("pangvel"), o.get("pangvel")) < 1e-9, chunk
    assert list(g.update_count()) == [o.update_count()]
    st_rad = st["R"]
    for s in (o, g):
        s.set("prad", st_rad)
    assert abs(g.rigid_kinetic_energy()[0] - o.rigid_kinetic_energy()) < 1e-10 * o.rigid_kinetic_energy()
    g.close()


def test_rigid_fire_parity(oracle_mod):
    print("case rigid_fire", flush=True)
    st = ring_block(gap=0.5)
    o, g = pair(oracle_mod, st, "harmonic", "vertexVertex")
    # the engine's list build records its positions as the list positions, as the drivers'
    # resetLastPositions after calcNeighborList does (initRigidFIRE does not reset them)
    o.reset_last_positions()
    dt = 1e-2 * 2 * np.sqrt(st["a0"].mean() / np.pi)
    params = [0.2, 0.5, 1.1, 0.99, dt, 10 * dt, 0.2]
    o.init_rigid_fire(params, 20, 0)
    g.init_rigid_fire(params, 20, 0)
    for chunk in (1, 2, 27):
        o.rigid_fire(chunk)
        g.rigid_fire(chunk)
        assert rel_err(g.get("pos"), o.get("pos"), scale=1.0) < 1e-10, chunk
        assert rel_err(g.get("pvel"), o.get("pvel")) < 1e-8, chunk
    fs = g.rigid_fire_state()
    assert fs["iters"][0] == 30
    assert list(g.update_count()) == [o.update_count()]
    assert abs(g.rigid_max_unbalanced_force()[0] - o.rigid_max_unbalanced_force()) < 1e-8 * o.rigid_max_unbalanced_force()
    # to a minimum: both relax by orders of magnitude
    f0 = o.rigid_max_unbalanced_force()
    o.rigid_fire(2000)
    g.rigid_fire(2000)
    assert g.rigid_max_unbalanced_force()[0] < 1e-3 * f0
    assert abs(g.potential_energy()[0] - o.potential_energy()) < 1e-6 * abs(o.potential_energy()) + 1e-14
    print("rebuilds: GPU", int(g.update_count()[0]), "oracle", o.update_count(), flush=True)
    g.close()


def test_rigid_batched_systems(oracle_mod):
    """two systems in one set: each follows its own oracle (system id = its index)"""
    print("case rigid_batched", flush=True)
    from cudaring_amd import DPMSet
    sts = [ring_block(gap=0.8, seed=2), ring_block(gap=0.7, seed=4)]
    nvp = np.concatenate([s["nvp"] for s in sts])
    g = DPMSet(nvp, rings_per_system=[len(s["nvp"]) for s in sts])
    cat = {k: np.concatenate([s[k] for s in sts]) for k in ("pos", "rad", "l0", "theta0", "a0")}
    cat["box"] = np.concatenate([s["box"] for s in sts])
    g.load_state(cat)
    g.set_types("harmonic", "vertexVertex", "harmonic", "passive")
    g.set_energy_costs(0, 0, 0, 1)
    g.set_dt(2e-3)
    g.set_seed(5)
    size = 2 * np.mean([s["rad"].mean() for s in sts])
    g.calc_neighbor_list(g.set_displacement_cutoff(0.5, size))
    g.calc_rigid_force_energy()
    g.init_rigid_langevin(1e-2, 1.0)
    g.calc_rigid_force_energy()
    g.rigid_langevin(30)
    pos = g.get("pos").reshape(-1, 2)
    k = 0
    for sid, st in enumerate(sts):
        o = oracle_from_state(oracle_mod, st, costs=(0, 0, 0, 1), dt=2e-3, seed=5, system_id=sid)
        o.calc_neighbor_list(o.set_displacement_cutoff(0.5, size))
        o.calc_rigid_force_energy()
        o.init_rigid_langevin(1e-2, 1.0)
        o.calc_rigid_force_energy()
        o.rigid_loop(30)
        n = int(st["nvp"].sum())
        assert rel_err(pos[k:k + n].ravel(), o.get("pos"), scale=1.0) < 1e-11, sid
        k += n
    g.close()

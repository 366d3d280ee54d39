"""GPU parity of the remaining hot-path rows: the all-to-all interaction (cudaKernel.cuh:725-757),
NVERescale (Simulator.cu:215-238), adjustTemperature / adjustKineticEnergy (DPM2D.cu:1161-1200)
and getTotalForceMagnitude (DPM2D.cu:986-995), against the oracle on identical states."""
import numpy as np
import pytest

from cudaring_amd import DPMSet, packing
from tests.conftest import rel_err
from tests.helpers import gpu_from_state, oracle_from_state

pytestmark = pytest.mark.gpu

TOL_F = 1e-12


def pair(oracle_mod, st, pot="harmonic", inter="vertexVertex", **kw):
    o = oracle_from_state(oracle_mod, st, potential=pot, interaction=inter, **kw)
    g = gpu_from_state(st, potential=pot, interaction=inter, **kw)
    size = 2 * st["rad"].mean()
    c = o.set_displacement_cutoff(0.3, size)
    assert g.set_displacement_cutoff(0.3, size) == c
    o.calc_neighbor_list(c)
    g.calc_neighbor_list(c)
    return o, g


@pytest.mark.parametrize("pot", ["harmonic", "wca", "lj", "adhesive"])
def test_all_to_all_interaction(oracle_mod, pot):
    print(f"case all_to_all pot={pot}", flush=True)
    st = packing.random_system(10, 0.8, 303)
    o, g = pair(oracle_mod, st, pot, "all")
    o.calc_force_energy()
    g.calc_force_energy()
    assert rel_err(g.get("force"), o.get("force")) < TOL_F
    assert rel_err(g.get("energy"), o.get("energy")) < TOL_F
    assert abs(g.potential_energy()[0] - o.potential_energy()) <= TOL_F * abs(o.potential_energy())
    # with a complete list the listed vertex-vertex interaction gives the same forces
    _, g2 = pair(oracle_mod, st, pot, "vertexVertex")
    g2.calc_neighbor_list(g2.set_displacement_cutoff(50.0, 2 * st["rad"].mean()))
    g2.calc_force_energy()
    assert rel_err(g2.get("force"), g.get("force")) < TOL_F


def test_nve_rescale_window(oracle_mod):
    print("case nve_rescale_window", flush=True)
    st = packing.lattice_system(16, 0.8, 19)
    o, g = pair(oracle_mod, st, dt=0.005, costs=(1, 1, 0.01, 1))
    o.calc_force_energy(); g.calc_force_energy()
    o.init_nve_rescale(2e-3); g.init_nve_rescale(2e-3)
    assert rel_err(g.get("vel"), o.get("vel")) < 1e-14
    assert abs(g.temperature()[0] / 2e-3 - 1) < 1e-13
    for chunk in (1, 4, 25):
        o.nve_rescale(chunk); g.nve_rescale(chunk)
        assert rel_err(g.get("pos"), o.get("pos"), scale=1.0) < 1e-11
        assert rel_err(g.get("vel"), o.get("vel")) < 1e-9
    assert abs(g.temperature()[0] - o.temperature()) < 1e-9 * o.temperature()


def batched(oracle_mod, sizes=(9, 14), seed=40):
    systems = [packing.lattice_system(n, 0.78, seed + n) for n in sizes]
    cat = packing.concat_systems(systems)
    g = DPMSet(cat["nvp"], rings_per_system=cat["rings_per_system"])
    g.load_state(cat)
    g.set_types("harmonic", "vertexVertex")
    g.set_energy_costs(1, 1, 0.01, 1)
    g.set_dt(0.005)
    g.set_seed(5)
    c = g.set_displacement_cutoff(0.3, 2 * cat["rad"].mean())
    g.calc_neighbor_list(c)
    g.calc_force_energy()
    orcs = []
    for k, st in enumerate(systems):
        o = oracle_from_state(oracle_mod, st, costs=(1, 1, 0.01, 1), dt=0.005, seed=5, system_id=k)
        o.calc_neighbor_list(o.set_displacement_cutoff(0.3, 2 * cat["rad"].mean()))
        o.calc_force_energy()
        orcs.append(o)
    return g, orcs, systems


def test_total_force_magnitude_and_max_force(oracle_mod):
    print("case total_force_magnitude", flush=True)
    g, orcs, _ = batched(oracle_mod)
    tf = g.total_force_magnitude()
    mf = g.max_unbalanced_force()
    for k, o in enumerate(orcs):
        assert abs(tf[k] - o.total_force_magnitude()) <= 1e-13 * o.total_force_magnitude()
        assert abs(mf[k] - o.max_unbalanced_force()) <= 1e-15 * o.max_unbalanced_force()


def test_adjust_temperature(oracle_mod):
    print("case adjust_temperature", flush=True)
    g, orcs, _ = batched(oracle_mod)
    g.init_nve(3e-2)
    for k, o in enumerate(orcs):
        o.init_nve(3e-2)
    # the set's systems draw from their own streams: velocities equal the oracle's per system
    v = g.get("vel").reshape(-1, 2)
    off = 0
    for o in orcs:
        assert rel_err(v[off:off + o.nv].ravel(), o.get("vel")) < 1e-15
        off += o.nv
    g.adjust_temperature(7e-3)
    for o in orcs:
        o.adjust_temperature(7e-3)
    T = g.temperature()
    v = g.get("vel").reshape(-1, 2)
    off = 0
    for k, o in enumerate(orcs):
        assert abs(T[k] / 7e-3 - 1) < 1e-13
        assert rel_err(v[off:off + o.nv].ravel(), o.get("vel")) < 1e-14
        off += o.nv


def test_adjust_kinetic_energy(oracle_mod):
    print("case adjust_kinetic_energy", flush=True)
    g, orcs, _ = batched(oracle_mod)
    g.init_nve(3e-2)
    for o in orcs:
        o.init_nve(3e-2)
    U = g.potential_energy()
    K = g.kinetic_energy()
    # system 0 can absorb the change, system 1 cannot (K < dE): it is left alone
    target = np.array([U[0] + 0.45 * K[0], U[1] - 0.5])
    v_before = g.get("vel")
    sc = g.adjust_kinetic_energy(target)
    so = [o.adjust_kinetic_energy(t) for o, t in zip(orcs, target)]
    assert sc[0] == pytest.approx(so[0], rel=1e-13) and so[0] == pytest.approx(np.sqrt(0.45), rel=1e-12)
    assert sc[1] == 0.0 and so[1] == 0.0
    v = g.get("vel").reshape(-1, 2)
    n0 = orcs[0].nv
    assert rel_err(v[:n0].ravel(), orcs[0].get("vel")) < 1e-14
    assert np.array_equal(v[n0:].ravel(), v_before.reshape(-1, 2)[n0:].ravel())
    E = g.potential_energy() + g.kinetic_energy()
    assert abs(E[0] - target[0]) < 1e-12 * (abs(U[0]) + K[0])

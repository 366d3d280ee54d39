"""GPU against the oracle for the observables (contacts, hexatic order, MSD, ISF, per-particle
stress, stress tensor and pressure) on the same configurations."""
import numpy as np
import pytest

from cudaring_amd import packing
from tests.helpers import gpu_from_state, oracle_from_state
from tests.test_oracle_observables import lattice_rings

pytestmark = pytest.mark.gpu


def pair(oracle_mod, st, skin=0.3, **kw):
    o = oracle_from_state(oracle_mod, st, **kw)
    g = gpu_from_state(st, **kw)
    size = 2 * st["rad"].mean()
    c1 = o.set_displacement_cutoff(skin, size)
    c2 = g.set_displacement_cutoff(skin, size)
    o.calc_neighbor_list(c1)
    g.calc_neighbor_list(c2)
    return o, g


def test_contacts_and_hexatic(oracle_mod):
    print("case contacts", flush=True)
    for st in (packing.lattice_system(20, 0.8, 5, perturb=0.2), lattice_rings("triangular"), lattice_rings("square")):
        o, g = pair(oracle_mod, st)
        lo, co = o.contacts()
        lg, cg = g.contacts()
        np.testing.assert_array_equal(cg, co)
        for p in range(len(co)):
            np.testing.assert_array_equal(lg[p, :cg[p]], lo[p, :co[p]])  # same order of first appearance
        assert abs(g.hexatic_order()[0] - o.hexatic_order()) < 1e-13
        g.close()


def test_geometric_contacts(oracle_mod):
    print("case geometric contacts", flush=True)
    for st in (packing.lattice_system(20, 0.8, 5, perturb=0.2), lattice_rings("triangular")):
        o, g = pair(oracle_mod, st)
        lo, co = o.geometric_contacts(1.0)
        lg, cg = g.geometric_contacts()
        assert lg.shape[1] == 8
        np.testing.assert_array_equal(cg, co)
        for p in range(len(co)):
            np.testing.assert_array_equal(lg[p, :cg[p]], lo[p, :co[p]])
        assert abs(g.hexatic_order()[0] - o.hexatic_order()) < 1e-13  # psi6 from these contacts
        g.close()


def test_msd_isf_after_dynamics(oracle_mod):
    print("case msd_isf", flush=True)
    st = packing.lattice_system(16, 0.8, 9, perturb=0.1)
    o, g = pair(oracle_mod, st, dt=0.005, seed=4)
    o.calc_force_energy(); g.calc_force_energy()
    o.init_langevin(1e-2, 1.0); g.init_langevin(1e-2, 1.0)
    o.langevin(50); g.langevin(50)
    for name in ("vertex_msd", "particle_msd", "vertex_isf"):
        a, b = getattr(g, name)()[0], getattr(o, name)()
        assert abs(a - b) <= 1e-10 * abs(b), name
    assert abs(g.particle_isf(2.5)[0] - o.particle_isf(2.5)) < 1e-10
    g.close()


def test_stress_and_pressure(oracle_mod):
    print("case stress", flush=True)
    for st in (lattice_rings("square", m=4, gap_vertex_diameters=0.25, rad_scale=0.3),
               packing.lattice_system(16, 0.9, 13, perturb=0.2)):
        o, g = pair(oracle_mod, st, costs=(1, 1, 0.1, 1))
        o.calc_force_energy(); g.calc_force_energy()
        po, pg = o.per_particle_stress(), g.per_particle_stress()
        scale = np.abs(po).max()
        assert scale > 0
        assert np.abs(pg - po).max() < 1e-12 * scale
        np.testing.assert_allclose(g.stress_tensor()[0], o.stress_tensor(), rtol=1e-12, atol=1e-14 * scale)
        assert abs(g.pressure()[0] - o.pressure()) <= 1e-12 * abs(o.pressure()) + 1e-18
        g.close()

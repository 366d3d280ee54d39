"""Error behaviour of the C ABI (each call returns a DPM_ERR_* code and the binding raises):
ring sizes outside 3..64 vertices, empty systems, field size mismatches, a force call before
the neighbour list exists, and rigid rings outside the periodic passive vertexVertex /
vertexSmooth case.  A failed call leaves the set usable."""
import numpy as np
import pytest

from tests.conftest import rel_err
from tests.helpers import gpu_from_state, oracle_from_state, small_state

pytestmark = pytest.mark.gpu


def test_create_rejects_ring_sizes():
    print("case create_ring_sizes", flush=True)
    from cudaring_amd import DPMSet
    from cudaring_amd.dpm import DPMError
    DPMSet(np.array([16, 65])).close()  # a ring of more than 64 vertices is cut into slices
    with pytest.raises(DPMError, match="unsupported"):
        DPMSet(np.array([2, 16]))
    with pytest.raises(DPMError, match="bad argument"):
        DPMSet(np.array([16, 16]), rings_per_system=[2, 0])


def test_calls_fail_cleanly_and_set_recovers(oracle_mod):
    print("case call_errors", flush=True)
    from cudaring_amd.dpm import DPMError
    st = small_state(8, 0.6, 11)
    g = gpu_from_state(st)
    with pytest.raises(DPMError, match="bad argument"):
        g.set("pos", np.zeros(3))
    with pytest.raises(DPMError, match="bad state"):
        g.calc_force_energy()
    g.set_geometry("fixedBox")
    g.calc_neighbor_list(g.set_displacement_cutoff(0.5, 2 * st["rad"].mean()))
    with pytest.raises(DPMError, match="unsupported"):
        g.calc_rigid_force_energy()
    g.set_geometry("normal")
    g.set_types("harmonic", "all")
    with pytest.raises(DPMError, match="unsupported"):
        g.calc_rigid_force_energy()
    # after the failures the set still computes the oracle's forces
    g.set_types("harmonic", "vertexVertex")
    o = oracle_from_state(oracle_mod, st)
    size = 2 * st["rad"].mean()
    o.calc_neighbor_list(o.set_displacement_cutoff(0.5, size))
    g.calc_neighbor_list(g.set_displacement_cutoff(0.5, size))
    o.calc_force_energy()
    g.calc_force_energy()
    assert rel_err(g.get("force"), o.get("force")) < 1e-12
    g.close()

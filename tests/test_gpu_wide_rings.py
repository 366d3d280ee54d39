"""Rings of more than 64 vertices (setPolySizeDistribution, DPM2D.cu:1545-1588, sets no upper
bound): each is cut into tiles of 64-vertex slices whose ring sums come from an extra launch,
and whose ring neighbours and (vertex-segment) neighbour shares across a slice boundary are read
from memory.  Checked against the oracle, which has no ring-size limit: forces, energies and
shapes, Langevin and FIRE windows through in-loop rebuilds, and the paths such sets refuse.

The rings below have equal preferred areas and 16 to 200 vertices (so the vertex radii differ),
including 64 (one full tile), 65 (a slice of one vertex), 128 (two full slices) and 200."""
import numpy as np
import pytest

from cudaring_amd import DPMError, packing
from tests.conftest import rel_err
from tests.helpers import gpu_from_state, oracle_from_state

pytestmark = pytest.mark.gpu

NVP = np.array([65, 128, 16, 200, 64, 97, 22, 150, 130])


def wide_state(gap_scale, seed=3, calA0=1.1):
    """3 x 3 rings of unit preferred area on a square lattice whose spacing leaves a gap of
    gap_scale x (mean vertex diameter) between the regular polygons (negative: overlapping)."""
    rng = np.random.default_rng(seed)
    n = NVP.astype(np.float64)
    a0 = np.ones(len(NVP))
    l0r = 2.0 * np.sqrt(np.pi * calA0 * n * np.tan(np.pi / n) / np.pi * a0) / n
    l0 = np.repeat(l0r, NVP)
    theta0 = np.repeat(2.0 * np.pi / n, NVP)
    rad = 0.5 * l0
    R = np.sqrt(2.0 / (n * np.sin(2.0 * np.pi / n))).max()
    a = 2 * R + gap_scale * 2 * rad.mean()
    idx = rng.permutation(9)
    centres = np.stack([(idx % 3 + 0.5) * a, (idx // 3 + 0.5) * a], axis=1)
    pos = packing.place_rings(NVP, a0, centres, rng=rng, jitter=0.05, l0=l0)
    return {"nvp": NVP.copy(), "box": np.array([3 * a, 3 * a]), "pos": pos.ravel(), "rad": rad, "l0": l0,
            "theta0": theta0, "a0": a0}


def pair(oracle_mod, st, pot, inter, skin=0.5, **kw):
    o = oracle_from_state(oracle_mod, st, potential=pot, interaction=inter, **kw)
    g = gpu_from_state(st, potential=pot, interaction=inter, **kw)
    size = 2 * st["rad"].mean()
    c = o.set_displacement_cutoff(skin, size)
    assert g.set_displacement_cutoff(skin, size) == c
    o.calc_neighbor_list(c)
    g.calc_neighbor_list(c)
    return o, g


@pytest.mark.parametrize("pot,inter", [("harmonic", "vertexVertex"), ("harmonic", "vertexSmooth"),
                                       ("wca", "vertexSmooth")])
def test_wide_ring_forces(oracle_mod, pot, inter):
    print(f"case wide forces {pot} {inter}", flush=True)
    st = wide_state(-0.3 if pot == "harmonic" else 0.05)
    o, g = pair(oracle_mod, st, pot, inter)
    assert [sorted(r) for r in g.neighbors()[0]] == [sorted(r) for r in o.neighbors()[0]]
    o.calc_force_energy()
    g.calc_force_energy()
    f = o.get("force")
    assert np.abs(f).max() > 0
    assert rel_err(g.get("force"), f) < 1e-12
    assert rel_err(g.get("energy"), o.get("energy")) < 1e-12
    if inter == "vertexSmooth":
        assert rel_err(g.get("penergy"), o.get("penergy")) < 1e-12
    assert abs(g.potential_energy()[0] - o.potential_energy()) <= 1e-12 * abs(o.potential_energy())
    for name in ("area", "perimeter", "theta", "length", "ppos"):
        assert rel_err(g.get(name), o.get(name)) < 1e-13, name


@pytest.mark.parametrize("inter", ["vertexVertex", "vertexSmooth"])
def test_wide_ring_langevin_window(oracle_mod, inter):
    """120 Langevin steps at a skin small enough that the loop rebuilds rows inside it."""
    print(f"case wide langevin {inter}", flush=True)
    st = wide_state(-0.3)
    o, g = pair(oracle_mod, st, "harmonic", inter, skin=0.05, dt=2e-3)
    for s in (o, g):
        s.calc_force_energy()
        s.init_langevin(2e-2, 1.0)
    g.set("vel", o.get("vel"))
    g.langevin(120)
    o.langevin(120)
    assert int(o.update_count()) > 0
    assert int(g.update_count()[0]) == int(o.update_count())
    assert rel_err(g.get("pos"), o.get("pos"), 1.0) < 1e-11
    assert rel_err(g.get("vel"), o.get("vel")) < 1e-9
    assert [sorted(r) for r in g.neighbors()[0]] == [sorted(r) for r in o.neighbors()[0]]


def test_wide_ring_fire(oracle_mod):
    """FIRE on a wide-ring set runs the multi-launch kernels (the one-workgroup-per-system kernel
    refuses such sets) and follows the oracle's iterations."""
    print("case wide fire", flush=True)
    st = wide_state(-0.3)
    o, g = pair(oracle_mod, st, "harmonic", "vertexSmooth")
    p7 = (0.2, 0.5, 1.1, 0.99, 1e-3, 1e-2, 0.2)
    for s in (o, g):
        s.calc_force_energy()
        s.init_fire(p7, 20, 0)
    with pytest.raises(DPMError):
        g.set_fire_kernel(1, 4)
    g.fire(200)
    o.fire(200)
    assert rel_err(g.get("pos"), o.get("pos"), 1.0) < 1e-10
    fs, ofs = g.fire_state(), o.fire_state()
    assert fs["numStep"][0] == ofs["numStep"] and abs(fs["dt"][0] - ofs["dt"]) <= 1e-14 * ofs["dt"]


def test_wide_ring_refused_paths():
    st = wide_state(-0.3)
    g = gpu_from_state(st)
    cut = g.set_displacement_cutoff(0.5, 2 * st["rad"].mean())
    g.calc_neighbor_list(cut)
    g.calc_force_energy()
    with pytest.raises(DPMError, match="64 vertices"):
        g.set_types("harmonic", "vertexVertex", "harmonic", "active")
    with pytest.raises(DPMError, match="64 vertices"):
        g.init_rigid_nve(1e-3)
    with pytest.raises(DPMError, match="64 vertices"):
        g.nve_forward(10, 5)

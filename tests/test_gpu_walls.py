"""GPU against the oracle with walls (fixedBox, fixedSides, roundBox; cudaKernel.cuh:76-215,
1221-1410; DPM2D.cu:2090-2103): forces, energies and wall forces at a configuration whose outer
rings touch the walls, and a Langevin trajectory on the same Philox stream."""
import numpy as np
import pytest

from cudaring_amd import packing
from tests.conftest import rel_err
from tests.helpers import gpu_from_state, oracle_from_state

pytestmark = pytest.mark.gpu


def walled_state(geometry, seed=3):
    """4 x 4 rings on a square grid; the largest rings' outer vertices sit about one vertex radius
    from the walls (inside the WCA cut, 2^(1/6) radii)."""
    rng = np.random.default_rng(seed)
    nvp = packing.ring_sizes(16, rng=rng)
    a0, l0, th0, rad = packing.ring_shape_params(nvp)
    R = np.sqrt(2.0 * a0 / (nvp * np.sin(2.0 * np.pi / nvp)))
    r = rad.max()
    step = 2 * R.max() + 2.3 * r  # facing vertices 0.3 radii apart
    margin = R.max() + 1.0 * r
    g = (np.arange(4) * step) + margin
    centers = np.array([(x, y) for y in g for x in g]) + rng.uniform(-0.1, 0.1, (16, 2)) * r
    L = 2 * margin + 3 * step
    if geometry == "roundBox":
        centers = centers - L / 2  # the round box is centred on the origin
    pos = packing.place_rings(nvp, a0, centers)
    st = {"nvp": nvp, "box": np.array([L, L]), "pos": pos.reshape(-1), "rad": rad, "l0": l0, "theta0": th0,
          "a0": a0}
    # round box: the farthest vertex one radius inside the wall
    box_radius = np.sqrt((pos ** 2).sum(1)).max() + 1.0 * r if geometry == "roundBox" else None
    return st, box_radius


def prepare(oracle_mod, geometry, pot="wca", inter="vertexSmooth"):
    st, R = walled_state(geometry)
    o = oracle_from_state(oracle_mod, st, potential=pot, interaction=inter, costs=(1, 1, 0.1, 1), dt=0.002, seed=9)
    g = gpu_from_state(st, potential=pot, interaction=inter, costs=(1, 1, 0.1, 1), dt=0.002, seed=9)
    for s in (o, g):
        s.set_geometry(geometry, ew=2.0, box_radius=R)
    size = 2 * st["rad"].mean()
    c1 = o.set_displacement_cutoff(0.5, size)
    c2 = g.set_displacement_cutoff(0.5, size)
    o.calc_neighbor_list(c1)
    g.calc_neighbor_list(c2)
    return o, g, st


@pytest.mark.parametrize("inter", ["vertexVertex", "vertexSmooth"])
@pytest.mark.parametrize("geometry", ["fixedBox", "fixedSides", "roundBox"])
def test_wall_force_parity(oracle_mod, geometry, inter):
    print(f"case wall_force {geometry} {inter}", flush=True)
    o, g, st = prepare(oracle_mod, geometry, inter=inter)
    o.calc_force_energy()
    g.calc_force_energy()
    wf = o.get("wallforce")
    assert np.abs(wf).max() > 1e-3  # some vertices press on a wall
    scale = np.abs(wf).max()
    assert np.abs(g.get("wallforce") - wf).max() < 1e-12 * scale
    assert rel_err(g.get("force"), o.get("force")) < 1e-12
    assert abs(g.potential_energy()[0] - o.potential_energy()) < 1e-12 * abs(o.potential_energy())
    g.close()


@pytest.mark.parametrize("geometry", ["fixedBox", "fixedSides", "roundBox"])
def test_wall_langevin_trajectory(oracle_mod, geometry):
    print(f"case wall_langevin {geometry}", flush=True)
    o, g, st = prepare(oracle_mod, geometry)
    o.calc_force_energy()
    g.calc_force_energy()
    o.init_langevin(1e-2, 1.0)
    g.init_langevin(1e-2, 1.0)
    for chunk in (2, 40):
        o.langevin(chunk)
        g.langevin(chunk)
        assert rel_err(g.get("pos"), o.get("pos"), scale=1.0) < 1e-11
        assert rel_err(g.get("vel"), o.get("vel")) < 1e-9
    g.close()


@pytest.mark.parametrize("geometry", ["normal", "fixedBox", "fixedSides", "roundBox"])
def test_scale_box_parity(oracle_mod, geometry):
    """scaleBox (DPM2D.cu:1693-1719) then a neighbour rebuild and forces, as compressNVE.cpp does
    between compression steps"""
    print(f"case scale_box {geometry}", flush=True)
    o, g, st = prepare(oracle_mod, geometry)
    o.calc_force_energy()
    g.calc_force_energy()
    for s in (1.01, 1.02):
        o.scale_box(s)
        g.scale_box(s)
        if geometry == "roundBox":
            assert abs(g.box_radius() - o.box_radius()) == 0.0
        else:
            assert np.array_equal(g.get("box"), o.get("box"))
        assert rel_err(g.get("ppos"), o.get("ppos"), scale=1.0) < 1e-13
        assert rel_err(g.get("pos"), o.get("pos"), scale=1.0) < 1e-13
        size = 2 * st["rad"].mean()
        o.calc_neighbor_list(o.set_displacement_cutoff(0.5, size))
        g.calc_neighbor_list(g.set_displacement_cutoff(0.5, size))
        o.calc_force_energy()
        g.calc_force_energy()
        assert rel_err(g.get("force"), o.get("force")) < 1e-12
    g.close()


def test_scale_box_batched_systems(oracle_mod):
    """two periodic systems in one set, each shrunk by its own scale (dpm_scale_box takes one
    scale per system), against one oracle each"""
    print("case scale_box batched", flush=True)
    from cudaring_amd import DPMSet
    from tests.test_oracle_rigid import ring_block
    sts = [ring_block(gap=1.0, seed=2), ring_block(gap=1.4, seed=6)]
    nvp = np.concatenate([s["nvp"] for s in sts])
    g = DPMSet(nvp, rings_per_system=[len(s["nvp"]) for s in sts])
    cat = {k: np.concatenate([s[k] for s in sts]) for k in ("pos", "rad", "l0", "theta0", "a0", "box")}
    g.load_state(cat)
    g.set_types("wca", "vertexSmooth", "harmonic", "passive")
    g.set_energy_costs(1, 1, 0.1, 1)
    size = 2 * np.mean([s["rad"].mean() for s in sts])
    g.calc_neighbor_list(g.set_displacement_cutoff(0.5, size))
    g.calc_force_energy()
    scales = [1.02, 1.05]
    g.scale_box(scales)
    g.calc_neighbor_list(g.set_displacement_cutoff(0.5, size))
    g.calc_force_energy()
    pos, force, box = g.get("pos").reshape(-1, 2), g.get("force").reshape(-1, 2), g.get("box").reshape(-1, 2)
    k = 0
    for sid, st in enumerate(sts):
        o = oracle_from_state(oracle_mod, st, potential="wca", interaction="vertexSmooth", costs=(1, 1, 0.1, 1),
                              system_id=sid)
        o.calc_neighbor_list(o.set_displacement_cutoff(0.5, size))
        o.calc_force_energy()
        o.scale_box(scales[sid])
        o.calc_neighbor_list(o.set_displacement_cutoff(0.5, size))
        o.calc_force_energy()
        n = int(st["nvp"].sum())
        assert np.array_equal(box[sid], o.get("box"))
        assert rel_err(pos[k:k + n].ravel(), o.get("pos"), scale=1.0) < 1e-13, sid
        assert rel_err(force[k:k + n].ravel(), o.get("force")) < 1e-12, sid
        k += n
    g.close()

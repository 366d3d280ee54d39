"""Pins the oracle's wall geometries (fixedBox, fixedSides, roundBox; cudaKernel.cuh:76-215,
1221-1410; DPM2D.cu:2090-2103): the wall energy's gradient is the wall force (central
differences), its known values (ew at a distance of one vertex radius, zero at the WCA cut),
the force on the wall balances the force on the vertices, distances lose their periodic image
along walled directions, and a round wall at distance d acts like a flat one at distance d."""
import numpy as np

from cudaring_amd import packing
from tests.helpers import oracle_from_state

WCA_CUT = 1.122462048


def one_ring(oracle_mod, center, box, geometry, ew=7.0, box_radius=None, costs=(0, 0, 0, 0), n=16):
    nvp = np.array([n])
    a0, l0, th0, rad = packing.ring_shape_params(nvp)
    pos = packing.place_rings(nvp, a0, np.array([center], dtype=float))
    st = {"nvp": nvp, "box": np.array(box, dtype=float), "pos": pos.reshape(-1), "rad": rad, "l0": l0,
          "theta0": th0, "a0": a0}
    o = oracle_from_state(oracle_mod, st, potential="wca", costs=costs)
    o.set_geometry(geometry, ew=ew, box_radius=box_radius)
    o.calc_neighbor_list(o.set_displacement_cutoff(0.5, 2 * rad.mean()))
    return o, st


def test_wall_energy_known_values(oracle_mod):
    # a ring whose lowest vertex sits exactly one radius above the floor: that vertex carries ew
    o, st = one_ring(oracle_mod, (5.0, 5.0), (10.0, 10.0), "fixedSides", ew=7.0)
    pos = o.get("pos").reshape(-1, 2)
    r = st["rad"][0]
    k = int(np.argmin(pos[:, 1]))
    pos[:, 1] += r - pos[k, 1]
    o.set("pos", pos.reshape(-1))
    o.calc_force_energy()
    e = o.get("energy")
    assert abs(e[k] - 7.0) < 1e-12
    # lifted to the WCA cut (plus a hair): no wall energy left
    pos[:, 1] += (WCA_CUT - 1) * r + 1e-12
    o.set("pos", pos.reshape(-1))
    o.calc_force_energy()
    assert np.abs(o.get("energy")).max() == 0.0


def ring_geometry(n=16):
    a0, l0, _, rad = packing.ring_shape_params(np.array([n]))
    return np.sqrt(2.0 * a0[0] / (n * np.sin(2.0 * np.pi / n))), rad[0]


def test_wall_force_is_energy_gradient(oracle_mod):
    Rg, r = ring_geometry()
    d = Rg + 0.9 * r  # the outermost vertex 0.9 radii from the wall
    for geometry, center, box, R in [("fixedBox", (d, d + 0.3 * r), (6.0, 6.0), None),
                                     ("fixedSides", (3.0, d), (6.0, 6.0), None),
                                     ("roundBox", (0.0, 2.2 - d), (6.0, 6.0), 2.2)]:
        o, st = one_ring(oracle_mod, center, box, geometry, ew=3.0, box_radius=R)
        o.calc_force_energy()
        F = o.get("force")
        assert np.abs(F).max() > 1e-3, geometry  # the ring touches the wall
        x = o.get("pos").copy()
        h = 1e-6
        for i in np.argsort(-np.abs(F))[:6]:
            xp, xm = x.copy(), x.copy()
            xp[i] += h
            xm[i] -= h
            o.set("pos", xp); o.calc_force_energy(); up = o.potential_energy()
            o.set("pos", xm); o.calc_force_energy(); um = o.potential_energy()
            fd = -(up - um) / (2 * h)
            assert abs(fd - F[i]) < 1e-6 * max(1.0, abs(F[i])), (geometry, i, fd, F[i])
        o.set("pos", x)
        o.calc_force_energy()
        # with zero shape costs and one ring, every force is a wall force: the wall takes -F
        np.testing.assert_allclose(o.get("wallforce"), -o.get("force"), rtol=0, atol=1e-12)


def test_round_wall_matches_flat_wall(oracle_mod):
    # the same ring shape, once at distance d below a flat floor-less ceiling and once at
    # distance d inside a large round wall along +y: the normal wall forces agree to O(l0 / R)
    R = 400.0
    Rg, r = ring_geometry()
    d = Rg + 0.9 * r
    o1, st = one_ring(oracle_mod, (0.0, R - d), (2 * R, 2 * R), "roundBox", box_radius=R)
    o2, _ = one_ring(oracle_mod, (5.0, 10.0 - d), (10.0, 10.0), "fixedSides")
    o1.calc_force_energy(); o2.calc_force_energy()
    f1, f2 = o1.get("wallforce").reshape(-1, 2), o2.get("wallforce").reshape(-1, 2)
    assert np.abs(f2[:, 1]).max() > 1e-3
    np.testing.assert_allclose(f1[:, 1], f2[:, 1], rtol=2e-2, atol=1e-6 * np.abs(f2).max())


def test_walls_remove_periodic_images(oracle_mod):
    # two rings touching only through the x boundary: they interact in the normal and
    # fixedSides geometries (x periodic) and not in fixedBox
    nvp = np.array([16, 16])
    a0, l0, th0, rad = packing.ring_shape_params(nvp)
    Rg = np.sqrt(2.0 * a0[0] / (16 * np.sin(2.0 * np.pi / 16)))
    L = 20.0
    c = np.array([[Rg + 0.3 * rad[0], 10.0], [L - Rg - 0.3 * rad[0], 10.0]])
    pos = packing.place_rings(nvp, a0, c).reshape(-1)
    st = {"nvp": nvp, "box": np.array([L, L]), "pos": pos, "rad": rad, "l0": l0, "theta0": th0, "a0": a0}
    pair = {}
    for geometry in ("normal", "fixedSides", "fixedBox"):
        o = oracle_from_state(oracle_mod, st, potential="harmonic", costs=(0, 0, 0, 1))
        o.set_geometry(geometry, ew=0.0)
        o.calc_neighbor_list(o.set_displacement_cutoff(0.5, 2 * rad.mean()))
        o.calc_force_energy()
        pair[geometry] = o.potential_energy()
    assert pair["normal"] > 0 and pair["fixedSides"] == pair["normal"]
    assert pair["fixedBox"] == 0.0

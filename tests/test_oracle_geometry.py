"""Pins the oracle's minimum image, ring topology and ring geometry (H1-H4)."""
import numpy as np
import pytest


def test_pbc_distance_half_away_from_zero(oracle_mod):
    pbc = oracle_mod.pbc_distance
    assert pbc(0.1, 9.9, 10.0) == pytest.approx(0.2)
    assert pbc(9.9, 0.1, 10.0) == pytest.approx(-0.2)
    # exactly half a box: round(0.5) = 1 (half away from zero)
    assert pbc(5.0, 0.0, 10.0) == -5.0
    assert pbc(0.0, 5.0, 10.0) == 5.0
    assert pbc(23.0, 1.0, 10.0) == pytest.approx(2.0)


def test_next_previous_ids(oracle_mod):
    o = oracle_mod.Oracle([3, 4, 5])
    L = o.L
    # ring 0: 0,1,2 ; ring 1: 3..6 ; ring 2: 7..11
    assert [L.orc_next_id(o.h, v, 0) for v in range(3)] == [1, 2, 0]
    assert [L.orc_previous_id(o.h, v, 0) for v in range(3)] == [2, 0, 1]
    assert [L.orc_next_id(o.h, v, 1) for v in range(3, 7)] == [4, 5, 6, 3]
    assert [L.orc_previous_id(o.h, v, 1) for v in range(3, 7)] == [6, 3, 4, 5]
    assert [L.orc_next_id(o.h, v, 2) for v in range(7, 12)] == [8, 9, 10, 11, 7]
    assert [L.orc_previous_id(o.h, v, 2) for v in range(7, 12)] == [11, 7, 8, 9, 10]


def polygon(n, R, cx, cy, phase=0.0):
    a = 2 * np.pi * np.arange(n) / n + phase
    return np.stack([R * np.cos(a) + cx, R * np.sin(a) + cy], 1)


@pytest.mark.parametrize("n", [3, 4, 16, 22, 64])
def test_regular_polygon_closed_forms(oracle_mod, n):
    R = 0.7
    o = oracle_mod.Oracle([n])
    o.set("box", [10.0, 10.0])
    o.set("pos", polygon(n, R, 3.0, 4.0, 0.1).ravel())
    o.L.orc_calc_particle_shape(o.h)
    o.L.orc_calc_particle_positions(o.h)
    A = 0.5 * n * R * R * np.sin(2 * np.pi / n)
    P = 2 * n * R * np.sin(np.pi / n)
    assert o.get("area")[0] == pytest.approx(A, rel=1e-13)
    assert o.get("perimeter")[0] == pytest.approx(P, rel=1e-13)
    # counter-clockwise rings have theta = -2 pi / n (calcAngle sign convention)
    np.testing.assert_allclose(o.get("theta"), -2 * np.pi / n, atol=1e-12)
    np.testing.assert_allclose(o.get("length"), 2 * R * np.sin(np.pi / n), rtol=1e-13)
    np.testing.assert_allclose(o.get("ppos"), [3.0, 4.0], atol=1e-12)


def test_ring_across_boundary_and_image_shift(oracle_mod):
    n, R, L = 16, 0.8, 6.0
    base = polygon(n, R, 0.3, 5.9)  # straddles x = 0 and y = L
    o = oracle_mod.Oracle([n])
    o.set("box", [L, L])
    o.set("pos", base.ravel())
    o.L.orc_calc_particle_shape(o.h)
    o.L.orc_calc_particle_positions(o.h)
    A0, P0, th0 = o.get("area")[0], o.get("perimeter")[0], o.get("theta").copy()
    # wrap half the vertices into the primary box: geometry must not change
    w = base.copy()
    w[::2] = np.mod(w[::2], L)
    o.set("pos", w.ravel())
    o.L.orc_calc_particle_shape(o.h)
    assert o.get("area")[0] == pytest.approx(A0, rel=1e-12)
    assert o.get("perimeter")[0] == pytest.approx(P0, rel=1e-12)
    np.testing.assert_allclose(o.get("theta"), th0, atol=1e-12)
    # translation by a box vector
    o.set("pos", (base + np.array([L, -2 * L])).ravel())
    o.L.orc_calc_particle_shape(o.h)
    assert o.get("area")[0] == pytest.approx(A0, rel=1e-12)


def test_ring_wider_than_half_box(oracle_mod):
    # the reference walks the ring one minimum-image segment at a time, so a ring wider than
    # L/2 keeps its true area and perimeter
    n, R, L = 64, 2.26, 8.9
    o = oracle_mod.Oracle([n])
    o.set("box", [L, L])
    o.set("pos", polygon(n, R, 1.0, 7.5).ravel())
    o.L.orc_calc_particle_shape(o.h)
    o.L.orc_calc_particle_positions(o.h)
    assert o.get("area")[0] == pytest.approx(0.5 * n * R * R * np.sin(2 * np.pi / n), rel=1e-12)
    assert o.get("perimeter")[0] == pytest.approx(2 * n * R * np.sin(np.pi / n), rel=1e-12)
    np.testing.assert_allclose(o.get("ppos"), [1.0, 7.5], atol=1e-12)

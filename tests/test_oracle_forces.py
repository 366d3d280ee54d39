"""Pins the oracle's shape force, pair potentials, vertex-vertex and smooth
interactions (H5-H9) by closed forms, finite differences and invariants."""
import numpy as np
import pytest

from cudaring_amd import packing
from tests.helpers import oracle_from_state

H = 1e-6


def total_energy(o, pos):
    o.set("pos", pos)
    o.calc_force_energy()
    return o.potential_energy()


def fd_forces(o, pos, idx, h=H):
    """-dU/dx by the fourth-order central stencil (8(U+h - U-h) - (U+2h - U-2h)) / 12h."""
    out = []
    for i in idx:
        u = {}
        for k in (-2, -1, 1, 2):
            p = pos.copy(); p[i] += k * h
            u[k] = total_energy(o, p)
        out.append(-(8 * (u[1] - u[-1]) - (u[2] - u[-2])) / (12 * h))
    total_energy(o, pos)
    return np.array(out)


def deformed_ring_state(seed=0, n=16):
    rng = np.random.default_rng(seed)
    st = packing.random_system(1, 0.3, seed)
    pos = st["pos"].reshape(-1, 2)
    pos += 0.03 * (rng.random(pos.shape) - 0.5)
    st["pos"] = pos.ravel()
    st["a0"] = st["a0"] * 1.05
    st["theta0"] = st["theta0"] * (1 + 0.1 * rng.random(st["theta0"].shape))
    return st


@pytest.mark.parametrize("costs", [(1, 1, 0.01, 1), (100, 20, 10, 1), (0, 0, 1, 1)])
def test_shape_force_is_gradient(oracle_mod, costs):
    st = deformed_ring_state()
    o = oracle_from_state(oracle_mod, st, costs=costs, interaction="all")
    pos = st["pos"].copy()
    o.set("pos", pos)
    o.calc_force_energy()
    F = o.get("force")
    idx = list(range(0, len(pos), 3))
    Ffd = fd_forces(o, pos, idx)
    scale = np.abs(F).max()
    assert np.max(np.abs(F[idx] - Ffd)) / scale < 1e-7


def test_shape_rest_polygon_zero_force(oracle_mod):
    n, R = 22, 1.1
    a = 2 * np.pi * np.arange(n) / n
    pos = np.stack([R * np.cos(a) + 4, R * np.sin(a) + 4], 1).ravel()
    o = oracle_mod.Oracle([n])
    o.set("box", [20, 20]); o.set("pos", pos)
    o.set("a0", [0.5 * n * R * R * np.sin(2 * np.pi / n)])
    o.set("l0", np.full(n, 2 * R * np.sin(np.pi / n)))
    o.set("theta0", np.full(n, 2 * np.pi / n))  # reference convention: +2 pi / n
    o.set("rad", np.full(n, 0.05))
    o.set_types("harmonic", "all")
    o.set_energy_costs(1, 1, 0.3, 1)
    o.calc_force_energy()
    assert np.abs(o.get("force")).max() < 1e-12
    # theta = -2 pi / n against theta0 = +2 pi / n: bending offset 1/2 eb n (4 pi / n)^2
    assert o.potential_energy() == pytest.approx(0.5 * 0.3 * n * (4 * np.pi / n) ** 2, rel=1e-12)


def test_shape_force_sums_to_zero_per_ring(oracle_mod):
    st = packing.random_system(6, 0.4, 5)
    o = oracle_from_state(oracle_mod, st, costs=(3, 2, 0.5, 0), interaction="all")
    o.calc_force_energy()
    F = o.get("force").reshape(-1, 2)
    off = np.concatenate([[0], np.cumsum(st["nvp"])])
    for p in range(len(st["nvp"])):
        assert np.abs(F[off[p]:off[p + 1]].sum(0)).max() < 1e-12


def gm(o, d, rs):
    import ctypes
    e = ctypes.c_double(0.0)
    g = o.L.orc_pair_grad_multiple_and_energy(o.h, d, rs, ctypes.byref(e))
    return g, e.value


def test_pair_potential_closed_forms(oracle_mod):
    o = oracle_mod.Oracle([3])
    o.set_energy_costs(1, 1, 1, 2.0)
    o.set_attraction(0.1, 0.2)
    sig = 0.3
    o.set_types("harmonic")
    g, e = gm(o, 0.9 * sig, sig)
    assert g == pytest.approx(2.0 * 0.1 / sig) and e == pytest.approx(0.25 * 2.0 * 0.01)
    o.set_types("wca")
    g, e = gm(o, sig, sig)
    assert e == pytest.approx(0.5 * 2.0) and g == pytest.approx(24 * 2.0 / sig)   # U_pair(sigma) = ec
    g, e = gm(o, 2 ** (1 / 6) * sig * (1 - 1e-14), sig)
    assert abs(g) < 1e-9 and abs(e) < 1e-12
    g, e = gm(o, 1.2 * sig, sig)
    assert g == 0 and e == 0
    o.set_types("adhesive")
    l1, l2, ec = 0.1, 0.2, 2.0
    gi, ei = gm(o, (1 + l1) * sig * (1 - 1e-13), sig)
    go, eo = gm(o, (1 + l1) * sig, sig)
    assert gi == pytest.approx(go, rel=1e-9) and ei == pytest.approx(eo, abs=1e-12)
    assert 2 * eo == pytest.approx(-0.5 * ec * l1 * (l2 - l1), rel=1e-9)          # well depth per pair
    g, e = gm(o, (1 + l2) * sig * (1 - 1e-14), sig)
    assert abs(g) < 1e-9 and abs(e) < 1e-12
    g, e = gm(o, 0.5 * sig, sig)
    assert 2 * e == pytest.approx(0.5 * ec * (0.25 - l1 * l2))
    o.set_types("lj")
    o.set_lj_cutoff(1.5)
    g, e = gm(o, 1.5 * sig * (1 - 1e-14), sig)
    assert abs(g) < 1e-9 and abs(e) < 1e-12


@pytest.mark.parametrize("pot", ["harmonic", "wca", "lj", "adhesive"])
def test_pair_gradient_multiple_is_derivative(oracle_mod, pot):
    o = oracle_mod.Oracle([3])
    o.set_energy_costs(1, 1, 1, 1.3)
    o.set_attraction(0.1, 0.2)
    o.set_types(pot)
    o.set_lj_cutoff(1.5)
    sig = 0.25
    for d in np.linspace(0.8, 1.45, 23) * sig:
        g, _ = gm(o, d, sig)
        _, ep = gm(o, d + 1e-7, sig)
        _, em = gm(o, d - 1e-7, sig)
        # two directed visits store half the pair energy each
        assert g == pytest.approx(-(2 * ep - 2 * em) / 2e-7, rel=2e-5, abs=1e-6)


@pytest.mark.parametrize("pot", ["harmonic", "wca", "lj", "adhesive"])
def test_vertex_list_equals_all_pairs(oracle_mod, pot):
    st = packing.random_system(10, 0.75, 11)
    o = oracle_from_state(oracle_mod, st, potential=pot)
    cut = o.set_displacement_cutoff(0.5, 2 * st["rad"].mean())
    o.calc_neighbor_list(cut)
    o.calc_force_energy()
    F1, E1 = o.get("force"), o.get("energy")
    o.set_types(pot, "all")
    o.calc_force_energy()
    F2, E2 = o.get("force"), o.get("energy")
    np.testing.assert_array_equal(F1, F2)
    np.testing.assert_array_equal(E1, E2)


@pytest.mark.parametrize("inter", ["vertexVertex", "vertexSmooth"])
@pytest.mark.parametrize("pot", ["harmonic", "wca", "adhesive", "lj"])
def test_interaction_newton_third_law(oracle_mod, pot, inter):
    st = packing.random_system(10, 0.8, 13)
    o = oracle_from_state(oracle_mod, st, potential=pot, interaction=inter)
    cut = o.set_displacement_cutoff(0.5, 2 * st["rad"].mean())
    o.calc_neighbor_list(cut)
    o.calc_force_energy()
    F = o.get("force").reshape(-1, 2)
    assert np.abs(F.sum(0)).max() / np.abs(F).max() < 1e-11


@pytest.mark.parametrize("inter", ["vertexVertex", "vertexSmooth"])
@pytest.mark.parametrize("pot", ["harmonic", "wca", "adhesive"])
def test_interaction_force_is_gradient(oracle_mod, pot, inter):
    # rings on a lattice at moderate overlap; finite differences away from branch switches
    st = packing.lattice_system(9, 0.86, 4, perturb=0.3)
    o = oracle_from_state(oracle_mod, st, potential=pot, interaction=inter, costs=(1, 1, 0.01, 1))
    cut = o.set_displacement_cutoff(1.0, 2 * st["rad"].mean())
    o.calc_neighbor_list(cut)
    pos = st["pos"].copy()
    o.set("pos", pos)
    o.calc_force_energy()
    F = o.get("force")
    Fs = F.copy()
    # interaction-only force: subtract the shape part computed without neighbours
    o.set_types(pot, "all")
    o.set_energy_costs(1, 1, 0.01, 0.0)
    o.calc_force_energy()
    Fshape = o.get("force")
    o.set_types(pot, inter)
    o.set_energy_costs(1, 1, 0.01, 1)
    Fint = Fs - Fshape
    active = np.argsort(-np.abs(Fint))[:24]
    Ffd = fd_forces(o, pos, active)
    err = np.abs(Fs[active] - Ffd)
    # skip coordinates whose finite-difference stencil crosses a projection branch
    ok = err / np.abs(F).max() < 1e-7
    assert ok.mean() > 0.9, (err / np.abs(F).max())


def test_ring_costs_gradient_and_uniform_limit(oracle_mod):
    """Per-ring ka, kl, kb: F = -grad U with different costs on each ring, and uniform per-ring
    costs reproduce the global ones exactly."""
    st = packing.random_system(3, 0.35, 4)
    rng = np.random.default_rng(5)
    pos = st["pos"].reshape(-1, 2) + 0.02 * (rng.random((len(st["rad"]), 2)) - 0.5)
    st["pos"] = pos.ravel()
    st["a0"] = st["a0"] * np.array([1.05, 0.97, 1.02])
    o = oracle_from_state(oracle_mod, st, costs=(1, 1, 0.01, 1), interaction="all")
    o.calc_force_energy()
    f_glob, u_glob = o.get("force"), o.potential_energy()
    o.set_ring_costs([1, 1, 1], [1, 1, 1], [0.01, 0.01, 0.01])
    o.calc_force_energy()
    assert np.array_equal(o.get("force"), f_glob) and o.potential_energy() == u_glob
    o.set_ring_costs([3.0, 0.5, 1.5], [2.0, 0.7, 1.1], [0.3, 0.02, 0.1])
    p = st["pos"].copy()
    o.set("pos", p)
    o.calc_force_energy()
    F = o.get("force")
    assert np.max(np.abs(F - f_glob)) > 1e-3
    idx = list(range(0, len(p), 5))
    assert np.max(np.abs(F[idx] - fd_forces(o, p, idx))) / np.abs(F).max() < 1e-7
    o.set_ring_costs()
    o.calc_force_energy()
    assert np.array_equal(o.get("force"), f_glob)

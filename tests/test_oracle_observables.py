"""Pins the oracle's observables (contacts, hexatic order, MSD, ISF, stress and pressure;
DPM2D.cu:976-983, 1284-1298, 1486-1510, 2181-2198, 3408-3428; cudaKernel.cuh
1525-1565, 1791-1822, 1935-1976): contacts are the particles within the list radius (brute force
over positions) and symmetric; psi6 is 1 on a triangular lattice and 0 on a square one; a rigid
translation gives MSD = |t|^2 / (Lx Ly) and ISF = sin(k|t|)/(k|t|); the per-particle stresses
sum to the particle-level virial sum over overlapping pairs of (c_q - c_p) (x) f."""
import numpy as np

from cudaring_amd import packing
from tests.helpers import oracle_from_state


def lattice_rings(kind, m=6, n=16, gap_vertex_diameters=0.8, rad_scale=1.0):
    """m x m identical regular rings on a periodic triangular or square lattice; rad_scale
    shrinks the vertex radii (gaps stay in units of the unscaled diameter)."""
    nvp = np.full(m * m, n)
    a0, l0, th0, rad = packing.ring_shape_params(nvp)
    R = np.sqrt(2.0 * a0[0] / (n * np.sin(2.0 * np.pi / n)))
    a = 2 * R + gap_vertex_diameters * 2 * rad[0]
    if kind == "triangular":
        centers = np.array([(a * (i + 0.5 * (j % 2)), a * np.sqrt(3) / 2 * j) for j in range(m) for i in range(m)])
        box = np.array([a * m, a * np.sqrt(3) / 2 * m])
    else:
        centers = np.array([(a * i, a * j) for j in range(m) for i in range(m)], dtype=float)
        box = np.array([a * m, a * m])
    centers = centers + 0.37 * a
    pos = packing.place_rings(nvp, a0, centers)
    return {"nvp": nvp, "box": box, "pos": pos.reshape(-1), "rad": rad * rad_scale, "l0": l0, "theta0": th0,
            "a0": a0}


def listed(oracle_mod, st, skin=0.3, **kw):
    o = oracle_from_state(oracle_mod, st, **kw)
    cut = o.set_displacement_cutoff(skin, 2 * st["rad"].mean())
    o.calc_neighbor_list(cut)
    return o, cut


def test_contacts_are_listed_particles(oracle_mod):
    st = packing.lattice_system(12, 0.8, 5, perturb=0.2)
    o, cut = listed(oracle_mod, st)
    lst, cnt = o.contacts()
    nvp = st["nvp"]
    ring = np.repeat(np.arange(len(nvp)), nvp)
    pos = st["pos"].reshape(-1, 2)
    L = st["box"]
    d = pos[:, None, :] - pos[None, :, :]
    d -= L * np.round(d / L)
    dist = np.sqrt((d ** 2).sum(-1))
    rs = st["rad"][:, None] + st["rad"][None, :]
    near = -(1 - dist / rs) < cut  # the list predicate, pair by pair
    for p in range(len(nvp)):
        mine = ring == p
        expect = set(np.unique(ring[np.any(near[mine], axis=0) & ~mine]))
        got = lst[p, :cnt[p]]
        assert set(got) == expect and len(got) == len(expect)
        assert np.all(lst[p, cnt[p]:] == -1)
        for q in got:
            assert p in lst[q, :cnt[q]]


def test_hexatic_order_lattices(oracle_mod):
    o, _ = listed(oracle_mod, lattice_rings("triangular"))
    o.contacts()
    assert abs(o.hexatic_order() - 1.0) < 1e-12
    o, _ = listed(oracle_mod, lattice_rings("square"))
    _, cnt = o.contacts()
    assert cnt.min() >= 4
    assert o.hexatic_order() < 1e-12


def test_msd_isf_rigid_translation(oracle_mod):
    st = packing.lattice_system(9, 0.7, 3)
    o, _ = listed(oracle_mod, st)
    o.set_initial_positions()
    t = np.array([0.031, -0.047])
    o.set("pos", (st["pos"].reshape(-1, 2) + t).reshape(-1))
    L = st["box"]
    assert abs(o.vertex_msd() - (t @ t) / (L[0] * L[1])) < 1e-15
    assert abs(o.particle_msd() - (t @ t) / (L[0] * L[1])) < 1e-15
    k = np.pi / (2 * st["rad"][0])
    dn = np.sqrt(t @ t)
    assert abs(o.vertex_isf() - np.sin(k * dn) / (k * dn)) < 1e-14
    assert abs(o.particle_isf(3.0) - np.sin(3.0 * dn) / (3.0 * dn)) < 1e-14


def test_stress_sums_to_particle_virial(oracle_mod):
    # square lattice of rings whose vertices are 0.3 segment lengths in radius: the facing
    # vertices of neighbouring rings, half a segment apart, are the only overlapping pairs
    st = lattice_rings("square", m=4, gap_vertex_diameters=0.25, rad_scale=0.3)
    o, _ = listed(oracle_mod, st, costs=(1, 1, 0.1, 1))
    o.calc_force_energy()
    per = o.per_particle_stress()
    total = o.stress_tensor()
    np.testing.assert_allclose(total, per.sum(0), rtol=0, atol=1e-15)
    # independent virial: pairs (v, j) of different rings with overlap, centres c
    pos = st["pos"].reshape(-1, 2)
    L = st["box"]
    ring = np.repeat(np.arange(len(st["nvp"])), st["nvp"])
    c = o.get("ppos").reshape(-1, 2)
    rad = st["rad"]
    expect = np.zeros(4)
    noverlap = np.zeros(len(pos), dtype=int)
    for v in range(len(pos)):
        for j in range(v + 1, len(pos)):
            if ring[j] == ring[v]:
                continue
            dv = pos[j] - pos[v]
            dv -= L * np.round(dv / L)
            dist = np.sqrt(dv @ dv)
            rs = rad[v] + rad[j]
            ov = 1 - dist / rs
            if ov > 0:
                noverlap[v] += 1
                noverlap[j] += 1
                f = (ov / rs) * dv / (dist * rs)
                dc = c[ring[j]] - c[ring[v]]
                dc -= L * np.round(dc / L)
                expect += np.array([dc[0] * f[0], dc[0] * f[1], dc[1] * f[0], dc[1] * f[1]])
    assert noverlap.max() == 1 and noverlap.sum() > 0
    np.testing.assert_allclose(total, expect, rtol=1e-12, atol=1e-15)
    assert abs(o.pressure() - (total[0] + total[3]) / (2 * len(pos))) < 1e-18


def test_geometric_contacts_within_one_radius_sum(oracle_mod):
    """calcGeometricContacts (DPM2D.cu:3430-3450): rings with a listed vertex pair at
    d < 2 radSum (gap 1), by brute force over positions, at most 8 per ring; a list built with a
    skin below the gap bounds them, so the test uses skin 1.2 (list radius above 2 radSum)"""
    st = packing.lattice_system(12, 0.8, 5, perturb=0.2)
    o, cut = listed(oracle_mod, st, skin=1.2)
    lst, cnt = o.geometric_contacts(1.0)
    assert lst.shape[1] == 8
    nvp = st["nvp"]
    ring = np.repeat(np.arange(len(nvp)), nvp)
    pos = st["pos"].reshape(-1, 2)
    L = st["box"]
    d = pos[:, None, :] - pos[None, :, :]
    d -= L * np.round(d / L)
    dist = np.sqrt((d ** 2).sum(-1))
    rs = st["rad"][:, None] + st["rad"][None, :]
    close = -(1 - dist / rs) < 1.0
    for p in range(len(nvp)):
        mine = ring == p
        expect = set(np.unique(ring[np.any(close[mine], axis=0) & ~mine]))
        got = lst[p, :cnt[p]]
        assert len(got) == min(len(expect), 8) and set(got) <= expect
        assert np.all(lst[p, cnt[p]:] == -1)
    assert cnt.max() > 0

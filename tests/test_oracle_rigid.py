"""Pins the oracle's rigid-ring path (calcRigidForceEnergy DPM2D.cu:2128-2142, cudaKernel.cuh:1412-1511;
RigidNVE / RigidLangevin Simulator.cu:241-371; rigid FIRE FIRE.cu:173-298):

- the ring force is minus the energy's gradient under a translation of the ring, and the ring
  torque (torque / moment of inertia times the moment of inertia) minus its derivative under a
  rotation about the centre (central differences), for vertexVertex and vertexSmooth;
- the moment of inertia of a regular n-gon of circumradius R is n R^2;
- rigid NVE moves rings rigidly, conserves U + sum |v|^2 / 2 + sum I w^2 / 2 to second order in
  the time step and retraces its path when all velocities are reversed;
- on an isolated ring the rigid Langevin kick is the linear map v -> c^2 v + (1 + c) z per step
  with c = (1 - h gamma)(1 + h^2 / 2) and z the step's unit Gaussian pair (the reference adds
  the raw draw, unscaled by temperature or time step, in both half kicks);
- the injected ring velocities have zero mean and the angular velocities start at zero;
- rigid FIRE lowers the energy and the largest ring force / torque by orders of magnitude.
"""
import numpy as np
import pytest

from cudaring_amd import packing
from tests.helpers import oracle_from_state


def ring_block(n_side=3, gap=0.6, seed=2, nvp=None):
    """n_side^2 rings on a square grid, facing vertices `gap` radii apart (overlapping for < 2)."""
    rng = np.random.default_rng(seed)
    nvp = nvp if nvp is not None else packing.ring_sizes(n_side * n_side, rng=rng)
    a0, l0, th0, rad = packing.ring_shape_params(nvp)
    R = np.sqrt(2.0 * a0 / (nvp * np.sin(2.0 * np.pi / nvp)))
    r = rad.max()
    step = 2 * R.max() + gap * r
    g = (np.arange(n_side) + 0.5) * step
    centers = np.array([(x, y) for y in g for x in g]) + rng.uniform(-0.3, 0.3, (len(nvp), 2)) * r
    pos = packing.place_rings(nvp, a0, centers)
    # a random initial orientation per ring
    out = []
    k0 = 0
    for p, n in enumerate(nvp):
        q = pos[k0:k0 + n] - centers[p]
        t = rng.uniform(0, 2 * np.pi)
        c, s = np.cos(t), np.sin(t)
        out.append(np.c_[q[:, 0] * c - q[:, 1] * s, q[:, 0] * s + q[:, 1] * c] + centers[p])
        k0 += n
    pos = np.vstack(out)
    L = n_side * step
    return {"nvp": nvp, "box": np.array([L, L]), "pos": pos.reshape(-1), "rad": rad, "l0": l0, "theta0": th0,
            "a0": a0, "R": R}


def rigid_oracle(oracle_mod, st, pot="harmonic", inter="vertexVertex", dt=1e-3):
    o = oracle_from_state(oracle_mod, st, potential=pot, interaction=inter, costs=(0, 0, 0, 1), dt=dt)
    o.calc_neighbor_list(o.set_displacement_cutoff(0.5, 2 * st["rad"].mean()))
    o.calc_rigid_force_energy()
    return o


def energy_at(o, pos):
    o.set("pos", pos.reshape(-1))
    o.calc_rigid_force_energy()
    return o.potential_energy()


@pytest.mark.parametrize("pot,inter", [("harmonic", "vertexVertex"), ("wca", "vertexVertex"),
                                       ("harmonic", "vertexSmooth"), ("wca", "vertexSmooth")])
def test_ring_force_and_torque_are_energy_derivatives(oracle_mod, pot, inter):
    st = ring_block(gap=0.6 if pot == "harmonic" else 1.05)
    o = rigid_oracle(oracle_mod, st, pot, inter)
    F = o.get("pforce").reshape(-1, 2).copy()
    T = o.get("ptorque").copy()
    c = o.get("ppos").reshape(-1, 2).copy()
    pos0 = o.get("pos").reshape(-1, 2).copy()
    assert np.abs(F).max() > 1e-3  # the rings touch
    first = np.r_[0, np.cumsum(st["nvp"])[:-1]]
    h = 1e-6
    for p in range(len(st["nvp"])):
        sl = slice(first[p], first[p] + st["nvp"][p])
        rel = pos0[sl] - c[p]
        I = (rel ** 2).sum()
        for dim in range(2):
            e = []
            for sgn in (1, -1):
                q = pos0.copy()
                q[sl, dim] += sgn * h
                e.append(energy_at(o, q))
            dE = (e[0] - e[1]) / (2 * h)
            assert abs(dE + F[p, dim]) < 1e-6 * max(1.0, np.abs(F).max()), (p, dim)
        e = []
        for sgn in (1, -1):
            t = sgn * h
            q = pos0.copy()
            q[sl] = c[p] + np.c_[rel[:, 0] * np.cos(t) - rel[:, 1] * np.sin(t),
                                 rel[:, 0] * np.sin(t) + rel[:, 1] * np.cos(t)]
            e.append(energy_at(o, q))
        dE = (e[0] - e[1]) / (2 * h)
        assert abs(dE + T[p] * I) < 1e-6 * max(1.0, np.abs(T * I).max()), p


def test_moment_of_inertia_of_regular_rings(oracle_mod):
    st = ring_block(gap=3.0)  # apart: no forces
    o = rigid_oracle(oracle_mod, st)
    assert np.abs(o.get("pforce")).max() == 0.0 and np.abs(o.get("ptorque")).max() == 0.0
    # force-free rings spun at angular velocity w: every vertex stays on its ring's circumcircle
    o.set("prad", st["R"])
    o.init_rigid_nve(0.0)
    w = np.linspace(-1, 1, len(st["nvp"]))
    o.set("pangvel", w)
    o.rigid_loop(25)
    c = o.get("ppos").reshape(-1, 2)
    pos = o.get("pos").reshape(-1, 2)
    first = np.r_[0, np.cumsum(st["nvp"])[:-1]]
    for p, n in enumerate(st["nvp"]):
        d = np.linalg.norm(pos[first[p]:first[p] + n] - c[p], axis=1)
        np.testing.assert_allclose(d, st["R"][p], rtol=1e-12)
        # n R^2, the ring's moment of inertia, from the rotated vertices
        assert abs((d ** 2).sum() - n * st["R"][p] ** 2) < 1e-12 * n * st["R"][p] ** 2
    np.testing.assert_allclose(o.get("pangle"), w * 25 * 1e-3, rtol=1e-12)
    # the rigid kinetic energy of DPM2D.cu:1149-1159 uses the particle radius as lever arm
    assert abs(o.rigid_kinetic_energy() - 0.5 * ((w * st["R"]) ** 2).sum()) < 1e-14


def hamiltonian(o, st):
    c = o.get("ppos").reshape(-1, 2)
    pos = o.get("pos").reshape(-1, 2)
    first = np.r_[0, np.cumsum(st["nvp"])[:-1]]
    I = np.array([((pos[first[p]:first[p] + n] - c[p]) ** 2).sum() for p, n in enumerate(st["nvp"])])
    v = o.get("pvel")
    w = o.get("pangvel")
    return o.potential_energy() + 0.5 * (v ** 2).sum() + 0.5 * (I * w ** 2).sum()


def test_rigid_nve_is_rigid_and_conserves_energy(oracle_mod):
    drifts = []
    for dt in (4e-3, 2e-3):
        st = ring_block(gap=0.8)
        o = rigid_oracle(oracle_mod, st, "harmonic", "vertexVertex", dt=dt)
        o.init_rigid_nve(1e-3)
        v = o.get("pvel").reshape(-1, 2)
        assert np.abs(v.mean(0)).max() < 1e-15 and np.abs(o.get("pangvel")).max() == 0.0
        o.set("pangvel", np.linspace(-0.2, 0.2, len(st["nvp"])))
        # the init clears the ring torques (initRotationalVariables): recompute them, or the
        # first half kick of a spinning ring misses its torque, an O(dt) energy offset
        o.calc_rigid_force_energy()
        H0 = hamiltonian(o, st)
        pos0 = o.get("pos").reshape(-1, 2).copy()
        o.rigid_loop(int(round(0.4 / dt)))
        H1 = hamiltonian(o, st)
        drifts.append(abs(H1 - H0))
        # rigid: distances inside every ring are unchanged
        pos = o.get("pos").reshape(-1, 2)
        first = np.r_[0, np.cumsum(st["nvp"])[:-1]]
        for p, n in enumerate(st["nvp"]):
            a, b = pos0[first[p]:first[p] + n], pos[first[p]:first[p] + n]
            da = np.linalg.norm(a[:, None] - a[None], axis=2)
            db = np.linalg.norm(b[:, None] - b[None], axis=2)
            assert np.abs(da - db).max() < 1e-12
        assert np.abs(pos - pos0).max() > 1e-3  # it moved
    assert drifts[1] < 0.35 * drifts[0] + 1e-14, drifts  # second order: about a quarter


def test_rigid_nve_is_time_reversible(oracle_mod):
    st = ring_block(gap=0.8)
    o = rigid_oracle(oracle_mod, st, "harmonic", "vertexVertex", dt=2e-3)
    o.init_rigid_nve(1e-3)
    o.set("pangvel", np.linspace(-0.2, 0.2, len(st["nvp"])))
    o.calc_rigid_force_energy()
    p0 = o.get("pos").copy()
    o.rigid_loop(100)
    assert np.abs(o.get("pos") - p0).max() > 1e-3
    o.set("pvel", -o.get("pvel"))
    o.set("pangvel", -o.get("pangvel"))
    o.rigid_loop(100)
    assert np.abs(o.get("pos") - p0).max() < 1e-12


def test_rigid_langevin_kick_on_isolated_ring(oracle_mod):
    st = ring_block(n_side=1, gap=3.0, nvp=np.array([16]))
    st["box"] = np.array([20.0, 20.0])
    gamma, dt, seed = 0.7, 0.01, 11
    o = oracle_from_state(oracle_mod, st, costs=(0, 0, 0, 1), dt=dt, seed=seed)
    o.calc_neighbor_list(o.set_displacement_cutoff(0.5, 2 * st["rad"].mean()))
    o.calc_rigid_force_energy()
    o.init_rigid_langevin(0.3, gamma)
    v = o.get("pvel").copy()
    assert np.all(v == 0.0)  # one ring: the injected velocity minus its mean
    h = 0.5 * dt
    c = (1 - h * gamma) * (1 + 0.5 * h * h)
    x = o.get("ppos").copy()
    for n in range(5):
        z = np.array(oracle_mod.gauss_pair(seed, 0, 1, n, 0))
        x = x + dt * (c * v + z)  # position update after the first half kick
        v = c * c * v + (1 + c) * z
        o.rigid_loop(1)
        np.testing.assert_allclose(o.get("pvel"), v, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(o.get("ppos"), x, rtol=1e-12)
    assert np.abs(o.get("pangvel")).max() == 0.0


def test_rigid_fire_relaxes(oracle_mod):
    st = ring_block(gap=0.5)
    o = rigid_oracle(oracle_mod, st, "harmonic", "vertexVertex")
    e0 = o.potential_energy()
    f0 = o.rigid_max_unbalanced_force()
    dt = 1e-2 * 2 * np.sqrt(st["a0"].mean() / np.pi)
    o.init_rigid_fire([0.2, 0.5, 1.1, 0.99, dt, 10 * dt, 0.2], 20, 0)
    o.rigid_fire(3000)
    e1 = o.potential_energy()
    f1 = o.rigid_max_unbalanced_force()
    assert e1 < 1e-3 * e0 and f1 < 1e-3 * f0, (e0, e1, f0, f1)

"""Pins the oracle's BAOAB, Brownian and driven-Brownian integrators and the activeShape
rest-length update (Simulator.cu:112-203, DPM2D.cu:2037-2056, 3713-3753) by properties the
schemes fix: the gamma -> 0 limit of BAOAB is velocity Verlet; its O step is the exact
Ornstein-Uhlenbeck update (stationary variance T, relaxation exp(-gamma t)); Brownian
displacements obey the Einstein relation <dx^2> = 2 T t / gamma; the overdamped step is
gradient descent (dU = -dt |F|^2 / gamma to first order); the rest lengths follow the linear
recursion of their relaxation and diffuse with <dl0^2> = 2 l0Diff t."""
import numpy as np

from cudaring_amd import packing
from tests.helpers import oracle_from_state


def dense(oracle_mod, dt, seed=5, noise="langevin", costs=(1, 1, 0.01, 1)):
    st = packing.lattice_system(9, 0.8, seed)
    o = oracle_from_state(oracle_mod, st, costs=costs, dt=dt, seed=seed)
    o.set_noise_type(noise)
    cut = o.set_displacement_cutoff(0.3, 2 * st["rad"].mean())
    o.calc_neighbor_list(cut)
    o.calc_force_energy()
    return o, st


def free_rings(oracle_mod, nrings, dt, seed, noise):
    """Rings far apart with every cost zero: no forces at all."""
    st = packing.lattice_system(nrings, 0.02, seed)
    o = oracle_from_state(oracle_mod, st, costs=(0, 0, 0, 0), dt=dt, seed=seed)
    o.set_noise_type(noise)
    cut = o.set_displacement_cutoff(0.3, 2 * st["rad"].mean())
    o.calc_neighbor_list(cut)
    o.calc_force_energy()
    return o, st


def ring_means(o, st, field):
    a = o.get(field).reshape(-1, 2)
    off = np.concatenate([[0], np.cumsum(st["nvp"])])
    return np.array([a[off[i]:off[i + 1]].mean(0) for i in range(len(st["nvp"]))])


def test_baoab_gamma_zero_is_velocity_verlet(oracle_mod):
    a, _ = dense(oracle_mod, 0.01, noise="baoab")
    b, _ = dense(oracle_mod, 0.01)
    a.init_langevin(1e-2, 0.0, zero_momentum=True)
    b.init_nve(1e-2, zero_momentum=True)
    np.testing.assert_array_equal(a.get("vel"), b.get("vel"))
    a.langevin(200)
    b.nve(200)
    scale = np.abs(b.get("pos")).max()
    assert np.abs(a.get("pos") - b.get("pos")).max() < 1e-12 * scale
    assert np.abs(a.get("vel") - b.get("vel")).max() < 1e-10


def test_baoab_free_mean_drift_at_zero_temperature(oracle_mod):
    # T = 0: v_k = c^k v_0 with c = exp(-gamma dt) (the OU relaxation), and each step moves by
    # dt/2 (v before O + v after O): x_n - x_0 = dt/2 v_0 (1 + c)(1 - c^n)/(1 - c)
    o, st = free_rings(oracle_mod, 16, 0.05, 3, "baoab")
    gamma, n = 0.7, 40
    o.init_langevin(0.0, gamma)
    rng = np.random.default_rng(0)
    v0 = np.repeat(rng.normal(size=(16, 2)), st["nvp"], axis=0).reshape(-1)
    o.set("vel", v0)
    x0 = o.get("pos").copy()
    o.langevin(n)
    c = np.exp(-gamma * 0.05)
    expect = 0.5 * 0.05 * v0 * (1 + c) * (1 - c ** n) / (1 - c)
    assert np.abs(o.get("pos") - x0 - expect).max() < 1e-12
    assert np.abs(o.get("vel") - c ** n * v0).max() < 1e-13


def test_baoab_ou_variance(oracle_mod):
    # from rest, the O steps give Var(v) = T (1 - exp(-2 gamma t)) exactly, per ring and component
    o, st = free_rings(oracle_mod, 200, 0.05, 4, "baoab")
    T, gamma, n = 0.3, 2.0, 6
    o.init_langevin(T, gamma)
    o.set("vel", np.zeros(2 * o.nv))
    o.langevin(n)
    v = ring_means(o, st, "vel")
    var = v.var()
    expect = T * (1 - np.exp(-2 * gamma * n * 0.05))
    # 400 samples: the sample variance has relative spread sqrt(2/400) = 7%
    assert abs(var / expect - 1) < 0.25, (var, expect)
    # every vertex of a ring carries the ring's velocity (per-ring noise)
    vv = o.get("vel").reshape(-1, 2)
    off = np.concatenate([[0], np.cumsum(st["nvp"])])
    assert max(np.ptp(vv[off[i]:off[i + 1]], axis=0).max() for i in range(200)) == 0.0


def test_brownian_einstein_relation(oracle_mod):
    o, st = free_rings(oracle_mod, 200, 0.05, 6, "brownian")
    T, gamma, n = 0.5, 3.0, 10
    o.init_langevin(T, gamma)
    x0 = ring_means(o, st, "pos")
    o.langevin(n)
    d = ring_means(o, st, "pos") - x0
    msd = (d ** 2).mean()
    expect = 2 * T * n * 0.05 / gamma
    assert abs(msd / expect - 1) < 0.25, (msd, expect)


def test_driven_brownian_is_noiseless_brownian(oracle_mod):
    a, _ = dense(oracle_mod, 1e-3, noise="drivenBrownian")
    b, _ = dense(oracle_mod, 1e-3, noise="brownian")
    a.init_langevin(0.4, 2.0)
    b.init_langevin(0.0, 2.0)
    a.langevin(50)
    b.langevin(50)
    np.testing.assert_array_equal(a.get("pos"), b.get("pos"))


def test_overdamped_step_is_gradient_descent(oracle_mod):
    # one driven-Brownian step: U(x + dt F / gamma) - U(x) = -dt |F|^2 / gamma + O(dt^2)
    gamma = 2.0
    ratios = []
    for dt in (1e-5, 1e-6):
        o, _ = dense(oracle_mod, dt, noise="drivenBrownian", costs=(1, 1, 0.1, 1))
        o.init_langevin(0.0, gamma)
        F = o.get("force")
        U0 = o.potential_energy()
        o.langevin(1)
        o.calc_force_energy()
        ratios.append((o.potential_energy() - U0) / (-dt * (F @ F) / gamma))
    assert abs(ratios[1] - 1) < 1e-4 and abs(ratios[1] - 1) < abs(ratios[0] - 1), ratios


def test_active_shape_rest_length_recursion(oracle_mod):
    # no noise: l0 - l0Rest grows by (1 + dt / tau) per force call
    o, st = dense(oracle_mod, 1e-3)
    o.set_types("harmonic", "vertexVertex", "harmonic", "activeShape")
    tau = 0.05
    o.set_active_tension(tau, 0.0)
    rest = o.get("l0").copy()
    delta = 1e-3 * np.sin(np.arange(o.nv))
    o.set("l0", rest + delta)
    for _ in range(7):
        o.calc_force_energy()
    np.testing.assert_allclose(o.get("l0") - rest, delta * (1 + 1e-3 / tau) ** 7, rtol=1e-9, atol=1e-15)


def test_active_shape_rest_length_diffusion(oracle_mod):
    # tau -> infinity: l0 performs a per-ring random walk with <dl0^2> = 2 l0Diff t
    o, st = free_rings(oracle_mod, 400, 1e-3, 8, "langevin")
    o.set_types("harmonic", "vertexVertex", "harmonic", "activeShape")
    D = 0.2
    o.set_active_tension(1e300, D)
    l0 = o.get("l0").copy()
    n = 30
    for _ in range(n):
        o.calc_force_energy()
    dl = o.get("l0") - l0
    off = np.concatenate([[0], np.cumsum(st["nvp"])])
    per_ring = np.array([dl[off[i]] for i in range(400)])
    assert max(np.ptp(dl[off[i]:off[i + 1]]) for i in range(400)) < 1e-15
    msd = (per_ring ** 2).mean()
    expect = 2 * D * n * 1e-3
    assert abs(msd / expect - 1) < 0.25, (msd, expect)

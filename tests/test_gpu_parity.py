"""GPU parity: libdpmring.so (through the C ABI) against the CPU oracle on identical
seeded states.  Bars: forces and energies within 1e-12 relative (max |diff| over
the largest term), neighbour lists identical as sets, short NVE / Langevin / FIRE
windows within 1e-10 before chaotic divergence."""
import numpy as np
import pytest

from cudaring_amd import packing
from tests.conftest import rel_err
from tests.helpers import gpu_from_state, oracle_from_state, relaxed_state

pytestmark = pytest.mark.gpu

TOL_F = 1e-12
POTS = ["harmonic", "wca", "lj", "adhesive"]
INTERS = ["vertexVertex", "vertexSmooth"]


def prepare(oracle_mod, st, pot, inter, skin=0.5, **kw):
    o = oracle_from_state(oracle_mod, st, potential=pot, interaction=inter, **kw)
    g = gpu_from_state(st, potential=pot, interaction=inter, **kw)
    size = 2 * st["rad"].mean()
    c1 = o.set_displacement_cutoff(skin, size)
    c2 = g.set_displacement_cutoff(skin, size)
    assert c1 == c2
    o.calc_neighbor_list(c1)
    g.calc_neighbor_list(c2)
    return o, g


@pytest.mark.parametrize("inter", INTERS)
@pytest.mark.parametrize("pot", POTS)
def test_force_energy_parity(oracle_mod, pot, inter):
    print(f"case force_energy pot={pot} inter={inter}", flush=True)
    st = packing.random_system(12, 0.8, 101)
    o, g = prepare(oracle_mod, st, pot, inter)
    o.calc_force_energy()
    g.calc_force_energy()
    assert rel_err(g.get("force"), o.get("force")) < TOL_F
    assert rel_err(g.get("energy"), o.get("energy")) < TOL_F
    if inter == "vertexSmooth":
        assert rel_err(g.get("penergy"), o.get("penergy")) < TOL_F
    assert abs(g.potential_energy()[0] - o.potential_energy()) <= TOL_F * abs(o.potential_energy())
    for f in ("area", "perimeter", "theta", "length", "ppos"):
        assert rel_err(g.get(f), o.get(f)) < 1e-13, f


@pytest.mark.parametrize("nrings,phi", [(3, 0.5), (12, 0.8), (80, 0.85)])
def test_neighbor_sets(oracle_mod, nrings, phi):
    print(f"case neighbor_sets nrings={nrings}", flush=True)
    st = packing.random_system(nrings, phi, 7 + nrings)
    o, g = prepare(oracle_mod, st, "wca", "vertexSmooth")
    ro, _, _ = o.neighbors()
    rg, _, _ = g.neighbors()
    assert [sorted(r) for r in rg] == [sorted(r) for r in ro]
    assert rg == ro  # rows are ascending on both sides


def test_ring_size_edges(oracle_mod):
    print("case ring_size_edges", flush=True)
    rng = np.random.default_rng(5)
    nvp = np.array([3, 64, 5, 33, 64, 3, 17, 40])
    a0, l0, th0, rad = packing.ring_shape_params(nvp, 1.1, n_min=16)
    L = np.sqrt(packing.preferred_phi_area(nvp, a0, rad) / 0.6)
    pos = packing.place_rings(nvp, a0, rng.random((len(nvp), 2)) * L, rng=rng, jitter=1e-3, l0=l0)
    st = {"nvp": nvp, "box": np.array([L, L]), "pos": pos.ravel(), "rad": rad, "l0": l0, "theta0": th0, "a0": a0}
    for inter in INTERS:
        o, g = prepare(oracle_mod, st, "harmonic", inter)
        o.calc_force_energy()
        g.calc_force_energy()
        assert rel_err(g.get("force"), o.get("force")) < TOL_F
        assert rel_err(g.get("energy"), o.get("energy")) < TOL_F


@pytest.mark.parametrize("pot,inter", [("harmonic", "vertexVertex"), ("wca", "vertexSmooth")])
def test_nve_window(oracle_mod, pot, inter):
    print(f"case nve_window {pot} {inter}", flush=True)
    st = packing.lattice_system(16, 0.8, 17) if pot == "harmonic" else relaxed_state(16, 0.8, 17)
    o, g = prepare(oracle_mod, st, pot, inter, dt=0.005)
    o.calc_force_energy(); g.calc_force_energy()
    o.init_nve(1e-2, zero_momentum=True); g.init_nve(1e-2, zero_momentum=True)
    assert rel_err(g.get("vel"), o.get("vel")) < 1e-15
    for chunk in (1, 7, 42):
        o.nve(chunk); g.nve(chunk)
        assert rel_err(g.get("pos"), o.get("pos"), scale=1.0) < 1e-11
        assert rel_err(g.get("vel"), o.get("vel")) < 1e-9
        assert rel_err(g.get("force"), o.get("force")) < 1e-9
    assert list(g.update_count()) == [o.update_count()]


@pytest.mark.parametrize("inter", INTERS)
def test_langevin_window(oracle_mod, inter):
    print(f"case langevin_window {inter}", flush=True)
    st = relaxed_state(16, 0.85, 23, final=("wca", inter))
    o, g = prepare(oracle_mod, st, "wca", inter, dt=0.002, seed=99)
    o.calc_force_energy(); g.calc_force_energy()
    o.init_langevin(1e-2, 1.0); g.init_langevin(1e-2, 1.0)
    for chunk in (1, 5, 30):
        o.langevin(chunk); g.langevin(chunk)
        assert rel_err(g.get("pos"), o.get("pos"), scale=1.0) < 1e-11
        assert rel_err(g.get("vel"), o.get("vel")) < 1e-9
    assert abs(g.temperature()[0] - o.temperature()) < 1e-9 * o.temperature()


def test_active_force(oracle_mod):
    print("case active_force", flush=True)
    st = relaxed_state(16, 0.8, 31)
    st["pangle"] = np.random.default_rng(1).uniform(-np.pi, np.pi, 16)
    o, g = prepare(oracle_mod, st, "wca", "vertexSmooth", particle="active", dt=0.01)
    o.set_self_propulsion(0.1, 100.0); g.set_self_propulsion(0.1, 100.0)
    for _ in range(3):
        o.calc_force_energy(); g.calc_force_energy()
        assert rel_err(g.get("pangle"), o.get("pangle"), scale=np.pi) < 1e-14
        assert rel_err(g.get("force"), o.get("force")) < TOL_F
    o.init_langevin(1e-3, 1.0); g.init_langevin(1e-3, 1.0)
    o.langevin(20); g.langevin(20)
    assert rel_err(g.get("pos"), o.get("pos"), scale=1.0) < 1e-11


def test_fire_sequence(oracle_mod):
    print("case fire_sequence", flush=True)
    st = packing.lattice_system(16, 0.86, 41, perturb=0.3)
    o, g = prepare(oracle_mod, st, "harmonic", "vertexVertex", costs=(1, 1, 0.1, 1))
    o.calc_force_energy(); g.calc_force_energy()
    dt = 0.01 * o.mean_particle_size()
    prm = [0.2, 0.5, 1.1, 0.99, dt, 10 * dt, 0.2]
    o.init_fire(prm, 20, 0); g.init_fire(prm, 20, 0)
    for chunk in (1, 10, 60):
        o.fire(chunk); g.fire(chunk)
        fo, fg = o.fire_state(), g.fire_state()
        assert fg["numStep"][0] == fo["numStep"]
        assert fg["dt"][0] == pytest.approx(fo["dt"], rel=1e-14)
        assert fg["alpha"][0] == pytest.approx(fo["alpha"], rel=1e-14)
        assert rel_err(g.get("pos"), o.get("pos"), scale=1.0) < 1e-11
    assert g.max_unbalanced_force()[0] == pytest.approx(o.max_unbalanced_force(), rel=1e-8)


def test_batched_set_matches_each_system(oracle_mod):
    print("case batched_set", flush=True)
    systems = [packing.random_system(n, 0.75, 200 + n) for n in (5, 12, 9)]
    cat = packing.concat_systems(systems)
    g = gpu_from_state(cat, potential="harmonic", interaction="vertexSmooth", seed=5,
                       rings_per_system=cat["rings_per_system"])
    size = 2 * np.mean(cat["rad"])
    cut = g.set_displacement_cutoff(0.4, size)
    g.calc_neighbor_list(cut)
    g.calc_force_energy()
    g.init_langevin(1e-2, 1.0)
    g.langevin(10)
    F, X = g.get("force").reshape(-1, 2), g.get("pos").reshape(-1, 2)
    U = g.potential_energy()
    off = g.vert_off
    for s, st in enumerate(systems):
        o = oracle_from_state(oracle_mod, st, potential="harmonic", interaction="vertexSmooth", seed=5, system_id=s)
        o.set_displacement_cutoff(0.4, size)
        o.calc_neighbor_list(cut)
        o.calc_force_energy()
        o.init_langevin(1e-2, 1.0)
        o.langevin(10)
        assert rel_err(X[off[s]:off[s + 1]].ravel(), o.get("pos"), scale=1.0) < 1e-11
        assert rel_err(F[off[s]:off[s + 1]].ravel(), o.get("force")) < 1e-9
        assert U[s] == pytest.approx(o.potential_energy(), rel=1e-10)


def test_observables(oracle_mod):
    print("case observables", flush=True)
    st = packing.random_system(20, 0.8, 77)
    o, g = prepare(oracle_mod, st, "harmonic", "vertexVertex")
    o.calc_force_energy(); g.calc_force_energy()
    o.init_nve(1e-2); g.init_nve(1e-2)
    assert g.kinetic_energy()[0] == pytest.approx(o.kinetic_energy(), rel=1e-13)
    assert g.temperature()[0] == pytest.approx(o.temperature(), rel=1e-13)
    assert g.max_unbalanced_force()[0] == pytest.approx(o.max_unbalanced_force(), rel=1e-12)
    assert g.phi()[0] == pytest.approx(o.phi(), rel=1e-13)
    assert g.preferred_phi()[0] == pytest.approx(o.preferred_phi(), rel=1e-13)
    np.testing.assert_allclose(g.shapes(), o.shapes(), rtol=1e-12)
    g.scale_vertices(1.01); o.scale_vertices(1.01)
    assert rel_err(g.get("pos"), o.get("pos"), scale=1.0) < 1e-13
    assert rel_err(g.get("a0"), o.get("a0")) < 1e-15


@pytest.mark.parametrize("inter", INTERS)
def test_split_invariance(oracle_mod, inter):
    print(f"case split_invariance {inter}", flush=True)
    st = packing.random_system(40, 0.8, 55)
    o, g = prepare(oracle_mod, st, "wca", inter)
    o.calc_force_energy()
    ref = o.get("force")
    for split in (1, 2, 3, 4, 8):
        g.set_launch_config(split)
        g.calc_force_energy()
        assert rel_err(g.get("force"), ref) < TOL_F, split


def test_loop_tuner_keeps_the_trajectory(oracle_mod):
    """The loop-timed waves-per-tile tuner runs stretches of the caller's own NVE loop under each
    candidate shape (a shape change re-captures the loop's graph): the trajectory it leaves is the
    oracle's for the same number of steps, and the chosen shape is the one set."""
    from cudaring_amd import autotune
    print("case loop_tuner", flush=True)
    st = packing.lattice_system(16, 0.8, 17)
    o, g = prepare(oracle_mod, st, "harmonic", "vertexVertex", dt=0.005)
    o.calc_force_energy(); g.calc_force_energy()
    o.init_nve(1e-2, zero_momentum=True); g.init_nve(1e-2, zero_momentum=True)
    best, times = autotune.tune_waves_per_tile_loop(g, lambda: g.nve(20), candidates=(2, 4, 8), reps=2)
    assert best in (2, 4, 8) and sorted(times) == [2, 4, 8]
    o.nve(3 * 3 * 20)
    assert rel_err(g.get("pos"), o.get("pos"), scale=1.0) < 1e-11
    assert rel_err(g.get("vel"), o.get("vel")) < 1e-9
    assert list(g.update_count()) == [o.update_count()]


@pytest.mark.parametrize("persistent", [True, False])
def test_batched_fire_minimize(oracle_mod, persistent):
    print(f"case batched_fire_minimize persistent={persistent}", flush=True)
    # square ring counts on a dilute lattice: unjammed, every system converges
    systems = [packing.lattice_system(n, 0.5, 300 + n, perturb=0.05) for n in (4, 9, 16, 25)]
    cat = packing.concat_systems(systems)
    g = gpu_from_state(cat, potential="harmonic", interaction="vertexVertex", costs=(1, 1, 0.1, 1),
                       rings_per_system=cat["rings_per_system"])
    size = 2 * np.mean(cat["rad"])
    cut = g.set_displacement_cutoff(0.3, size)
    g.calc_neighbor_list(cut)
    g.calc_force_energy()
    dt = 0.01 * 2.0
    prm = [0.2, 0.5, 1.1, 0.99, dt, 10 * dt, 0.2]
    g.init_fire(prm, 20, 0)
    g.set_fire_kernel(persistent)
    it = g.fire_minimize(1e-10, 100000)
    fmax = g.max_unbalanced_force()
    assert np.all(fmax <= 1e-10), fmax
    U = g.potential_energy()
    for s, st in enumerate(systems):
        o = oracle_from_state(oracle_mod, st, potential="harmonic", interaction="vertexVertex", costs=(1, 1, 0.1, 1))
        o.set_displacement_cutoff(0.3, size)
        o.calc_neighbor_list(cut)
        o.calc_force_energy()
        o.init_fire(prm, 20, 0)
        ito = o.fire_minimize(1e-10, 100000)
        assert abs(int(it[s]) - ito) <= max(5, 0.02 * ito), (it[s], ito)
        assert U[s] == pytest.approx(o.potential_energy(), rel=1e-9)


def test_cfg1_nve_full_length(oracle_mod):
    """The plain NVE configuration at full length: 32 bidisperse rings (16/22, calA0 1.10,
    ka = kl = 1, kb = 0.01), harmonic vertex-vertex, phi 0.80 with uniform random centres,
    T = 1e-3 with zero net momentum, dt = 0.01 sigma, skin 0.3 sigma, 1e4 steps.  The first 100
    steps match the oracle step for step; over 1e4 steps both conserve energy to the same
    O(dt^2) level."""
    print("case cfg1_nve", flush=True)
    st = packing.random_system(32, 0.80, 20261202 + 1)
    size = 2 * float(np.mean(st["rad"]))
    sigma = 2 * np.sqrt(np.mean(st["a0"]) / np.pi)
    o, g = prepare(oracle_mod, st, "harmonic", "vertexVertex", skin=0.3, dt=0.01 * sigma)
    o.calc_force_energy(); g.calc_force_energy()
    o.init_nve(1e-3, zero_momentum=True); g.init_nve(1e-3, zero_momentum=True)
    e0 = o.potential_energy() + o.kinetic_energy()
    o.nve(100); g.nve(100)
    assert rel_err(g.get("pos"), o.get("pos"), scale=1.0) < 1e-11
    o.nve(9900); g.nve(9900)
    eo = o.potential_energy() + o.kinetic_energy()
    eg = g.potential_energy()[0] + g.kinetic_energy()[0]
    # the random start releases its overlap energy into heat; E stays within 1% of E0
    assert abs(eo - e0) < 1e-2 * abs(e0) and abs(eg - e0) < 1e-2 * abs(e0)
    assert np.all(np.isfinite(g.get("pos")))


@pytest.mark.parametrize("waves", [0, 16])
@pytest.mark.parametrize("pot,inter", [("harmonic", "vertexVertex"), ("wca", "vertexSmooth")])
def test_persistent_fire_matches_step_kernels(pot, inter, waves):
    """fire_minimize on one workgroup per system (tol 0, 30 iterations) against 30 iterations of
    the split FIRE step kernels, on a batched set with rebuilds along the way.  The two paths add
    the pair forces in different orders; at dt_max the stiff WCA trajectory amplifies that
    roundoff about 1e3-fold per 10 iterations (1e-16 at iteration 20, 7e-13 in the velocities at
    30, 2e-9 at 40), so the positions are compared at 30 iterations."""
    print(f"case persistent_fire {pot} {inter} waves={waves}", flush=True)
    systems = [relaxed_state(n, 0.8, 500 + n, final=(pot, inter)) for n in (9, 16)]
    cat = packing.concat_systems(systems)
    sets = []
    for _ in range(2):
        g = gpu_from_state(cat, potential=pot, interaction=inter, costs=(1, 1, 0.1, 1),
                           rings_per_system=cat["rings_per_system"])
        cut = g.set_displacement_cutoff(0.1, 2 * np.mean(cat["rad"]))
        g.calc_neighbor_list(cut)
        g.calc_force_energy()
        dt = 0.01 * 2 * np.sqrt(np.mean(cat["a0"]) / np.pi)  # dt_max = 10 dt stays stable
        g.init_fire([0.2, 0.5, 1.1, 0.99, dt, 10 * dt, 0.2], 20, 0)
        sets.append(g)
    a, b = sets
    a.set_fire_kernel(True, waves)  # 0: one wave per tile of the larger system; 16: the 1024-lane build
    it = a.fire_minimize(0.0, 30)
    b.fire(30)
    assert list(it) == [30, 30]
    fa, fb = a.fire_state(), b.fire_state()
    assert list(fa["numStep"]) == list(fb["numStep"])
    np.testing.assert_allclose(fa["dt"], fb["dt"], rtol=1e-13)
    np.testing.assert_allclose(fa["alpha"], fb["alpha"], rtol=1e-13)
    assert rel_err(a.get("pos"), b.get("pos"), scale=1.0) < 1e-11
    assert rel_err(a.get("vel"), b.get("vel")) < 1e-8
    assert rel_err(a.get("force"), b.get("force")) < 1e-8
    assert int(a.update_count().sum()) == int(b.update_count().sum()) and int(a.update_count().sum()) > 0


def test_nvt_workload_full_size(oracle_mod):
    """The NVT Langevin workload at its full size: 1024 bidisperse rings (16/22, 19,456 vertices),
    WCA vertex-segment, phi 0.85, T = 1e-2, gamma = 1, dt = 1e-3 sigma, skin 0.5 sigma, prepared
    on the GPU (compression + FIRE) and ordered by ring centre.  Forces and energy at the start
    to 1e-12; 150 Langevin steps, across neighbour-list rebuilds, step for step with the oracle."""
    print("case nvt_full_size", flush=True)
    st = relaxed_state(1024, 0.85, 20261202 + 2)
    st, _ = packing.spatial_order(st)
    assert int(st["nvp"].sum()) == 19456
    sigma = 2 * np.sqrt(np.mean(st["a0"]) / np.pi)
    o, g = prepare(oracle_mod, st, "wca", "vertexSmooth", dt=1e-3 * sigma, seed=11)
    o.calc_force_energy(); g.calc_force_energy()
    assert rel_err(g.get("force"), o.get("force")) < TOL_F
    assert abs(g.potential_energy()[0] - o.potential_energy()) < 1e-12 * abs(o.potential_energy())
    o.init_langevin(1e-2, 1.0); g.init_langevin(1e-2, 1.0)
    o.reset_update_count(); g.reset_update_count()
    for chunk in (50, 100):
        o.langevin(chunk); g.langevin(chunk)
        assert rel_err(g.get("pos"), o.get("pos"), scale=1.0) < 1e-11, chunk
        assert rel_err(g.get("vel"), o.get("vel")) < 1e-9, chunk
        # only a loop's last step stores the force, the energies and the ring geometry: those of
        # the final positions
        assert rel_err(g.get("force"), o.get("force")) < 1e-9, chunk
        assert rel_err(g.get("energy"), o.get("energy")) < 1e-9, chunk
        assert rel_err(g.get("penergy"), o.get("penergy")) < 1e-9, chunk
        for name in ("area", "perimeter", "theta", "length", "ppos"):
            assert rel_err(g.get(name), o.get(name), scale=1.0) < 1e-10, (name, chunk)
    assert g.update_count()[0] == o.update_count() >= 1
    g.close()


@pytest.mark.parametrize("pot,inter", [("harmonic", "vertexVertex"), ("adhesive", "vertexSmooth")])
def test_large_system_invariants(pot, inter):
    """Half a million vertices (27,000 bidisperse 16/22 rings at random centres, phi 0.8), beyond
    what the oracle finishes in seconds: the forces of a passive system sum to zero (every shape
    and pair term is translation invariant), and translating every vertex by a box fraction
    changes no force beyond roundoff."""
    print(f"case large_invariants {pot} {inter}", flush=True)
    st = packing.random_system(27000, 0.8, 20261202 + 7)
    assert int(st["nvp"].sum()) > 500000
    g = gpu_from_state(st, potential=pot, interaction=inter)
    cut = g.set_displacement_cutoff(0.5, 2 * st["rad"].mean())
    g.calc_neighbor_list(cut)
    g.calc_force_energy()
    f1 = g.get("force").reshape(-1, 2)
    assert np.abs(f1).max() > 1e-3
    assert np.all(np.abs(f1.sum(axis=0)) < 1e-12 * np.abs(f1).sum(axis=0))
    shift = np.array([0.31, 0.17]) * st["box"]
    g.set("pos", (st["pos"].reshape(-1, 2) + shift).ravel())
    g.calc_neighbor_list(cut)
    g.calc_force_energy()
    f2 = g.get("force").reshape(-1, 2)
    assert rel_err(f2.ravel(), f1.ravel()) < 1e-9
    g.close()

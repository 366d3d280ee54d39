"""The three benchmark configurations beyond the NVE and NVT ones, on the GPU against the oracle:

* batched FIRE compression: 16 of the 512 packings (N ~ U{16..256}, calA0 ~ U[1.0, 1.2]) drawn
  as the benchmark draws them, including the largest, the smallest and both calA0 ends; one
  delta phi = 1e-3 step (compressFIRE.cpp:104-131, 153) to max|F_c| <= 1e-12 in one batched
  set, every packing against the oracle minimising it alone;
* the adjoint: 256 rings at phi 0.85, harmonic, NVE at dt = 0.01 sigma checkpointed every 100;
  gradients of U(x_N) + mean p^2 / (4 pi a) along x0 and for ka, kl, kb and calA0 of one ring
  against central differences of the oracle's trajectory over 50 steps, the full
  2000 steps against the tangent-linear model (a dot-product test), and Hessian-vector products
  at a FIRE minimum against central differences of the oracle's forces;
* active + adhesive rings of 32 vertices in the vertex-segment model at phi 0.9: step for step
  with the oracle at 64 rings, the 16,384-ring system through its force balance (the shape and
  interaction forces cancel, so the total force is the self-propulsion's), and a set of 64
  systems against the oracle running each system alone.
"""
import numpy as np
import pytest

from cudaring_amd import packing, protocols
from tests.conftest import rel_err
from tests.helpers import gpu_from_state, oracle_from_state

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

# ---------------------------------------------------------------- batched FIRE compression ---
FIRE_SEED = 20261202 + 3
FIRE_PARAMS = [0.1, 0.5, 1.1, 0.99, 0.01, 0.1, 0.1]  # a_start, f_dec, f_inc, f_a, dt0, dt_max, a
FIRE_COSTS = (1.0, 1.0, 0.1, 1.0)
TOL, DPHI, MAXIT = 1e-12, 1e-3, 200000
FIRST = 100  # iterations compared step for step


def fire_sample(count=16):
    """`count` of the benchmark's 512 packings: the largest and smallest N, the calA0 extremes and
    the rest evenly spaced in order of N."""
    rng = np.random.default_rng(FIRE_SEED)
    sizes = rng.integers(16, 257, 512)
    calA0 = rng.uniform(1.0, 1.2, 512)
    order = np.argsort(sizes, kind="stable")
    pick = [int(np.argmax(sizes)), int(np.argmin(sizes)), int(np.argmin(calA0)), int(np.argmax(calA0))]
    for i in order[np.linspace(0, 511, count).round().astype(int)]:
        if len(pick) == count:
            break
        if int(i) not in pick:
            pick.append(int(i))
    return [(i, packing.lattice_system(int(sizes[i]), 0.5, FIRE_SEED + i, calA0=float(calA0[i]), perturb=0.15))
            for i in sorted(pick)]


def test_cfg3_batched_fire_compression_step(oracle_mod):
    print("case cfg3 batched FIRE compression step, 16 packings", flush=True)
    sample = fire_sample()
    systems = [st for _, st in sample]
    assert max(len(st["nvp"]) for st in systems) >= 250
    cat = packing.concat_systems(systems)
    g = gpu_from_state(cat, potential="harmonic", interaction="vertexVertex", costs=FIRE_COSTS,
                       rings_per_system=cat["rings_per_system"])
    size = 2 * float(np.mean(cat["rad"]))
    cut = g.set_displacement_cutoff(0.3, size)
    g.calc_neighbor_list(cut)
    g.calc_force_energy()
    g.init_fire(FIRE_PARAMS, 20, 0)
    g.fire_minimize(TOL, MAXIT)  # the dilute start, relaxed: the common start of both codes
    assert np.all(g.max_unbalanced_force() <= TOL)
    start = {k: g.get(k) for k in ("pos", "vel", "rad", "l0", "a0")}  # FIRE keeps its velocities
    phi = g.phi()
    g.scale_vertices(np.sqrt((phi + DPHI) / phi))
    g.calc_neighbor_list(cut)
    g.calc_force_energy()
    g.reset_last_positions()
    g.init_fire(FIRE_PARAMS, 20, 0)
    # the first iterations step for step: every packing's trajectory is the oracle's
    g.fire_minimize(0.0, FIRST)
    pos_first = g.get("pos").reshape(-1, 2)
    # then each packing to its own convergence: FIRE's restarts are decided by the sign of the
    # power P, which rounding flips once P is near zero, so the two codes' iteration counts and
    # paths separate from there; both must reach max|F_c| <= 1e-12 at the same energy
    it = g.fire_minimize(TOL, MAXIT)
    fmax, U = g.max_unbalanced_force(), g.potential_energy()
    assert np.all(fmax <= TOL), fmax
    voff, roff = g.vert_off, g.ring_off
    ratios = []
    for s, st in enumerate(systems):
        o = oracle_from_state(oracle_mod, st, potential="harmonic", interaction="vertexVertex", costs=FIRE_COSTS)
        vs, rs = slice(voff[s], voff[s + 1]), slice(roff[s], roff[s + 1])
        o.set("pos", start["pos"][2 * voff[s]:2 * voff[s + 1]])
        o.set("vel", start["vel"][2 * voff[s]:2 * voff[s + 1]])
        o.set("rad", start["rad"][vs])
        o.set("l0", start["l0"][vs])
        o.set("a0", start["a0"][rs])
        o.set_displacement_cutoff(0.3, size)
        o.calc_neighbor_list(cut)
        o.calc_force_energy()  # ring areas and centres of the relaxed start (phi, scaleVertices)
        p = o.phi()
        assert p == pytest.approx(phi[s], rel=1e-13)
        o.scale_vertices(np.sqrt((p + DPHI) / p))
        o.calc_neighbor_list(cut)
        o.calc_force_energy()
        o.init_fire(FIRE_PARAMS, 20, 0)
        o.fire(FIRST)
        L = st["box"][0]
        assert rel_err(pos_first[vs].ravel(), o.get("pos"), scale=L) < 1e-10, s
        ito = FIRST + o.fire_minimize(TOL, MAXIT)
        assert o.max_unbalanced_force() <= TOL
        assert U[s] == pytest.approx(o.potential_energy(), rel=1e-9, abs=1e-14), s
        ratios.append(int(it[s]) / ito)
    ratios = np.array(ratios)
    print("iterations GPU / oracle per packing:", np.round(ratios, 3), flush=True)
    assert np.all(np.abs(ratios - 1) < 0.35) and abs(float(np.mean(ratios)) - 1) < 0.1, ratios


# ---------------------------------------------------------------------------------- adjoint ---
ADJ_COSTS = (1.0, 1.0, 0.01, 1.0)
H = 1e-6


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device="cuda")


@pytest.fixture(scope="module")
def cfg4_state():
    from cudaring_amd import DPMSet
    s, st = protocols.prepared_state(lambda nvp: DPMSet(nvp), 256, 0.85, 20261202 + 4, costs=ADJ_COSTS,
                                     iters=1000)
    s.close()
    return st


def cfg4_set(st, dt):
    size = 2 * float(np.mean(st["rad"]))
    g = gpu_from_state(st, costs=ADJ_COSTS, dt=dt)
    cut = g.set_displacement_cutoff(0.3, size)
    g.calc_neighbor_list(cut)
    g.calc_force_energy()
    g.init_nve(1e-3, zero_momentum=True)
    return g, cut, size


def test_cfg4_adjoint_gradients_against_oracle_differences(oracle_mod, cfg4_state):
    """Gradients of U(x_N) + mean p^2 / (4 pi a) through 50 NVE steps (five checkpoint segments of
    10, recomputed in the reverse pass) against the oracle's central differences.
    The harmonic contact force has a kink at contact, so the loss is only once differentiable and
    the differences converge slowly as trajectories lengthen: on this packing, after 200 steps
    they give 3.991, 4.055, 4.059 at h = 1e-5, 1e-6, 1e-7 against the adjoint's 4.064 (the GPU's
    and the oracle's differences agree to 1e-9), after 50 steps h = 1e-7 agrees with the adjoint
    to 2e-8, and over 2000 steps the trajectory is chaotic (differences at h = 1e-7 ... 1e-10
    disagree by tens of percent).  The full length is checked by the dot-product test below."""
    print("case cfg4 adjoint against differences, 256 rings, 50 NVE steps", flush=True)
    from cudaring_amd.autograd import DPMAdjoint, l0_from_calA0, shape_parameter
    st = cfg4_state
    nr, nvp, a0 = len(st["nvp"]), st["nvp"], st["a0"]
    sigma = float(2 * np.sqrt(np.mean(a0) / np.pi))
    dt, steps, every = 0.01 * sigma, 50, 10
    g, cut, size = cfg4_set(st, dt)
    v0 = g.get("vel")
    ka0, kl0, kb0 = (np.full(nr, c) for c in ADJ_COSTS[:3])
    calA0 = np.full(nr, 1.10)
    adj = DPMAdjoint(g, cut)
    x0_t = dev(st["pos"]).requires_grad_()
    ka_t, kl_t, kb_t, c_t = (dev(a).requires_grad_() for a in (ka0, kl0, kb0, calA0))
    l0 = l0_from_calA0(c_t, nvp, a0)
    xN, _ = adj.trajectory(x0_t, dev(v0), ka_t, kl_t, kb_t, l0, steps, every=every)
    loss = adj.potential_energy(xN, ka_t, kl_t, kb_t, l0) + shape_parameter(xN, nvp, st["box"])
    loss.backward()

    o = oracle_from_state(oracle_mod, st, costs=ADJ_COSTS, dt=dt)
    o.set_displacement_cutoff(0.3, size)
    l0_base = l0.detach().cpu().numpy()

    def oracle_loss(x0=st["pos"], ka=ka0, kl=kl0, kb=kb0, l0_=l0_base):
        o.set_ring_costs(ka, kl, kb)
        o.set("l0", l0_)
        o.set("pos", x0)
        o.set("vel", v0)
        o.init_nve(0.0, read_state=True)
        o.calc_neighbor_list(cut)
        o.calc_force_energy()
        o.nve(steps)
        o.calc_force_energy()
        return o.potential_energy() + float(np.mean(o.shapes()[:, 2]))

    def cdiff(f, h=1e-7):
        return (f(h) - f(-h)) / (2 * h)

    rng = np.random.default_rng(4)
    d = rng.standard_normal(2 * g.nv)
    # contacts that open or close during the run sit on the kink of the harmonic force: a
    # difference whose perturbation moves such an event across the kink converges only as O(h)
    # (on this packing 0.18855, 0.18608, 0.185071 at h = 1e-6, 1e-7, 1e-8 against the adjoint's
    # 0.185072), so the positions are perturbed by 1e-8 (rounding of the loss: about 1e-8)
    fd = cdiff(lambda h: oracle_loss(x0=st["pos"] + h * d), 1e-8)
    gx = float(x0_t.grad.cpu().numpy() @ d)
    assert abs(gx - fd) <= 2e-5 * abs(fd), (gx, fd)
    p = 7
    for name, base, grad in (("ka", ka0, ka_t.grad), ("kl", kl0, kl_t.grad), ("kb", kb0, kb_t.grad)):
        def f(h, base=base, name=name):
            b = base.copy()
            b[p] += h
            return oracle_loss(**{name: b})
        fd = cdiff(f, 1e-6)
        # differences of a loss of order 1 at h = 1e-6 carry rounding of about 1e-9
        assert abs(float(grad[p]) - fd) <= 1e-4 * abs(fd) + 1e-8, (name, float(grad[p]), fd)

    def fc(h):
        e = np.zeros(nr)
        e[p] = h
        return oracle_loss(l0_=l0_from_calA0(dev(calA0 + e), nvp, a0).cpu().numpy())
    fd = cdiff(fc, 1e-6)
    assert abs(float(c_t.grad[p]) - fd) <= 1e-4 * abs(fd) + 1e-8, (float(c_t.grad[p]), fd)
    g.close()


def test_cfg4_full_length_dot_product(cfg4_state):
    """2000 NVE steps checkpointed every 100: the reverse pass (segments recomputed, rows rebuilt)
    against the tangent-linear model of velocity Verlet built from Hessian-vector products along
    the same trajectory, delta v' = delta v - dt/2 H(x) delta x, delta x1 = delta x + dt delta v',
    delta v1 = delta v' - dt/2 H(x1) delta x1: <a, J d> = <J^T a, d> for the linear functional
    a . x_N + b . v_N and a random direction d of (x_0, v_0)."""
    print("case cfg4 dot-product test, 256 rings, 2000 NVE steps", flush=True)
    st = cfg4_state
    sigma = float(2 * np.sqrt(np.mean(st["a0"]) / np.pi))
    dt, steps, every = 0.01 * sigma, 2000, 100
    g, cut, size = cfg4_set(st, dt)
    nv = g.nv
    x0, v0 = g.get("pos"), g.get("vel")
    rng = np.random.default_rng(8)
    ax, av = rng.standard_normal(2 * nv), rng.standard_normal(2 * nv)
    dx, dv = rng.standard_normal(2 * nv), rng.standard_normal(2 * nv)

    def soa(a):
        return dev(np.concatenate([a[0::2], a[1::2]]))

    def aos(t):
        a = t.cpu().numpy()
        out = np.empty(2 * nv)
        out[0::2], out[1::2] = a[:nv], a[nv:]
        return out

    g.nve_forward(steps, every)
    gx0, gv0 = (torch.empty(2 * nv, dtype=torch.float64, device="cuda") for _ in range(2))
    torch.cuda.synchronize()
    ax_t, av_t = soa(ax), soa(av)
    g.nve_backward(ax_t.data_ptr(), av_t.data_ptr(), gx0.data_ptr(), gv0.data_ptr(), 0, 0)
    adjoint = float(aos(gx0) @ dx + aos(gv0) @ dv)

    def hvp(u):
        ux, uy = dev(u[0::2]), dev(u[1::2])
        ox, oy = torch.empty_like(ux), torch.empty_like(uy)
        torch.cuda.synchronize()
        g.force_hvp(ux.data_ptr(), uy.data_ptr(), ox.data_ptr(), oy.data_ptr())
        out = np.empty(2 * nv)
        out[0::2], out[1::2] = ox.cpu().numpy(), oy.cpu().numpy()
        return out

    g.set("pos", x0)
    g.set("vel", v0)
    g.calc_neighbor_list(cut)
    g.calc_force_energy()
    h = 0.5 * dt
    for _ in range(steps):
        dvh = dv - h * hvp(dx)
        g.nve(1)
        dx = dx + dt * dvh
        dv = dvh - h * hvp(dx)
    tangent = float(ax @ dx + av @ dv)
    assert abs(adjoint - tangent) <= 1e-9 * max(abs(tangent), 1.0), (adjoint, tangent)
    g.close()


def test_cfg4_hvp_at_fire_minimum(oracle_mod, cfg4_state):
    print("case cfg4 Hessian-vector products at a FIRE minimum", flush=True)
    st = cfg4_state
    size = 2 * float(np.mean(st["rad"]))
    g = gpu_from_state(st, costs=ADJ_COSTS)
    cut = g.set_displacement_cutoff(0.3, size)
    g.calc_neighbor_list(cut)
    g.calc_force_energy()
    sigma = float(2 * np.sqrt(np.mean(st["a0"]) / np.pi))
    dt = 0.01 * sigma
    g.init_fire([0.2, 0.5, 1.1, 0.99, dt, 10 * dt, 0.2], 20, 0)
    g.fire_minimize(1e-12, 400000)
    assert g.max_unbalanced_force()[0] <= 1e-9
    g.calc_neighbor_list(cut)
    g.calc_force_energy()
    xmin = g.get("pos")
    o = oracle_from_state(oracle_mod, st, costs=ADJ_COSTS)
    o.set("pos", xmin)
    o.set_displacement_cutoff(0.3, size)
    o.calc_neighbor_list(cut)
    rng = np.random.default_rng(12)
    nv = g.nv
    # At a minimum many contacts rest at zero overlap (rings with no force left), on the kink of
    # the harmonic force; a difference at h = 1e-6 across such a contact averages its two sides.
    # The directions keep every vertex within 1e-4 of a kink still, so both differences stay on
    # the branch the Hessian-vector product is taken on.
    rows, _, _ = g.neighbors()
    pos, rad = xmin.reshape(-1, 2), st["rad"]
    L = st["box"]
    still = np.zeros(nv, dtype=bool)
    for v, row in enumerate(rows):
        for j in row:
            dvec = pos[v] - pos[j]
            dvec -= L * np.round(dvec / L)
            if abs(np.hypot(*dvec) / (rad[v] + rad[j]) - 1) < 1e-4:
                still[v] = still[j] = True
    print(f"vertices held still at kinks: {int(still.sum())} of {nv}", flush=True)
    for _ in range(2):
        u = rng.standard_normal(2 * nv)
        u.reshape(-1, 2)[still] = 0.0
        ux, uy = dev(u[0::2]), dev(u[1::2])
        ox, oy = torch.empty_like(ux), torch.empty_like(uy)
        torch.cuda.synchronize()
        g.force_hvp(ux.data_ptr(), uy.data_ptr(), ox.data_ptr(), oy.data_ptr())
        hu = np.empty(2 * nv)
        hu[0::2], hu[1::2] = ox.cpu().numpy(), oy.cpu().numpy()
        o.set("pos", xmin + H * u)
        o.calc_force_energy()
        fp = o.get("force")
        o.set("pos", xmin - H * u)
        o.calc_force_energy()
        fm = o.get("force")
        fd = -(fp - fm) / (2 * H)
        assert np.max(np.abs(hu - fd)) <= 1e-6 * np.abs(fd).max()
        # at a minimum the Hessian is positive semi-definite along any direction
        assert u @ hu >= -1e-9 * np.abs(u) @ np.abs(hu)
    g.close()


# --------------------------------------------------------------------- active + adhesive ---
ACT_COSTS = (1e5, 20.0, 10.0, 1.0)  # ea, el, eb, ec of the active driver (runActive.cpp:39)


def active_state(nrings, seed, phi=0.9, dphi=0.05, iters=600):
    """Size-bidisperse rings of 32 vertices compressed to preferred phi with harmonic FIRE, then
    relaxed with the production interaction and shape costs (as the benchmark prepares them)."""
    from cudaring_amd import DPMSet
    st = packing.size_bidisperse_system(nrings, 0.5, seed)
    s = DPMSet(st["nvp"])
    s.load_state(st)
    s.set_types("harmonic", "vertexVertex")
    s.set_energy_costs(1, 1, 0.01, 1)
    protocols.compress(s, 0.5, phi, dphi=dphi, iters=iters)
    s.set_types("adhesive", "vertexSmooth")
    s.set_attraction(0.1, 0.2)
    s.set_energy_costs(*ACT_COSTS)
    size = 2 * float(np.mean(s.get("rad")))
    s.calc_neighbor_list(s.set_displacement_cutoff(0.5, size))
    s.calc_force_energy()
    protocols.fire_relax(s, 0.002 * size, 2 * iters)
    out = dict(st)
    for k in ("pos", "rad", "l0", "a0"):
        out[k] = s.get(k)
    s.close()
    return out


def active_setup(obj, st, size, sigma, angles, oracle=False, system_id=0):
    obj.set_types("adhesive", "vertexSmooth", "harmonic", "active")
    obj.set_attraction(0.1, 0.2)
    obj.set_energy_costs(*ACT_COSTS)
    obj.set_self_propulsion(0.1 / sigma, 100.0 * sigma)
    obj.set("pangle", angles)
    obj.set_dt(1e-3 * sigma)
    if oracle:
        obj.set_rng(9, system_id)
    else:
        obj.set_seed(9)
    cut = obj.set_displacement_cutoff(0.5, size)
    obj.calc_neighbor_list(cut)
    obj.calc_force_energy()
    obj.init_langevin(1e-4, 1.0 / sigma)
    return cut


def test_cfg5_active_adhesive_step_for_step(oracle_mod):
    print("case cfg5 active + adhesive, 64 rings x 32 vertices", flush=True)
    st = active_state(64, 20261202 + 5)
    size = 2 * float(np.mean(st["rad"]))
    sigma = float(2 * np.sqrt(np.mean(st["a0"]) / np.pi))
    angles = np.random.default_rng(2).uniform(-np.pi, np.pi, len(st["nvp"]))
    g = gpu_from_state(st)
    o = oracle_from_state(oracle_mod, st)
    active_setup(g, st, size, sigma, angles)
    active_setup(o, st, size, sigma, angles, oracle=True)
    # with ea = 1e5 and areas at their rest values the area force is a small difference of terms
    # of size ea l0 / a0: the error is normed by that scale
    term = ACT_COSTS[0] * float(np.mean(st["l0"])) / float(np.mean(st["a0"]))
    assert rel_err(g.get("force"), o.get("force"), scale=term) < 1e-13
    for chunk in (5, 25):
        g.langevin(chunk)
        o.langevin(chunk)
        assert rel_err(g.get("pos"), o.get("pos"), scale=1.0) < 1e-11
        assert rel_err(g.get("pangle"), o.get("pangle"), scale=np.pi) < 1e-12
    assert rel_err(g.get("force"), o.get("force"), scale=term) < 1e-11
    assert g.potential_energy()[0] == pytest.approx(o.potential_energy(), rel=1e-10)
    g.close()


def test_cfg5_adjoint_through_active_adhesive_nve(oracle_mod):
    """The reverse pass on the active + adhesive vertex-segment model (16 rings x 32 vertices):
    40 NVE steps with self-propulsion, whose ring angles advance by one draw per step, recorded
    with checkpoints every 15 steps (each recomputed segment restarts from its checkpoint's angles
    and RNG counters); gradients of L = c . x_N + d . v_N with respect to x0, v0 and per-ring ka,
    kl, kb against the oracle's central differences.  The self-propulsion force depends on
    neither positions nor parameters, so it enters the gradients through the trajectory only."""
    print("case cfg5 adjoint, 16 rings x 32 vertices, 40 active NVE steps", flush=True)
    torch = pytest.importorskip("torch")
    st = active_state(16, 20261202 + 7)
    nr, nv = len(st["nvp"]), int(np.sum(st["nvp"]))
    size = 2 * float(np.mean(st["rad"]))
    sigma = float(2 * np.sqrt(np.mean(st["a0"]) / np.pi))
    dt, steps, every = 1e-3 * sigma, 40, 15
    rng = np.random.default_rng(8)
    angles = rng.uniform(-np.pi, np.pi, nr)
    v0 = 1e-3 * rng.standard_normal(2 * nv)
    ka, kl, kb = (c * (0.8 + 0.4 * rng.random(nr)) for c in ACT_COSTS[:3])

    def setup(obj, x0, v, ka_, kl_, kb_, oracle):
        obj.set_types("adhesive", "vertexSmooth", "harmonic", "active")
        obj.set_attraction(0.1, 0.2)
        obj.set_energy_costs(*ACT_COSTS)
        obj.set_ring_costs(ka_, kl_, kb_)
        obj.set_self_propulsion(0.1 / sigma, 100.0 * sigma)
        obj.set_dt(dt)
        if oracle:
            obj.set_rng(9, 0)
        else:
            obj.set_seed(9)
        obj.set("pangle", angles)
        obj.set("pos", x0)
        obj.set("vel", v)
        obj.init_nve(0.0, read_state=True)
        cut = obj.set_displacement_cutoff(0.5, size)
        obj.calc_neighbor_list(cut)
        obj.calc_force_energy()

    g = gpu_from_state(st)
    setup(g, st["pos"], v0, ka, kl, kb, False)
    g.nve_forward(steps, every)
    xg, vg, ag = g.get("pos"), g.get("vel"), g.get("pangle")
    o = oracle_from_state(oracle_mod, st)

    def final(x0=st["pos"], v=v0, ka_=ka, kl_=kl, kb_=kb):
        setup(o, x0, v, ka_, kl_, kb_, True)
        o.nve(steps)
        return o.get("pos"), o.get("vel")

    xo, vo = final()
    assert rel_err(xg, xo, scale=1.0) < 1e-11
    assert rel_err(ag, o.get("pangle"), scale=np.pi) < 1e-13
    cx, cv = rng.standard_normal(2 * nv), rng.standard_normal(2 * nv)

    def soa(a):
        return torch.as_tensor(np.concatenate([a[0::2], a[1::2]]), device="cuda")

    def aos(t):
        a = t.cpu().numpy()
        return np.stack([a[:nv], a[nv:]], 1).ravel()

    axt, avt = soa(cx), soa(cv)
    gx, gv = (torch.empty(2 * nv, dtype=torch.float64, device="cuda") for _ in range(2))
    gring = torch.empty(3 * nr, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    g.nve_backward(axt.data_ptr(), avt.data_ptr(), gx.data_ptr(), gv.data_ptr(), gring.data_ptr(), 0)
    gx, gv, gring = aos(gx), aos(gv), gring.cpu().numpy()
    # the set is left where the forward pass ended, angles included
    np.testing.assert_array_equal(g.get("pos"), xg)
    np.testing.assert_array_equal(g.get("pangle"), ag)

    def L(**kw):
        x, v = final(**kw)
        return cx @ x + cv @ v

    d = rng.standard_normal(2 * nv)
    # contacts opening or closing across a difference step sit on the adhesive force's kinks:
    # positions are perturbed by 1e-8 (as for the harmonic configuration above)
    fd = (L(x0=st["pos"] + 1e-8 * d) - L(x0=st["pos"] - 1e-8 * d)) / 2e-8
    assert abs(gx @ d - fd) <= 2e-5 * abs(fd), (gx @ d, fd)
    fd = (L(v=v0 + 1e-8 * d) - L(v=v0 - 1e-8 * d)) / 2e-8
    assert abs(gv @ d - fd) <= 2e-5 * abs(fd), (gv @ d, fd)
    p = 5
    for q, (name, base) in enumerate((("ka_", ka), ("kl_", kl), ("kb_", kb))):
        h = 1e-6 * abs(base[p])
        up, dn = base.copy(), base.copy()
        up[p] += h
        dn[p] -= h
        fd = (L(**{name: up}) - L(**{name: dn})) / (2 * h)
        assert abs(gring[q * nr + p] - fd) <= 1e-4 * abs(fd) + 1e-9, (name, gring[q * nr + p], fd)
    g.close()


def test_cfg5_full_size_force_balance():
    print("case cfg5 full size, 16384 rings x 32 vertices", flush=True)
    st = active_state(16384, 20261202 + 5, dphi=0.1, iters=300)
    size = 2 * float(np.mean(st["rad"]))
    sigma = float(2 * np.sqrt(np.mean(st["a0"]) / np.pi))
    angles = np.random.default_rng(3).uniform(-np.pi, np.pi, len(st["nvp"]))
    g = gpu_from_state(st)
    active_setup(g, st, size, sigma, angles)
    for _ in range(2):
        F = g.get("force").reshape(-1, 2)
        th = g.get("pangle")
        f0 = 0.1 / sigma
        drive = f0 * np.stack([np.cos(th), np.sin(th)], 1) * st["nvp"][:, None]
        scale = np.abs(F).sum(0).max()
        assert np.all(np.abs(F.sum(0) - drive.sum(0)) <= 1e-12 * scale), (F.sum(0), drive.sum(0))
        g.langevin(20)
        g.calc_force_energy()
    assert np.all(np.isfinite(g.get("pos")))
    g.close()


def test_cfg5_batched_systems_match_each_alone(oracle_mod):
    print("case cfg5 batched, 64 systems", flush=True)
    rng = np.random.default_rng(20261202 + 6)
    sizes = rng.integers(8, 33, 64)
    systems = [active_state(int(n), 20261202 + 100 + i, dphi=0.1, iters=300) for i, n in enumerate(sizes)]
    cat = packing.concat_systems(systems)
    size = 2 * float(np.mean(cat["rad"]))
    sigma = float(2 * np.sqrt(np.mean(cat["a0"]) / np.pi))
    angles = rng.uniform(-np.pi, np.pi, len(cat["nvp"]))
    g = gpu_from_state(cat, rings_per_system=cat["rings_per_system"])
    active_setup(g, cat, size, sigma, angles)
    g.langevin(20)
    X = g.get("pos").reshape(-1, 2)
    U = g.potential_energy()
    voff, roff = g.vert_off, g.ring_off
    for s, st in enumerate(systems):
        o = oracle_from_state(oracle_mod, st)
        active_setup(o, st, size, sigma, angles[roff[s]:roff[s + 1]], oracle=True, system_id=s)
        o.langevin(20)
        assert rel_err(X[voff[s]:voff[s + 1]].ravel(), o.get("pos"), scale=1.0) < 1e-11, s
        assert U[s] == pytest.approx(o.potential_energy(), rel=1e-9), s
    g.close()

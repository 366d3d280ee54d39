"""Adjoint entry points on the GPU: Hessian-vector products, energy parameter gradients and the
reverse pass of NVE, against central finite differences of the oracle (h = 1e-6) and the
identities the exact derivatives satisfy (symmetry of H, uniform-cost limit)."""
import numpy as np
import pytest

from cudaring_amd import packing
from tests.helpers import gpu_from_state, oracle_from_state

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
H = 1e-6
COSTS = (1.0, 1.0, 0.01, 1.0)


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device="cuda")


def soft_state(nrings=9, phi=0.62, seed=4):
    """A packing with ring contacts and deformed rings (non-zero forces everywhere)."""
    st = packing.lattice_system(nrings, phi, seed, perturb=0.1)
    rng = np.random.default_rng(seed)
    pos = st["pos"].reshape(-1, 2)
    pos += 0.02 * (rng.random(pos.shape) - 0.5)
    st["pos"] = pos.ravel()
    st["a0"] = st["a0"] * (1 + 0.05 * rng.random(len(st["a0"])))
    return st


def ring_costs(n, seed=1):
    rng = np.random.default_rng(seed)
    return (0.5 + rng.random(n), 0.5 + rng.random(n), 0.005 + 0.02 * rng.random(n))


def gpu_hvp(g, u):
    nv = g.nv
    ux, uy = dev(u[0::2]), dev(u[1::2])
    ox, oy = torch.empty_like(ux), torch.empty_like(uy)
    torch.cuda.synchronize()
    g.force_hvp(ux.data_ptr(), uy.data_ptr(), ox.data_ptr(), oy.data_ptr())
    out = np.empty(2 * nv)
    out[0::2], out[1::2] = ox.cpu().numpy(), oy.cpu().numpy()
    return out


def oracle_force(o, pos):
    o.set("pos", pos)
    o.calc_force_energy()
    return o.get("force")


@pytest.mark.parametrize("pot", ["harmonic", "wca", "lj", "adhesive"])
def test_hvp_matches_oracle_fd(oracle_mod, pot):
    st = soft_state()
    ka, kl, kb = ring_costs(len(st["nvp"]))
    g = gpu_from_state(st, potential=pot, costs=COSTS)
    g.set_ring_costs(ka, kl, kb)
    cut = g.set_displacement_cutoff(0.5, 2 * float(np.mean(st["rad"])))
    g.calc_neighbor_list(cut)
    g.calc_force_energy()
    o = oracle_from_state(oracle_mod, st, potential=pot, costs=COSTS)
    o.set_ring_costs(ka, kl, kb)
    o.set_displacement_cutoff(0.5, 2 * float(np.mean(st["rad"])))
    o.calc_neighbor_list(cut)
    rng = np.random.default_rng(7)
    u = rng.standard_normal(2 * g.nv)
    hu = gpu_hvp(g, u)
    fd = -(oracle_force(o, st["pos"] + H * u) - oracle_force(o, st["pos"] - H * u)) / (2 * H)
    scale = np.abs(fd).max()
    assert scale > 1e-3
    assert np.max(np.abs(hu - fd)) / scale < 1e-6, (pot, np.max(np.abs(hu - fd)) / scale)
    # exact symmetry of the Hessian
    w = rng.standard_normal(2 * g.nv)
    hw = gpu_hvp(g, w)
    assert abs(w @ hu - u @ hw) <= 1e-10 * max(abs(w @ hu), 1.0)


def test_energy_param_grad_matches_oracle_fd(oracle_mod):
    st = soft_state(seed=8)
    nr = len(st["nvp"])
    ka, kl, kb = ring_costs(nr, 3)
    g = gpu_from_state(st, costs=COSTS)
    g.set_ring_costs(ka, kl, kb)
    cut = g.set_displacement_cutoff(0.5, 2 * float(np.mean(st["rad"])))
    g.calc_neighbor_list(cut)
    g.calc_force_energy()
    gring = torch.empty(3 * nr, dtype=torch.float64, device="cuda")
    gl0 = torch.empty(g.nv, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    g.energy_param_grad(gring.data_ptr(), gl0.data_ptr())
    gring, gl0 = gring.cpu().numpy(), gl0.cpu().numpy()
    o = oracle_from_state(oracle_mod, st, costs=COSTS)
    o.calc_neighbor_list(cut)

    def U(ka_, kl_, kb_, l0_):
        o.set_ring_costs(ka_, kl_, kb_)
        o.set("l0", l0_)
        o.calc_force_energy()
        return o.potential_energy()

    for p in (0, nr // 2, nr - 1):
        for q, base in enumerate((ka, kl, kb)):
            up, dn = base.copy(), base.copy()
            up[p] += H
            dn[p] -= H
            args_up = [ka, kl, kb]
            args_dn = [ka, kl, kb]
            args_up[q], args_dn[q] = up, dn
            fd = (U(*args_up, st["l0"]) - U(*args_dn, st["l0"])) / (2 * H)
            assert abs(gring[q * nr + p] - fd) <= 1e-6 * max(abs(fd), 1e-3), (q, p, gring[q * nr + p], fd)
    for v in (0, 5, g.nv - 1):
        up, dn = st["l0"].copy(), st["l0"].copy()
        up[v] += H
        dn[v] -= H
        fd = (U(ka, kl, kb, up) - U(ka, kl, kb, dn)) / (2 * H)
        assert abs(gl0[v] - fd) <= 1e-6 * max(abs(fd), 1e-3), (v, gl0[v], fd)


def nve_setup(st, ka, kl, kb, dt):
    g = gpu_from_state(st, costs=COSTS, dt=dt)
    g.set_ring_costs(ka, kl, kb)
    size = 2 * float(np.mean(st["rad"]))
    cut = g.set_displacement_cutoff(0.3, size)
    return g, cut, size


def oracle_final(oracle_mod, st, ka, kl, kb, dt, size, cut, steps, x0, v0, l0):
    o = oracle_from_state(oracle_mod, st, costs=COSTS, dt=dt)
    o.set_ring_costs(ka, kl, kb)
    o.set("l0", l0)
    o.set("pos", x0)
    o.set("vel", v0)
    o.set_displacement_cutoff(0.3, size)
    o.init_nve(0.0, read_state=True)
    o.calc_neighbor_list(cut)
    o.calc_force_energy()
    o.nve(steps)
    o.calc_force_energy()
    return o.get("pos"), o.get("vel"), o.potential_energy()


@pytest.mark.parametrize("vscale", [0.05, 0.5])
def test_nve_backward_matches_oracle_fd(oracle_mod, vscale):
    """vscale 0.05: one Verlet list serves each recomputed segment; 0.5: vertices move past the
    skin within a segment, so every position runs the rebuild protocol."""
    st = soft_state(seed=12)
    nr, dt, steps, every = len(st["nvp"]), 5e-3, 120, 50
    ka, kl, kb = ring_costs(nr, 5)
    rng = np.random.default_rng(2)
    v0 = vscale * rng.standard_normal(2 * len(st["rad"]))
    g, cut, size = nve_setup(st, ka, kl, kb, dt)
    g.set("vel", v0)
    g.calc_neighbor_list(cut)
    g.calc_force_energy()
    g.nve_forward(steps, every)
    xg, vg = g.get("pos"), g.get("vel")
    xo, vo, _ = oracle_final(oracle_mod, st, ka, kl, kb, dt, size, cut, steps, st["pos"], v0, st["l0"])
    np.testing.assert_allclose(xg, xo, rtol=0, atol=1e-10)
    # L = c_x . x_N + c_v . v_N
    cx, cv = rng.standard_normal(2 * g.nv), rng.standard_normal(2 * g.nv)
    nv = g.nv
    ax = dev(np.concatenate([cx[0::2], cx[1::2]]))
    av = dev(np.concatenate([cv[0::2], cv[1::2]]))
    gx, gv = torch.empty_like(ax), torch.empty_like(ax)
    gring = torch.empty(3 * nr, dtype=torch.float64, device="cuda")
    gl0 = torch.empty(nv, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    g.nve_backward(ax.data_ptr(), av.data_ptr(), gx.data_ptr(), gv.data_ptr(), gring.data_ptr(), gl0.data_ptr())
    gx, gv = gx.cpu().numpy(), gv.cpu().numpy()
    gx = np.stack([gx[:nv], gx[nv:]], 1).ravel()
    gv = np.stack([gv[:nv], gv[nv:]], 1).ravel()
    gring, gl0 = gring.cpu().numpy(), gl0.cpu().numpy()
    # the set is left at the end of the forward pass
    np.testing.assert_array_equal(g.get("pos"), xg)

    def L(x0=st["pos"], v=v0, ka_=ka, kl_=kl, kb_=kb, l0=st["l0"]):
        x, vv, _ = oracle_final(oracle_mod, st, ka_, kl_, kb_, dt, size, cut, steps, x0, v, l0)
        return cx @ x + cv @ vv

    d = rng.standard_normal(2 * nv)
    fd = (L(x0=st["pos"] + H * d) - L(x0=st["pos"] - H * d)) / (2 * H)
    assert abs(gx @ d - fd) <= 1e-6 * abs(fd), (gx @ d, fd)
    fd = (L(v=v0 + H * d) - L(v=v0 - H * d)) / (2 * H)
    assert abs(gv @ d - fd) <= 1e-6 * abs(fd), (gv @ d, fd)
    p = 1
    for q, base in enumerate((ka, kl, kb)):
        up, dn = base.copy(), base.copy()
        up[p] += H
        dn[p] -= H
        kw_up = {"ka_": ka, "kl_": kl, "kb_": kb}
        kw_dn = dict(kw_up)
        key = ("ka_", "kl_", "kb_")[q]
        kw_up[key], kw_dn[key] = up, dn
        fd = (L(**kw_up) - L(**kw_dn)) / (2 * H)
        assert abs(gring[q * nr + p] - fd) <= 1e-5 * max(abs(fd), 1e-4), (q, gring[q * nr + p], fd)
    dl = rng.standard_normal(nv)
    fd = (L(l0=st["l0"] + H * dl) - L(l0=st["l0"] - H * dl)) / (2 * H)
    assert abs(gl0 @ dl - fd) <= 1e-5 * abs(fd), (gl0 @ dl, fd)


def test_autograd_energy_and_shape(oracle_mod):
    from cudaring_amd.autograd import DPMAdjoint, l0_from_calA0, shape_parameter
    st = soft_state(seed=21)
    nr, dt, steps = len(st["nvp"]), 5e-3, 100
    g, cut, size = nve_setup(st, *ring_costs(nr, 9), dt)
    adj = DPMAdjoint(g, cut)
    rng = np.random.default_rng(3)
    x0 = dev(st["pos"]).requires_grad_()
    v0 = dev(0.05 * rng.standard_normal(2 * g.nv))
    ka, kl, kb = (dev(a).requires_grad_() for a in ring_costs(nr, 9))
    calA0 = dev(np.full(nr, 1.1)).requires_grad_()
    a0 = st["a0"]

    def loss(x0_, calA0_):
        l0 = l0_from_calA0(calA0_, st["nvp"], a0)
        xN, vN = adj.trajectory(x0_, v0, ka, kl, kb, l0, steps, every=40)
        return adj.potential_energy(xN, ka, kl, kb, l0) + shape_parameter(xN, st["nvp"], st["box"])

    L = loss(x0, calA0)
    L.backward()
    d = dev(rng.standard_normal(2 * g.nv))
    with torch.no_grad():
        fd = (loss(x0 + H * d, calA0) - loss(x0 - H * d, calA0)) / (2 * H)
    assert abs(float(x0.grad @ d) - float(fd)) <= 1e-5 * abs(float(fd)), (float(x0.grad @ d), float(fd))
    e = torch.zeros(nr, dtype=torch.float64, device="cuda")
    e[2] = 1.0
    with torch.no_grad():
        fd = (loss(x0, calA0 + H * e) - loss(x0, calA0 - H * e)) / (2 * H)
    assert abs(float(calA0.grad[2]) - float(fd)) <= 1e-5 * max(abs(float(fd)), 1e-4)


def test_two_trajectories_in_one_loss():
    """Two trajectories of one set in one loss: each node's gradients equal those of its
    trajectory differentiated alone (the second forward replaced the first one's tape)."""
    from cudaring_amd.autograd import DPMAdjoint
    st = soft_state(seed=23)
    nr, dt, steps = len(st["nvp"]), 5e-3, 60
    g, cut, size = nve_setup(st, *ring_costs(nr, 5), dt)
    adj = DPMAdjoint(g, cut)
    rng = np.random.default_rng(11)
    ka, kl, kb = (dev(a) for a in ring_costs(nr, 5))
    l0 = dev(st["l0"])
    xa = dev(st["pos"])
    xb = dev(st["pos"] + 0.01 * rng.standard_normal(2 * g.nv))
    v0 = dev(0.05 * rng.standard_normal(2 * g.nv))
    wa, wb = dev(rng.standard_normal(2 * g.nv)), dev(rng.standard_normal(2 * g.nv))

    def alone(x, w):
        x = x.clone().requires_grad_()
        xN, _ = adj.trajectory(x, v0, ka, kl, kb, l0, steps, every=25)
        (xN @ w).backward()
        return x.grad.clone()

    ga, gb = alone(xa, wa), alone(xb, wb)
    xa_, xb_ = xa.clone().requires_grad_(), xb.clone().requires_grad_()
    xNa, _ = adj.trajectory(xa_, v0, ka, kl, kb, l0, steps, every=25)
    xNb, _ = adj.trajectory(xb_, v0, ka, kl, kb, l0, steps, every=25)
    (xNa @ wa + xNb @ wb).backward()
    scale = max(float(ga.abs().max()), float(gb.abs().max()))
    assert float((xa_.grad - ga).abs().max()) <= 1e-12 * scale
    assert float((xb_.grad - gb).abs().max()) <= 1e-12 * scale


@pytest.mark.parametrize("pot", ["harmonic", "wca", "lj", "adhesive"])
def test_smooth_hvp_matches_oracle_fd(oracle_mod, pot):
    """The vertex-segment interaction's Hessian-vector product (every branch: segment interiors,
    convex corners, concave notches) against central differences of the oracle's forces, on a
    packing relaxed with the same potential (no deep overlaps, where the stiff potentials'
    curvature makes differences meaningless); the WCA packing is then thermalized (relaxed, it
    parks pairs at the cutoff, where the force's derivative jumps within the difference step)."""
    print(f"case smooth hvp {pot}", flush=True)
    from tests.helpers import relaxed_state, thermalized_state
    st = relaxed_state(12, 0.8, 31, final=(pot, "vertexSmooth"))
    if pot == "wca":
        st = thermalized_state(st, pot, "vertexSmooth")
    ka, kl, kb = ring_costs(len(st["nvp"]), 2)
    g = gpu_from_state(st, potential=pot, interaction="vertexSmooth", costs=COSTS)
    g.set_ring_costs(ka, kl, kb)
    cut = g.set_displacement_cutoff(0.5, 2 * float(np.mean(st["rad"])))
    g.calc_neighbor_list(cut)
    g.calc_force_energy()
    o = oracle_from_state(oracle_mod, st, potential=pot, interaction="vertexSmooth", costs=COSTS)
    o.set_ring_costs(ka, kl, kb)
    o.set_displacement_cutoff(0.5, 2 * float(np.mean(st["rad"])))
    o.calc_neighbor_list(cut)
    rng = np.random.default_rng(17)
    for _ in range(2):
        u = rng.standard_normal(2 * g.nv)
        hu = gpu_hvp(g, u)
        fd = -(oracle_force(o, st["pos"] + H * u) - oracle_force(o, st["pos"] - H * u)) / (2 * H)
        scale = np.abs(fd).max()
        assert np.max(np.abs(hu - fd)) / scale < 1e-6, (pot, np.max(np.abs(hu - fd)) / scale)
    # the interaction part alone is checked too (shape costs zero): its Hessian is symmetric
    g.set_ring_costs(0 * ka, 0 * kl, 0 * kb)
    w = rng.standard_normal(2 * g.nv)
    hw, hu = gpu_hvp(g, w), gpu_hvp(g, u)
    assert abs(w @ hu - u @ hw) <= 1e-9 * max(abs(w @ hu), 1e-6)


def test_smooth_nve_backward_matches_oracle_fd(oracle_mod):
    """Gradients through 60 NVE steps of WCA vertex-segment rings (checkpoints every 20) against
    the oracle's central differences."""
    print("case smooth nve backward", flush=True)
    from tests.helpers import relaxed_state, thermalized_state
    st = thermalized_state(relaxed_state(9, 0.78, 12), "wca", "vertexSmooth")
    nr, dt, steps = len(st["nvp"]), 2e-3, 60
    ka, kl, kb = ring_costs(nr, 6)
    size = 2 * float(np.mean(st["rad"]))
    g = gpu_from_state(st, potential="wca", interaction="vertexSmooth", costs=COSTS, dt=dt)
    g.set_ring_costs(ka, kl, kb)
    cut = g.set_displacement_cutoff(0.5, size)
    g.calc_neighbor_list(cut)
    g.calc_force_energy()
    g.init_nve(1e-3, zero_momentum=True)
    x0, v0 = g.get("pos"), g.get("vel")
    g.nve_forward(steps, 20)
    rng = np.random.default_rng(5)
    nv = g.nv
    ax, av = rng.standard_normal(2 * nv), rng.standard_normal(2 * nv)

    def soa(a):
        return dev(np.concatenate([a[0::2], a[1::2]]))

    gx, gv = (torch.empty(2 * nv, dtype=torch.float64, device="cuda") for _ in range(2))
    axt, avt = soa(ax), soa(av)
    torch.cuda.synchronize()
    g.nve_backward(axt.data_ptr(), avt.data_ptr(), gx.data_ptr(), gv.data_ptr(), 0, 0)
    gxa = np.empty(2 * nv)
    gxa[0::2], gxa[1::2] = gx.cpu().numpy()[:nv], gx.cpu().numpy()[nv:]
    o = oracle_from_state(oracle_mod, st, potential="wca", interaction="vertexSmooth", costs=COSTS, dt=dt)
    o.set_ring_costs(ka, kl, kb)
    o.set_displacement_cutoff(0.5, size)

    def L(x):
        o.set("pos", x)
        o.set("vel", v0)
        o.init_nve(0.0, read_state=True)
        o.calc_neighbor_list(cut)
        o.calc_force_energy()
        o.nve(steps)
        return float(ax @ o.get("pos") + av @ o.get("vel"))

    d = rng.standard_normal(2 * nv)
    fd = (L(x0 + 1e-7 * d) - L(x0 - 1e-7 * d)) / 2e-7
    assert abs(gxa @ d - fd) <= 1e-5 * abs(fd), (gxa @ d, fd)


FIRE_P7 = (0.1, 0.5, 1.1, 0.99, 5e-3, 5e-2, 0.1)  # a_start, f_dec, f_inc, f_a, dt, dt_max, a


@pytest.mark.parametrize("interaction", ["vertexVertex", "vertexSmooth"])
def test_fire_backward_matches_oracle_fd(oracle_mod, interaction):
    """Gradients through 40 FIRE iterations (vertexFIRELoop, FIRE.cu:338-411) against the oracle's
    central differences.  The start velocities point against the forces, so the first iteration
    restarts; the parameters grow dt and shrink alpha after 5 uphill-free steps, and later
    restarts are counted on the oracle.  Harmonic contacts on a soft packing: their force
    derivative jumps only by ec / radSum^2 where a contact opens."""
    print(f"case fire backward {interaction}", flush=True)
    pot = "harmonic"
    st = soft_state(seed=14)
    nr, N = len(st["nvp"]), 40
    ka, kl, kb = ring_costs(nr, 4)
    size = 2 * float(np.mean(st["rad"]))
    g = gpu_from_state(st, potential=pot, interaction=interaction, costs=COSTS)
    g.set_ring_costs(ka, kl, kb)
    cut = g.set_displacement_cutoff(0.5, size)
    g.calc_neighbor_list(cut)
    g.calc_force_energy()
    rng = np.random.default_rng(8)
    f0 = g.get("force")
    v0 = -0.2 * f0 + 1e-3 * rng.standard_normal(f0.size)
    assert v0 @ f0 < 0
    g.set("vel", v0)
    g.init_fire(FIRE_P7, 5, 0)
    g.fire_forward(N)
    xg, vg = g.get("pos"), g.get("vel")
    o = oracle_from_state(oracle_mod, st, potential=pot, interaction=interaction, costs=COSTS)
    o.set_ring_costs(ka, kl, kb)
    o.set_displacement_cutoff(0.5, size)

    def run(x0=st["pos"], v=v0, ka_=ka, kl_=kl, kb_=kb, l0=st["l0"], count=False):
        o.set_ring_costs(ka_, kl_, kb_)
        o.set("l0", l0)
        o.set("pos", x0)
        o.set("vel", v)
        o.calc_neighbor_list(cut)
        o.calc_force_energy()
        o.init_fire(FIRE_P7, 5, 0)
        restarts = 0
        for _ in range(N if count else 1):
            o.fire(1 if count else N)
            restarts += count and o.fire_state()["numStep"] == 0
        return o.get("pos"), o.get("vel"), restarts

    xo, vo, restarts = run(count=True)
    print("restarts", restarts, flush=True)
    assert restarts >= 2, restarts  # the first iteration and at least one more
    np.testing.assert_allclose(xg, xo, rtol=0, atol=1e-11)
    np.testing.assert_allclose(vg, vo, rtol=0, atol=1e-10)
    nv = g.nv
    cx, cv = rng.standard_normal(2 * nv), rng.standard_normal(2 * nv)

    def soa(a):
        return dev(np.concatenate([a[0::2], a[1::2]]))

    ax, av = soa(cx), soa(cv)
    gx, gv = torch.empty_like(ax), torch.empty_like(ax)
    gring = torch.empty(3 * nr, dtype=torch.float64, device="cuda")
    gl0 = torch.empty(nv, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    g.fire_backward(ax.data_ptr(), av.data_ptr(), gx.data_ptr(), gv.data_ptr(), gring.data_ptr(), gl0.data_ptr())
    gx, gv = gx.cpu().numpy(), gv.cpu().numpy()
    gx = np.stack([gx[:nv], gx[nv:]], 1).ravel()
    gv = np.stack([gv[:nv], gv[nv:]], 1).ravel()
    gring, gl0 = gring.cpu().numpy(), gl0.cpu().numpy()
    np.testing.assert_array_equal(g.get("pos"), xg)  # the set stays at the end of the forward pass

    def L(**kw):
        x, v, _ = run(**kw)
        return cx @ x + cv @ v

    h = 1e-7
    d = rng.standard_normal(2 * nv)
    fd = (L(x0=st["pos"] + h * d) - L(x0=st["pos"] - h * d)) / (2 * h)
    assert abs(gx @ d - fd) <= 1e-5 * abs(fd), (gx @ d, fd)
    fd = (L(v=v0 + h * d) - L(v=v0 - h * d)) / (2 * h)
    assert abs(gv @ d - fd) <= 1e-5 * abs(fd), (gv @ d, fd)
    # the parameter derivatives are small against L (|L| ~ 1e2): a step of 1e-7 would leave
    # differences at the rounding level of L, so the parameters take 1e-5
    hp = 1e-5
    p = 2
    for q, base in enumerate((ka, kl, kb)):
        up, dn = base.copy(), base.copy()
        up[p] += hp
        dn[p] -= hp
        key = ("ka_", "kl_", "kb_")[q]
        fd = (L(**{key: up}) - L(**{key: dn})) / (2 * hp)
        assert abs(gring[q * nr + p] - fd) <= 1e-4 * max(abs(fd), 1e-3), (q, gring[q * nr + p], fd)
    dl = rng.standard_normal(nv)
    fd = (L(l0=st["l0"] + hp * dl) - L(l0=st["l0"] - hp * dl)) / (2 * hp)
    assert abs(gl0 @ dl - fd) <= 1e-4 * abs(fd), (gl0 @ dl, fd)

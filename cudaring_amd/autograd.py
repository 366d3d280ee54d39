"""torch.autograd over the adjoint entry points of libdpmring.so.

`DPMAdjoint` wraps one DPMSet (vertexVertex or vertexSmooth interaction, harmonic monomers,
passive or self-propelled rings) and exposes two differentiable functions of torch float64
tensors:

  trajectory(x0, v0, ka, kl, kb, l0, steps, every) -> (xN, vN)
      `steps` NVE steps (NVELoop, DPM2D.h:686) from (x0, v0).  The backward pass is the library's
      reverse pass: one Hessian-vector product per step, with the segments between checkpoints
      (every `every` steps) recomputed on the device.
  potential_energy(x, ka, kl, kb, l0) -> U
      getPotentialEnergy (DPM2D.h:417); dU/dx = -F, dU/dka, dkl, dkb and dU/dl0 from the library.

Positions and velocities are interleaved (x0, y0, x1, y1, ...) as in the reference; ka, kl, kb
are per ring, l0 per vertex.  The shape costs of the reference are global scalars
(cudaKernel.cuh:39-42); per-ring values are an extension that reproduces the global ones when
all rings share them.  calA0 enters through l0 = 2 sqrt(calA0 n tan(pi/n) a0) / n (DPM2D.cu:1575-
1581), which `l0_from_calA0` evaluates in torch so that autograd carries dL/dl0 on to calA0.
"""
import numpy as np
import torch

from .dpm import DPMSet


def _soa(t, nv):
    return t.reshape(nv, 2).t().contiguous().reshape(-1)


def _aos(t, nv):
    return t.reshape(2, nv).t().contiguous().reshape(-1)


def l0_from_calA0(calA0, nvp, a0):
    """Per-vertex rest length from per-ring calA0 and a0 (setPolySizeDistribution's rule)."""
    n = torch.as_tensor(np.asarray(nvp, dtype=np.float64), device=calA0.device)
    a0 = torch.as_tensor(np.asarray(a0, dtype=np.float64), device=calA0.device)
    l0r = 2.0 * torch.sqrt(calA0 * n * torch.tan(np.pi / n) * a0) / n
    return torch.repeat_interleave(l0r, torch.as_tensor(np.asarray(nvp), device=calA0.device))


def shape_parameter(x, nvp, box):
    """Mean P^2 / (4 pi A) over rings, differentiable in torch (rings walked by minimum-image
    hops, as calcParticleShape, cudaKernel.cuh:1571-1606)."""
    nvp = np.asarray(nvp)
    nv = int(nvp.sum())
    pos = x.reshape(nv, 2)
    first = np.concatenate([[0], np.cumsum(nvp)[:-1]])
    nxt = np.arange(nv) + 1
    nxt[first + nvp - 1] = first
    L = torch.as_tensor(np.asarray(box, dtype=np.float64), device=x.device)
    d = pos[torch.as_tensor(nxt, device=x.device)] - pos
    d = d - L * torch.round(d / L)
    seg = torch.sqrt((d * d).sum(1))
    ring = torch.as_tensor(np.repeat(np.arange(len(nvp)), nvp), device=x.device)
    perim = torch.zeros(len(nvp), dtype=x.dtype, device=x.device).index_add_(0, ring, seg)
    # walk positions relative to each ring's first vertex: exclusive prefix sums of the hops
    cs = torch.cumsum(d, 0)
    start = cs[torch.as_tensor(first, device=x.device)] - d[torch.as_tensor(first, device=x.device)]
    u = cs - d - start[ring]
    shoe = u[:, 0] * d[:, 1] - d[:, 0] * u[:, 1]
    area = 0.5 * torch.abs(torch.zeros(len(nvp), dtype=x.dtype, device=x.device).index_add_(0, ring, shoe))
    return (perim * perim / (4 * np.pi * area)).mean()


class DPMAdjoint:
    """Differentiable NVE dynamics and potential energy of one system held in `sset`."""

    def __init__(self, sset: DPMSet, cut_distance):
        if sset.nsys != 1:
            raise ValueError("DPMAdjoint holds one system")
        self.s = sset
        self.cut = cut_distance
        self._params = None

    def _set_params(self, ka, kl, kb, l0):
        vals = [t.detach().cpu().numpy().astype(np.float64) for t in (ka, kl, kb, l0)]
        if self._params is not None and all(np.array_equal(a, b) for a, b in zip(vals, self._params)):
            return
        self.s.set_ring_costs(vals[0], vals[1], vals[2])
        self.s.set("l0", vals[3])
        self._params = vals

    def _prepare(self, x, ka, kl, kb, l0):
        self._set_params(ka, kl, kb, l0)
        self.s.set("pos", x.detach().cpu().numpy())
        self.s.calc_neighbor_list(self.cut)
        self.s.calc_force_energy()

    def trajectory(self, x0, v0, ka, kl, kb, l0, steps, every=100):
        return _Trajectory.apply(x0, v0, ka, kl, kb, l0, self, int(steps), int(every))

    def potential_energy(self, x, ka, kl, kb, l0):
        return _Energy.apply(x, ka, kl, kb, l0, self)


class _Trajectory(torch.autograd.Function):
    """The set holds one checkpoint tape, the last nve_forward's: a node whose tape a later
    trajectory (or a direct nve_forward) has replaced records its own again before the reverse
    pass, from the inputs it keeps, so several trajectories can share one loss."""

    @staticmethod
    def _record(adj, x0, v0, ka, kl, kb, l0, steps, every):
        s = adj.s
        adj._prepare(x0, ka, kl, kb, l0)
        s.set("vel", v0.detach().cpu().numpy())
        s.nve_forward(steps, every)
        return s.tape_serial

    @staticmethod
    def forward(ctx, x0, v0, ka, kl, kb, l0, adj, steps, every):
        s = adj.s
        ctx.tape = _Trajectory._record(adj, x0, v0, ka, kl, kb, l0, steps, every)
        ctx.inputs = tuple(t.detach().clone() for t in (x0, v0, ka, kl, kb, l0))
        ctx.steps, ctx.every = steps, every
        ctx.adj = adj
        ctx.device = x0.device
        xN = torch.from_numpy(s.get("pos")).to(x0.device)
        vN = torch.from_numpy(s.get("vel")).to(x0.device)
        return xN, vN

    @staticmethod
    def backward(ctx, gx, gv):
        s = ctx.adj.s
        if s.tape_serial != ctx.tape:  # the set's tape is another trajectory's: record this one again
            if getattr(s, "particle", "passive") != "passive":
                # self-propelled rings: the angles and random draws the recorded trajectory started
                # from have moved on, so it cannot be recorded again
                raise RuntimeError("DPMAdjoint: with self-propelled rings only the last trajectory of a set can "
                                   "be differentiated (its tape was replaced by a later forward)")
            ctx.tape = _Trajectory._record(ctx.adj, *ctx.inputs, ctx.steps, ctx.every)
        nv, np_ = s.nv, s.np
        dev = torch.device("cuda")
        z = torch.zeros(2 * nv, dtype=torch.float64, device=dev)
        ax = _soa(gx.to(dev, torch.float64), nv) if gx is not None else z
        av = _soa(gv.to(dev, torch.float64), nv) if gv is not None else z.clone()
        gx0 = torch.empty(2 * nv, dtype=torch.float64, device=dev)
        gv0 = torch.empty_like(gx0)
        gring = torch.empty(3 * np_, dtype=torch.float64, device=dev)
        gl0 = torch.empty(nv, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()  # the library works on its own stream
        s.nve_backward(ax.data_ptr(), av.data_ptr(), gx0.data_ptr(), gv0.data_ptr(), gring.data_ptr(),
                       gl0.data_ptr())
        out = ctx.device
        return (_aos(gx0, nv).to(out), _aos(gv0, nv).to(out), gring[:np_].to(out), gring[np_:2 * np_].to(out),
                gring[2 * np_:].to(out), gl0.to(out), None, None, None)


class _Energy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, ka, kl, kb, l0, adj):
        s = adj.s
        adj._prepare(x, ka, kl, kb, l0)
        u = float(s.potential_energy()[0])
        dev = torch.device("cuda")
        gring = torch.empty(3 * s.np, dtype=torch.float64, device=dev)
        gl0 = torch.empty(s.nv, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        s.energy_param_grad(gring.data_ptr(), gl0.data_ptr())
        ctx.save_for_backward(torch.from_numpy(-s.get("force")).to(x.device), gring.to(x.device),
                              gl0.to(x.device))
        ctx.np = s.np
        return torch.tensor(u, dtype=torch.float64, device=x.device)

    @staticmethod
    def backward(ctx, g):
        gradU, gring, gl0 = ctx.saved_tensors
        n = ctx.np
        return g * gradU, g * gring[:n], g * gring[n:2 * n], g * gring[2 * n:], g * gl0, None

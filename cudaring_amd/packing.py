"""Seeded synthetic ring packings.

Sizing follows the reference's setPolySizeDistribution (DPM2D.cu:1545-1588 of
farceri/cudaRing): a ring of n vertices has rest area a0 = (n / n_min)^2,
calA0' = calA0 * n * tan(pi/n) / pi, rest length l0 = 2 sqrt(pi calA0' a0) / n,
rest angle theta0 = 2 pi / n and vertex radius l0 / 2.  Rings are regular
counter-clockwise polygons of area a0 (circumradius sqrt(2 a0 / (n sin(2 pi/n))),
as in initVerticesOnParticles, DPM2D.cu:1678-1691).  The box side comes from the
preferred packing fraction (getPreferredPhi, DPM2D.cu:1222-1238).
"""
import numpy as np


def ring_sizes(nrings, n_small=16, n_large=22, frac_large=0.5, rng=None):
    """Bidisperse vertex counts, 50/50 by default, in a seeded random order."""
    rng = rng or np.random.default_rng(0)
    nl = int(round(frac_large * nrings))
    nvp = np.array([n_large] * nl + [n_small] * (nrings - nl), dtype=np.int64)
    rng.shuffle(nvp)
    return nvp


def ring_shape_params(nvp, calA0=1.10, n_min=None):
    """a0[R], l0[V], theta0[V], rad[V] from the reference's sizing rule."""
    nvp = np.asarray(nvp, dtype=np.int64)
    n_min = n_min or int(nvp.min())
    calA0 = np.broadcast_to(np.asarray(calA0, dtype=np.float64), nvp.shape)
    n = nvp.astype(np.float64)
    a0 = (n / n_min) * (n / n_min)
    calA0tmp = calA0 * n * np.tan(np.pi / n) / np.pi
    l0r = 2.0 * np.sqrt(np.pi * calA0tmp * a0) / n
    l0 = np.repeat(l0r, nvp)
    theta0 = np.repeat(2.0 * np.pi / n, nvp)
    rad = 0.5 * l0
    return a0, l0, theta0, rad


def preferred_phi_area(nvp, a0, rad):
    """Numerator of getPreferredPhi: sum a0 + pi sum r_first^2 (n/2 - 1)."""
    first = np.concatenate([[0], np.cumsum(nvp)[:-1]])
    r = rad[first]
    return a0.sum() + np.pi * np.sum(r * r * (0.5 * nvp - 1))


def place_rings(nvp, a0, centers, rng=None, jitter=0.0, l0=None):
    """Regular CCW polygons of area a0 centred at `centers` (+ optional jitter * l0 * U(0,1))."""
    pos = []
    for p, n in enumerate(nvp):
        R = np.sqrt(2.0 * a0[p] / (n * np.sin(2.0 * np.pi / n)))
        k = np.arange(n)
        ang = 2.0 * np.pi * k / n
        xy = np.stack([R * np.cos(ang) + centers[p, 0], R * np.sin(ang) + centers[p, 1]], axis=1)
        if jitter > 0 and rng is not None:
            xy = xy + jitter * l0[0] * rng.random((n, 2))
        pos.append(xy)
    return np.concatenate(pos, axis=0)


def random_system(nrings, phi, seed, n_small=16, n_large=22, calA0=1.10, scale_to_unit=False):
    """A seeded bidisperse system with random centres at preferred packing fraction phi.

    Returns a dict with nvp, box, pos (AoS, 2V), rad, l0, theta0, a0.  Rings may overlap.
    """
    rng = np.random.default_rng(seed)
    nvp = ring_sizes(nrings, n_small, n_large, rng=rng)
    a0, l0, theta0, rad = ring_shape_params(nvp, calA0)
    L = np.sqrt(preferred_phi_area(nvp, a0, rad) / phi)
    centers = rng.random((nrings, 2)) * L
    pos = place_rings(nvp, a0, centers, rng=rng, jitter=1e-3, l0=l0)
    return {"nvp": nvp, "box": np.array([L, L]), "pos": pos.ravel(), "rad": rad, "l0": l0,
            "theta0": theta0, "a0": a0}


def lattice_system(nrings, phi, seed, n_small=16, n_large=22, calA0=1.10, perturb=0.05):
    """Rings on a jittered square lattice (no ring-ring crossings at moderate phi)."""
    rng = np.random.default_rng(seed)
    nvp = ring_sizes(nrings, n_small, n_large, rng=rng)
    a0, l0, theta0, rad = ring_shape_params(nvp, calA0)
    L = np.sqrt(preferred_phi_area(nvp, a0, rad) / phi)
    m = int(np.ceil(np.sqrt(nrings)))
    a = L / m
    idx = rng.permutation(m * m)[:nrings]
    centers = np.stack([(idx % m + 0.5) * a, (idx // m + 0.5) * a], axis=1)
    centers += perturb * a * (rng.random((nrings, 2)) - 0.5)
    pos = place_rings(nvp, a0, centers, rng=rng, jitter=1e-3, l0=l0)
    return {"nvp": nvp, "box": np.array([L, L]), "pos": pos.ravel(), "rad": rad, "l0": l0,
            "theta0": theta0, "a0": a0}


def size_bidisperse_system(nrings, phi, seed, n=32, ratio=1.4, calA0=1.10, perturb=0.1):
    """Rings of equal vertex count n in two sizes (area ratio ratio^2, 50/50), so l0 and the vertex
    radius differ between the classes; jittered square lattice at preferred phi."""
    rng = np.random.default_rng(seed)
    nvp = np.full(nrings, n, dtype=np.int64)
    big = rng.permutation(nrings) < nrings // 2
    a0 = np.where(big, ratio * ratio, 1.0)
    c = calA0 * n * np.tan(np.pi / n) / np.pi
    l0r = 2.0 * np.sqrt(np.pi * c * a0) / n
    l0 = np.repeat(l0r, n)
    theta0 = np.full(nrings * n, 2.0 * np.pi / n)
    rad = 0.5 * l0
    L = np.sqrt(preferred_phi_area(nvp, a0, rad) / phi)
    m = int(np.ceil(np.sqrt(nrings)))
    a = L / m
    idx = rng.permutation(m * m)[:nrings]
    centers = np.stack([(idx % m + 0.5) * a, (idx // m + 0.5) * a], axis=1)
    centers += perturb * a * (rng.random((nrings, 2)) - 0.5)
    pos = place_rings(nvp, a0, centers, rng=rng, jitter=1e-3, l0=l0)
    return {"nvp": nvp, "box": np.array([L, L]), "pos": pos.ravel(), "rad": rad, "l0": l0,
            "theta0": theta0, "a0": a0}


def _morton2(ix, iy):
    """Interleave the bits of two non-negative int arrays (16 bits each)."""
    def spread(v):
        v = v.astype(np.uint64) & 0xFFFF
        v = (v | (v << 8)) & 0x00FF00FF
        v = (v | (v << 4)) & 0x0F0F0F0F
        v = (v | (v << 2)) & 0x33333333
        v = (v | (v << 1)) & 0x55555555
        return v
    return spread(ix) | (spread(iy) << np.uint64(1))


def spatial_order(st, ncell=None):
    """The same system with its rings renumbered along a Morton curve of their centres, so that
    rings close in space are close in memory.  Returns (state, ring permutation)."""
    nvp = np.asarray(st["nvp"])
    L = np.asarray(st["box"], dtype=np.float64)
    first = np.concatenate([[0], np.cumsum(nvp)])
    pos = np.asarray(st["pos"]).reshape(-1, 2)
    centres = np.empty((len(nvp), 2))
    for p in range(len(nvp)):
        seg = pos[first[p]:first[p + 1]]
        d = seg - seg[0]
        d -= L * np.round(d / L)
        centres[p] = seg[0] + d.mean(0)
    centres -= L * np.floor(centres / L)
    m = ncell or int(2 ** np.ceil(np.log2(max(2, np.sqrt(len(nvp))))))
    cell = np.minimum((centres / L * m).astype(np.int64), m - 1)
    order = np.argsort(_morton2(cell[:, 0], cell[:, 1]), kind="stable")
    vidx = np.concatenate([np.arange(first[p], first[p + 1]) for p in order])
    out = dict(st)
    out["nvp"] = nvp[order]
    out["pos"] = pos[vidx].ravel()
    for k in ("rad", "l0", "theta0"):
        out[k] = np.asarray(st[k])[vidx]
    out["a0"] = np.asarray(st["a0"])[order]
    for k in ("vel",):
        if k in st and st[k] is not None:
            out[k] = np.asarray(st[k]).reshape(-1, 2)[vidx].ravel()
    if "pangle" in st and st["pangle"] is not None:
        out["pangle"] = np.asarray(st["pangle"])[order]
    return out, order


def concat_systems(systems):
    """Stack systems into one set: nvp, rings_per_system and concatenated fields."""
    out = {"nvp": np.concatenate([s["nvp"] for s in systems]),
           "rings_per_system": np.array([len(s["nvp"]) for s in systems], dtype=np.int64)}
    for k in ("pos", "rad", "l0", "theta0", "a0", "box"):
        out[k] = np.concatenate([np.asarray(s[k]).ravel() for s in systems])
    for k in ("vel", "pangle"):
        if all(k in s for s in systems):
            out[k] = np.concatenate([np.asarray(s[k]).ravel() for s in systems])
    return out

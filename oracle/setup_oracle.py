"""Oracle of the packing set-up path (test infrastructure only: imported by tests, never by the
product).

Pure-Python restatement of the reference's host-side initialisation, draw for draw:
  - drand48 / srand48: the POSIX 48-bit linear congruential generator (a = 0x5DEECE66D, c = 0xB,
    m = 2^48; srand48 sets the state to (seed << 16) | 0x330E).  Pinned against glibc's own
    drand48 through ctypes (tests/test_host_setup.py).
  - setPolySizeDistribution (DPM2D.cu:1545-1588): per ring one Box-Muller variate from two draws,
    n = floor(poly * nvp * g + nvp), clamped below at nvp; a0 = (n / nvp)^2,
    l0 = 2 sqrt(pi calA0 n tan(pi/n) / pi * a0) / n, theta0 = 2 pi / n, rad = l0 / 2.
  - setRandomParticles (DPM2D.cu:1601-1630): unit box, scale = sqrt(preferredPhi / phi0),
    a0 /= scale^2, centre x then y from drand48, disk radius
    extraRad sqrt(2 a0 / (n sin(2 pi / n))), l0 and rad /= scale.
  - initVerticesOnParticles (DPM2D.cu:1678-1691): vertex v of ring p at angle 2 pi v / n with v the
    GLOBAL vertex index, circumradius sqrt(2 a0 / (n sin(2 pi / n))), plus 1e-3 l0 drand48() on x
    and then on y.
  - soft-disk harmonic force (kernelCalcParticleInteraction, cudaKernel.cuh:1185-1217).
getPreferredPhi follows DPM2D.cu:1222-1238 with kernelCalcVertexArea (cudaKernel.cuh:1636-1642).
"""
import math

A48, C48, M48 = 0x5DEECE66D, 0xB, 1 << 48


class Drand48:
    def __init__(self, seed):
        self.x = ((int(seed) & 0xFFFFFFFF) << 16) | 0x330E

    def __call__(self):
        self.x = (A48 * self.x + C48) % M48
        return self.x / M48


def poly_sizes(num_particles, nvp, poly, rng):
    n = []
    for _ in range(num_particles):
        r1 = rng()
        r2 = rng()
        g = math.sqrt(-2.0 * math.log(r1)) * math.cos(2.0 * math.pi * r2)
        k = math.floor(poly * nvp * g + nvp)
        n.append(int(max(k, nvp)))
    return n


def ring_sizing(n, n_min, calA0):
    a0, l0, th0, rad = [], [], [], []
    for k in n:
        a = (k / n_min) * (k / n_min)
        c = calA0 * k * math.tan(math.pi / k) / math.pi
        length = 2.0 * math.sqrt(math.pi * c * a) / k
        a0.append(a)
        l0 += [length] * k
        th0 += [2.0 * math.pi / k] * k
        rad += [0.5 * length] * k
    return a0, l0, th0, rad


def preferred_phi(n, a0, rad, box=(1.0, 1.0)):
    first, v = 0, 0.0
    for k in n:
        r = rad[first]
        v += r * r * (0.5 * k - 1)
        first += k
    return (sum(a0) + math.pi * v) / (box[0] * box[1])


def random_particles(n, a0, l0, rad, phi0, extra_rad, rng):
    scale = math.sqrt(preferred_phi(n, a0, rad) / phi0)
    a0 = list(a0)
    pos, prad = [], []
    for p, k in enumerate(n):
        a0[p] /= scale * scale
        pos += [rng(), rng()]
        prad.append(extra_rad * math.sqrt(2.0 * a0[p] / (k * math.sin(2.0 * math.pi / k))))
    return a0, [x / scale for x in l0], [x / scale for x in rad], pos, prad


def scaled_random_particles(n, a0, rad, phi0, extra_rad, lx, ly, rng):
    """setScaledRandomParticles, DPM2D.cu:1631-1653: box grown to phi0, disks in 0.95 of it"""
    scale = math.sqrt(preferred_phi(n, a0, rad, (lx, ly)) / phi0)
    box = (lx * scale, ly * scale)
    pos, prad = [], []
    for p, k in enumerate(n):
        pos += [0.95 * box[0] * rng(), 0.95 * box[1] * rng()]
        prad.append(extra_rad * math.sqrt(2.0 * a0[p] / (k * math.sin(2.0 * math.pi / k))))
    return box, pos, prad


def round_scaled_random_particles(n, a0, rad, phi0, extra_rad, radius, rng):
    """setRoundScaledRandomParticles, DPM2D.cu:1655-1676: round box grown to phi0 (preferred phi
    over pi R^2), disks at radius 0.95 R sqrt(U), angle 2 pi U - pi"""
    area = preferred_phi(n, a0, rad) / (math.pi * radius * radius)
    R = radius * math.sqrt(area / phi0)
    pos, prad = [], []
    for p, k in enumerate(n):
        r = 0.95 * R * math.sqrt(rng())
        t = 2 * math.pi * rng() - math.pi
        pos += [r * math.cos(t), r * math.sin(t)]
        prad.append(extra_rad * math.sqrt(2.0 * a0[p] / (k * math.sin(2.0 * math.pi / k))))
    return R, pos, prad


def vertices_on_particles(n, a0, l0, ppos, rng):
    pos, v = [], 0
    for p, k in enumerate(n):
        R = math.sqrt(2.0 * a0[p] / (k * math.sin(2.0 * math.pi / k)))
        for _ in range(k):
            x = R * math.cos(2.0 * math.pi * v / k) + ppos[2 * p] + 1e-3 * l0[v] * rng()
            y = R * math.sin(2.0 * math.pi * v / k) + ppos[2 * p + 1] + 1e-3 * l0[v] * rng()
            pos += [x, y]
            v += 1
    return pos


def soft_disk_forces(pos, rad, box, ec=1.0):
    """All-pairs harmonic disk forces and per-disk energies (each directed pair stores half)."""
    num = len(rad)
    f = [0.0] * (2 * num)
    e = [0.0] * num
    for i in range(num):
        for j in range(num):
            if i == j:
                continue
            dx = pos[2 * i] - pos[2 * j]
            dy = pos[2 * i + 1] - pos[2 * j + 1]
            dx -= box[0] * round_half_away(dx / box[0])
            dy -= box[1] * round_half_away(dy / box[1])
            d = math.hypot(dx, dy)
            s = rad[i] + rad[j]
            ov = 1.0 - d / s
            if ov > 0:
                g = ec * ov / s
                f[2 * i] += g * dx / d
                f[2 * i + 1] += g * dy / d
                e[i] += 0.25 * ec * ov * ov
    return f, e


def round_half_away(x):
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)

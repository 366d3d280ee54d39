"""Packing set-up through the C++ class layer (DPM2D / FIRE / ioDPMFile), on the CPU.

The compressFIRE driver is run with a target packing fraction of 0, which stops after the
host-side set-up (sizes, random soft disks, soft-disk FIRE, vertices on the disks) and writes the
set-up files; no device is touched.  The files are compared with the set-up oracle draw for draw.
"""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from oracle import setup_oracle as so

BIN = os.path.join(ROOT, "cudaring_amd", "bin", "compressFIRE")


@pytest.fixture(scope="module")
def driver():
    if not os.path.exists(BIN):
        from cudaring_amd import build as b
        if not os.path.exists(b.LIB):
            b.build()
        b.build_host()
    return BIN


def run_setup(driver, tmp, np_, nvp, calA0, poly, seed, extra=()):
    out = os.path.join(tmp, "run")
    r = subprocess.run([driver, out, str(np_), str(nvp), str(calA0), str(poly), "0", str(seed), *extra],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return out


def load1(path):
    return [float(x) for x in open(path).read().split()]


def test_drand48_matches_libc():
    libc = ctypes.CDLL("libc.so.6")
    libc.drand48.restype = ctypes.c_double
    for seed in (0, 1, 7, 123456789):
        libc.srand48(ctypes.c_long(seed))
        r = so.Drand48(seed)
        assert all(libc.drand48() == r() for _ in range(500))


def test_round_half_away():
    assert so.round_half_away(0.5) == 1 and so.round_half_away(-0.5) == -1
    assert so.round_half_away(1.49) == 1 and so.round_half_away(-2.5) == -3


@pytest.mark.parametrize("np_,nvp,poly,seed", [(48, 16, 0.2, 7), (30, 12, 0.1, 99)])
def test_setup_matches_oracle(driver, tmp_path, np_, nvp, poly, seed):
    calA0, phi0 = 1.1, 0.2
    out = run_setup(driver, str(tmp_path), np_, nvp, calA0, poly, seed)
    d = os.path.join(out, "setup")
    rng = so.Drand48(seed)
    n = so.poly_sizes(np_, nvp, poly, rng)
    assert [int(x) for x in load1(os.path.join(d, "numVertexInParticleList.dat"))] == n
    a0, l0, th0, rad = so.ring_sizing(n, nvp, calA0)
    a0, l0, rad, ppos0, prad = so.random_particles(n, a0, l0, rad, phi0, 1.1, rng)
    # files carry 14 significant digits
    np.testing.assert_allclose(load1(os.path.join(d, "restAreas.dat")), a0, rtol=1e-13)
    np.testing.assert_allclose(load1(os.path.join(d, "restLengths.dat")), l0, rtol=1e-13)
    np.testing.assert_allclose(load1(os.path.join(d, "restAngles.dat")), th0, rtol=1e-13)
    np.testing.assert_allclose(load1(os.path.join(d, "radii.dat")), rad, rtol=1e-13)
    np.testing.assert_allclose(load1(os.path.join(d, "particleRad.dat")), prad, rtol=1e-13)
    assert load1(os.path.join(d, "boxSize.dat")) == [1.0, 1.0]
    # preferred packing fraction reached by the rest-size scaling
    assert abs(so.preferred_phi(n, a0, rad) - phi0) < 1e-12
    # vertices continue the same drand48 stream from the relaxed disk centres
    ppos = load1(os.path.join(out, "sp", "particlePos.dat"))
    pos = so.vertices_on_particles(n, a0, l0, ppos, rng)
    np.testing.assert_allclose(load1(os.path.join(d, "positions.dat")), pos, rtol=0, atol=1e-13)


def test_soft_disk_fire_relaxes(driver, tmp_path):
    np_, nvp, poly, seed = 48, 16, 0.2, 11
    out = run_setup(driver, str(tmp_path), np_, nvp, 1.1, poly, seed)
    rng = so.Drand48(seed)
    n = so.poly_sizes(np_, nvp, poly, rng)
    a0, l0, _, rad = so.ring_sizing(n, nvp, 1.1)
    _, _, _, ppos0, prad = so.random_particles(n, a0, l0, rad, 0.2, 1.1, rng)
    f0, e0 = so.soft_disk_forces(ppos0, prad, (1.0, 1.0))
    assert sum(e0) > 0  # the random start overlaps, so the minimiser had work to do
    ppos = load1(os.path.join(out, "sp", "particlePos.dat"))
    f, e = so.soft_disk_forces(ppos, prad, (1.0, 1.0))
    tol = 1e-10 / (sum(prad) / len(prad))
    assert max(abs(x) for x in f) <= tol
    assert sum(e) < 1e-20
    moved = [math.hypot(ppos[2 * i] - ppos0[2 * i], ppos[2 * i + 1] - ppos0[2 * i + 1]) for i in range(np_)]
    assert max(moved) > 1e-4


def test_file_format(driver, tmp_path):
    out = run_setup(driver, str(tmp_path), 20, 12, 1.1, 0.1, 3)
    lines = open(os.path.join(out, "setup", "positions.dat")).read().splitlines()
    for ln in lines:
        assert ln.endswith("\t") and ln.count("\t") == 2
        for tok in ln.split("\t")[:2]:
            assert tok == format(float(tok), ".14g")
    for tok in open(os.path.join(out, "setup", "radii.dat")).read().split():
        assert tok == format(float(tok), ".14g")
    params = dict(ln.split("\t") for ln in open(os.path.join(out, "sp", "params.dat")).read().splitlines())
    assert params["numParticles"] == "20"


def test_bidisperse_sizes(driver, tmp_path):
    out = run_setup(driver, str(tmp_path), 40, 16, 1.1, 0.0, 5, extra=("bi=16,22",))
    n = [int(x) for x in load1(os.path.join(out, "setup", "numVertexInParticleList.dat"))]
    assert sorted(set(n)) == [16, 22] and n.count(22) == 20
    a0 = load1(os.path.join(out, "setup", "restAreas.dat"))
    ratio = {k: a for k, a in zip(n, a0)}
    assert abs(ratio[22] / ratio[16] - (22 / 16) ** 2) < 1e-12


NVE_BIN = os.path.join(ROOT, "cudaring_amd", "bin", "compressNVE")


@pytest.mark.parametrize("box_type", ["periodic", "circle"])
def test_compress_nve_setup_matches_oracle(driver, tmp_path, box_type):
    """compressNVE's own packing (compressNVE.cpp:70-112) with maxStep 0: sizes, scaled random
    disks in the box (square, or round with its wall), scaleParticlePacking, soft-disk FIRE,
    vertices on the disks, scalePacking by the mean particle size -- draw for draw against the
    set-up oracle"""
    np_, shape0, lx, ly, seed = 24, 1.1, 2.0, 1.0, 13
    out = str(tmp_path)
    r = subprocess.run([NVE_BIN, out, "0", "1e-3", "1e-4", str(shape0), str(np_), "0", str(lx), str(ly), "wca",
                        box_type, "0.9", "0.002", str(seed)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    fire_line = [ln for ln in r.stdout.splitlines() if ln.startswith("FIRE: iteration")][-1]
    assert float(fire_line.split("maxUnbalancedForce: ")[1].split()[0]) <= 1e-12
    base = os.path.join(out, "circle", f"{np_}-A{shape0}") if box_type == "circle" else \
        os.path.join(out, f"{np_}-box{lx}{ly}-A{shape0}")
    assert os.path.isdir(base), os.listdir(out)
    rng = so.Drand48(seed)
    n = so.poly_sizes(np_, 32, 0.1, rng)
    d = os.path.join(base, "initial")
    assert [int(x) for x in load1(os.path.join(d, "numVertexInParticleList.dat"))] == n
    a0, l0, th0, rad = so.ring_sizing(n, 32, shape0)
    if box_type == "circle":
        R, ppos0, prad = so.round_scaled_random_particles(n, a0, rad, 0.1, 1.5, lx, rng)
    else:
        box, ppos0, prad = so.scaled_random_particles(n, a0, rad, 0.1, 1.5, lx, ly, rng)
    # scaleParticlePacking: every length by the mean disk radius
    s = sum(prad) / len(prad)
    prad = [x / s for x in prad]
    l0 = [x / s for x in l0]
    rad = [x / s for x in rad]
    a0 = [x / (s * s) for x in a0]
    ppos = load1(os.path.join(base, "fire", "particlePos.dat"))
    np.testing.assert_allclose(load1(os.path.join(base, "fire", "particleRad.dat")), prad, rtol=1e-13)
    pos = so.vertices_on_particles(n, a0, l0, ppos, rng)
    # scalePacking by the mean particle size 2 sqrt(<a0> / pi)
    m = 2 * math.sqrt(sum(a0) / len(a0) / math.pi)
    np.testing.assert_allclose(load1(os.path.join(d, "positions.dat")), [x / m for x in pos], rtol=0, atol=1e-12)
    np.testing.assert_allclose(load1(os.path.join(d, "restAreas.dat")), [x / (m * m) for x in a0], rtol=1e-13)
    np.testing.assert_allclose(load1(os.path.join(d, "radii.dat")), [x / m for x in rad], rtol=1e-13)
    if box_type == "circle":
        np.testing.assert_allclose(load1(os.path.join(d, "boxSize.dat")), [R / s / m], rtol=1e-13)
    else:
        np.testing.assert_allclose(load1(os.path.join(d, "boxSize.dat")), [b / s / m for b in box], rtol=1e-13)

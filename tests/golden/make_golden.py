"""Writes tests/golden/golden_small.npz: small seeded states and the oracle's outputs on them.

The reference ships no fixtures (no tests, no data files), so these vectors come from the
oracle (oracle/dpm_oracle.c) and pin it against later change; the GPU tests compare the library
with them.  Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from cudaring_amd import packing  # noqa: E402
from oracle import oracle  # noqa: E402

POTS = ["harmonic", "wca", "lj", "adhesive"]
INTERS = ["vertexVertex", "vertexSmooth"]
COSTS = (1.0, 1.0, 0.01, 1.0)


def make_oracle(st, pot="harmonic", inter="vertexVertex", dt=5e-3, seed=7):
    o = oracle.Oracle(st["nvp"])
    o.load_state(st)
    o.set_types(pot, inter)
    o.set_energy_costs(*COSTS)
    o.set_attraction(0.1, 0.2)
    o.set_lj_cutoff(1.5)
    o.set_dt(dt)
    o.set_rng(seed, 0)
    cut = o.set_displacement_cutoff(0.5, 2 * float(np.mean(st["rad"])))
    o.calc_neighbor_list(cut)
    return o


def main():
    oracle.build()
    out = {}
    st = packing.random_system(8, 0.75, 31)
    for k in ("nvp", "box", "pos", "rad", "l0", "theta0", "a0"):
        out[f"state_{k}"] = np.asarray(st[k])
    for pot in POTS:
        for inter in INTERS:
            o = make_oracle(st, pot, inter)
            o.calc_force_energy()
            key = f"{pot}_{inter}"
            out[f"force_{key}"] = o.get("force")
            out[f"energy_{key}"] = o.get("energy")
            out[f"penergy_{key}"] = o.get("penergy")
            out[f"U_{key}"] = np.array([o.potential_energy()])
    lat = packing.lattice_system(8, 0.8, 5)
    for k in ("nvp", "box", "pos", "rad", "l0", "theta0", "a0"):
        out[f"lattice_{k}"] = np.asarray(lat[k])
    o = make_oracle(lat)
    o.calc_force_energy()
    o.init_nve(1e-3, zero_momentum=True)
    o.nve(50)
    out["nve50_pos"], out["nve50_vel"] = o.get("pos"), o.get("vel")
    o = make_oracle(lat, seed=9)
    o.calc_force_energy()
    o.init_langevin(1e-2, 1.0)
    o.langevin(50)
    out["langevin50_pos"], out["langevin50_vel"] = o.get("pos"), o.get("vel")
    o = make_oracle(st)
    o.calc_force_energy()
    o.init_fire([0.2, 0.5, 1.1, 0.99, 5e-3, 5e-2, 0.2], 20, 0)
    o.fire(100)
    fs = o.fire_state()
    out["fire100_pos"] = o.get("pos")
    out["fire100_state"] = np.array([fs["dt"], fs["alpha"], fs["numStep"]], dtype=np.float64)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden_small.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    main()

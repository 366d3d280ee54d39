"""Shared builders for the oracle and the GPU set from one synthetic state."""
import numpy as np

from cudaring_amd import packing

CUTOFF = 0.5  # runNVT.cpp:37 displacement cutoff


def oracle_from_state(oracle_mod, st, potential="harmonic", interaction="vertexVertex", costs=(1, 1, 0.01, 1),
                      l1=0.1, l2=0.2, lj=1.5, dt=1e-3, seed=7, system_id=0, particle="passive"):
    o = oracle_mod.Oracle(st["nvp"])
    o.load_state(st)
    o.set_types(potential, interaction, "harmonic", particle)
    o.set_energy_costs(*costs)
    o.set_attraction(l1, l2)
    o.set_lj_cutoff(lj)
    o.set_dt(dt)
    o.set_rng(seed, system_id)
    return o


def gpu_from_state(st, potential="harmonic", interaction="vertexVertex", costs=(1, 1, 0.01, 1), l1=0.1, l2=0.2,
                   lj=1.5, dt=1e-3, seed=7, particle="passive", rings_per_system=None):
    from cudaring_amd import DPMSet
    g = DPMSet(st["nvp"], rings_per_system=rings_per_system)
    g.load_state(st)
    g.set_types(potential, interaction, "harmonic", particle)
    g.set_energy_costs(*costs)
    g.set_attraction(l1, l2)
    g.set_lj_cutoff(lj)
    g.set_dt(dt)
    g.set_seed(seed)
    return g


def small_state(nrings=12, phi=0.7, seed=3):
    return packing.random_system(nrings, phi, seed)


def relaxed_state(nrings, phi, seed, final=("wca", "vertexSmooth")):
    """A compressed, FIRE-relaxed packing prepared on the GPU (deterministic), so that a stiff
    potential starts without deep ring overlaps.  GPU tests only."""
    from cudaring_amd import DPMSet, protocols
    s, st = protocols.prepared_state(lambda nvp: DPMSet(nvp), nrings, phi, seed, dphi=0.05, iters=400,
                                     final=final, final_iters=1500)
    s.close()
    return st


def thermalized_state(st, potential, interaction, T=1e-3, steps=300, seed=11):
    """The positions of `st` after `steps` Langevin steps at temperature T (gamma 1) on the GPU.
    A packing relaxed under a purely repulsive potential leaves many pairs parked at the cutoff,
    where the force's derivative jumps, so finite differences there straddle a kink; thermal
    motion spreads the pair distances.  GPU tests only."""
    g = gpu_from_state(st, potential=potential, interaction=interaction, dt=1e-3, seed=seed)
    cut = g.set_displacement_cutoff(0.5, 2 * float(np.mean(st["rad"])))
    g.calc_neighbor_list(cut)
    g.calc_force_energy()
    g.init_langevin(T, 1.0)
    g.langevin(steps)
    out = dict(st)
    out["pos"] = g.get("pos")
    g.close()
    return out

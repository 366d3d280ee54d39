"""Autotuner for the launch parameters of a system set.

The step kernel runs one workgroup per ring tile; how many wavefronts share a
tile's neighbour gather (1..8) trades latency hiding against idle lanes and LDS
traffic, and the best value depends on the neighbour count and on how many
tiles the set has.  `tune_waves_per_tile` times the fused Langevin step kernel
for each candidate on the caller's own state (velocities are restored by the
library after timing) and keeps the fastest; `tune_waves_per_tile_loop` times a
stretch of the caller's own loop instead, rebuilds included.  `tune_skin` picks the neighbour
skin (setDisplacementCutoff's `cutoff`) that minimises the measured time of a
short stretch of the real loop, rebuilds included; `tune_fire_neighbors` picks the skin and the
cell size of the batched FIRE minimisation the same way.
"""
import time

import numpy as np


def tune_waves_per_tile(s, candidates=(1, 2, 4, 8), reps=100):
    """Time the step kernel for each waves-per-tile value; keep and return the fastest."""
    times = {}
    for w in candidates:
        s.set_launch_config(w)
        times[w] = min(s.time_step_kernel(reps)[0] for _ in range(2))
    best = min(times, key=times.get)
    s.set_launch_config(best)
    return best, times


def tune_waves_per_tile_loop(s, run, candidates=(2, 4, 8), reps=2):
    """Pick the waves per tile minimising the wall time of `run()` (a stretch of the caller's own
    loop, rebuilds included: a set whose list rebuilds every few steps spends much of its time in
    the row build, which more waves share).  The state is advanced while tuning, so call it on a
    warm-up stretch of the run."""
    times = {}
    for w in candidates:
        s.set_launch_config(w)
        run()  # capture the graph for this configuration
        s.synchronize()
        best = None
        for _ in range(reps):
            t0 = time.perf_counter()
            run()
            s.synchronize()
            el = time.perf_counter() - t0
            best = el if best is None else min(best, el)
        times[w] = best
    best = min(times, key=times.get)
    s.set_launch_config(best)
    return best, times


def tune_skin(s, size, candidates=(0.2, 0.3, 0.5, 0.8), steps=400, kind="langevin"):
    """Pick the displacement cutoff (skin) minimising the wall time of `steps` loop steps.

    The state is advanced while tuning, so call it on a warm-up stretch of the run."""
    times = {}
    run = s.langevin if kind == "langevin" else s.nve
    for c in candidates:
        cut = s.set_displacement_cutoff(c, size)
        s.calc_neighbor_list(cut)
        run(steps)  # capture the graph for this configuration
        s.synchronize()
        t0 = time.perf_counter()
        run(steps)
        s.synchronize()
        times[c] = time.perf_counter() - t0
    best = min(times, key=times.get)
    cut = s.set_displacement_cutoff(best, size)
    s.calc_neighbor_list(cut)
    return best, times


def tune_fire(s, params7, cut, candidates=((1, 2), (1, 4), (1, 8), (1, 16), (0, 0)), iters=200):
    """Pick the fire_minimize kernel (persistent workgroup per system with 2/4/8/16 waves, or graphs
    of the split step kernels) by timing `iters` iterations of each from the caller's state,
    which is restored after every trial and at the end."""
    keep = {k: s.get(k) for k in ("pos", "vel")}
    times = {}
    for persistent, waves in candidates:
        for k, v in keep.items():
            s.set(k, v)
        s.calc_neighbor_list(cut)
        s.calc_force_energy()
        s.init_fire(params7, 20, 0)
        s.set_fire_kernel(persistent, waves)
        s.synchronize()
        t0 = time.perf_counter()
        s.fire_minimize(0.0, iters)
        s.synchronize()
        times[(persistent, waves)] = time.perf_counter() - t0
    best = min(times, key=times.get)
    for k, v in keep.items():
        s.set(k, v)
    s.calc_neighbor_list(cut)
    s.calc_force_energy()
    s.set_fire_kernel(*best)
    return best, {f"{'persistent' if p else 'graph'}{w or ''}": round(t, 4) for (p, w), t in times.items()}


def tune_fire_neighbors(s, params7, size, skins=(0.2, 0.3, 0.5, 0.8), scales=(1.0, 1.5, 2.0), iters=200):
    """Pick the neighbour skin (setDisplacementCutoff's `cutoff`, in units of `size`) and the
    cell scale (dpm_set_cell_scale) minimising the time of `iters` fire_minimize iterations from
    the caller's state, which is restored after every trial and at the end.  The forces, and so
    the FIRE trajectory, do not depend on either: pairs the wider list adds lie beyond the
    potential's range."""
    keep = {k: s.get(k) for k in ("pos", "vel")}

    def restore(skin, scale):
        for k, v in keep.items():
            s.set(k, v)
        cut = s.set_displacement_cutoff(skin, size)
        s.set_cell_scale(scale)
        s.calc_neighbor_list(cut)
        s.calc_force_energy()
        s.init_fire(params7, 20, 0)
        return cut

    times = {}
    for skin in skins:
        for scale in scales:
            restore(skin, scale)
            s.synchronize()
            t0 = time.perf_counter()
            s.fire_minimize(0.0, iters)
            s.synchronize()
            times[(skin, scale)] = time.perf_counter() - t0
    best = min(times, key=times.get)
    cut = restore(*best)
    return best, cut, {f"skin{k[0]}_cells{k[1]}": round(t, 4) for k, t in times.items()}


def summary(times):
    return {str(k): round(float(v), 4) for k, v in sorted(times.items())}


def choose(values, times):
    return values[int(np.argmin([times[v] for v in values]))]

"""Benchmark of the DPM2D vertex engine on one MI355X.

Headline workload (N=1): Langevin NVT of 1024 bidisperse rings (16/22 vertices,
19,456 vertices), WCA vertex-segment ("vertexSmooth") interaction, preferred
packing fraction 0.85, T = 1e-2, gamma = 1, FP64, neighbour list rebuilt on skin
crossing.  A step is one velocity-Verlet Langevin step of the whole system; the
metric is vertex-steps/s = vertices * steps / wall time, with the state resident
on the GPU.  The synthetic packing is prepared (compressed with FIRE) before the
timed region.

One JSON line is printed with the roofline of the dominant kernel (the fused
step kernel, timed live with hipEvents on its stream) and a CPU baseline (the
oracle, single thread, on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "vertex-steps/s (NVT N=1024); batched FIRE packings/s vs plain; adjoint/fwd"
HBM_PEAK_GBS = 8000.0
NRINGS, PHI, T, GAMMA, SEED = 1024, 0.85, 1e-2, 1.0, 20261202 + 2
COSTS = (1.0, 1.0, 0.01, 1.0)  # ka, kl, kb, ec
SKIN = 0.5                     # setDisplacementCutoff(0.5, 2 <r_v>), runNVT.cpp:37,136
DT_UNITS = 1e-3                # dt = 1e-3 * sigma / sqrt(ec), sigma = mean particle size
CHUNK = 250                    # steps per graph replay


def alg_bytes_per_step(nv, np_, kbar):
    """Algorithmic HBM bytes of one step: 132 + 4 kbar per vertex and 16 per ring (SoA FP64,
    int32 ids: pos r+w 32, vel r+w 32, F read 16 + write 16, rad/l0/theta0 24, ring id 4,
    energy 8, neighbour ids 4 kbar; ring a0 + offset 16)."""
    return nv * (132.0 + 4.0 * kbar) + np_ * 16.0


def prepare(args):
    from cudaring_amd import DPMSet, autotune, protocols
    t0 = time.perf_counter()
    s, st = protocols.prepared_state(lambda nvp: DPMSet(nvp), NRINGS, PHI, SEED, costs=COSTS,
                                     final=("wca", "vertexSmooth"), iters=args.prep_iters,
                                     final_iters=3 * args.prep_iters)
    sigma = float(2 * np.sqrt(np.mean(s.get("a0")) / np.pi))  # getMeanParticleSize
    dt = DT_UNITS * sigma
    s.set_dt(dt)
    s.set_seed(SEED)
    size = 2 * float(np.mean(s.get("rad")))
    cut = s.set_displacement_cutoff(SKIN, size)
    s.calc_neighbor_list(cut)
    s.calc_force_energy()
    s.init_langevin(T, GAMMA)
    waves, wtimes = autotune.tune_waves_per_tile(s)
    return s, st, dt, time.perf_counter() - t0, {"waves_per_tile": waves, "kernel_us": autotune.summary(wtimes)}


def cpu_baseline(st, dt, steps):
    """The oracle (single thread) on `steps` Langevin steps of the same prepared state."""
    from oracle import oracle
    o = oracle.Oracle(st["nvp"])
    o.load_state(st)
    o.set_types("wca", "vertexSmooth")
    o.set_energy_costs(*COSTS)
    o.set_dt(dt)
    o.set_rng(SEED, 0)
    size = 2 * float(np.mean(st["rad"]))
    cut = o.set_displacement_cutoff(SKIN, size)
    o.calc_neighbor_list(cut)
    o.calc_force_energy()
    o.init_langevin(T, GAMMA)
    o.langevin(2)
    t0 = time.perf_counter()
    o.langevin(steps)
    el = time.perf_counter() - t0
    nv = int(np.sum(st["nvp"]))
    return {"value": nv * steps / el, "unit": "vertex-steps/s", "cores": 1, "kind": "port",
            "sample": f"{steps} Langevin steps of the same {NRINGS}-ring WCA vertexSmooth state "
                      f"(incl. its all-pairs neighbour rebuilds), one thread, {el:.1f} s"}


COPY_ELEMS = 1 << 27  # float64 elements of the calibration copy: 1 GiB read + 1 GiB written


def traffic_run(s, torch):
    """Launch sequence for `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE`: three copies of
    COPY_ELEMS doubles (the calibration kernel: 8 * COPY_ELEMS bytes read and written each), then
    50 back-to-back launches of the step kernel on the prepared state."""
    src = torch.rand(COPY_ELEMS, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    us, kbar = s.time_step_kernel(50)
    s.synchronize()
    print(json.dumps({"traffic_run": True, "copy_bytes_each_way": 8 * COPY_ELEMS, "step_launches": 50,
                      "kernel_us": us, "mean_neighbors": kbar,
                      "alg_bytes_per_launch": alg_bytes_per_step(s.nv, s.np, kbar)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--prep-iters", type=int, default=1000)
    ap.add_argument("--cpu-steps", type=int, default=1200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic", action="store_true",
                    help="counter-collection run: a copy of known size, then step-kernel launches only")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the path is single-GPU (systems shard across GPUs "
              "with no exchange; see DESIGN.md).  Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch  # device plumbing: synchronisation around the timed region
    s, st, dt, prep_s, tuned = prepare(args)
    nv, np_ = s.nv, s.np
    if args.traffic:
        traffic_run(s, torch)
        return

    done = 0
    while done < args.warmup:
        n = min(CHUNK, args.warmup - done)
        s.langevin(n)
        done += n
    s.synchronize()
    torch.cuda.synchronize()
    s.reset_update_count()
    t0 = time.perf_counter()
    done = 0
    while done < args.steps:
        n = min(CHUNK, args.steps - done)
        s.langevin(n)
        done += n
    s.synchronize()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    rebuilds = int(s.update_count()[0])
    temp = float(s.temperature()[0])
    if not np.isfinite(temp):
        raise RuntimeError("non-finite temperature after the timed run")

    value = nv * args.steps / el
    us, kbar = s.time_step_kernel(200)
    alg = alg_bytes_per_step(nv, np_, kbar)
    achieved = alg / (us * 1e-6) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
            "kernel": "k_step<vertexSmooth,WCA,kick|langevin|start>", "kernel_us": round(us, 3),
            "alg_bytes_per_launch": int(alg), "mean_neighbors": round(kbar, 3),
            "step_us": round(1e6 * el / args.steps, 3)}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(st, dt, args.cpu_steps)
    out = {
        "metric": METRIC, "value": value, "unit": "vertex-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "NVT Langevin, 1024 bidisperse rings 16/22 (19456 vertices), WCA "
                               "vertex-segment, phi=0.85, T=1e-2, gamma=1, dt=1e-3 sigma, skin 0.5",
                   "vertices": nv, "rings": np_, "dt": dt, "rebuilds_in_timed_region": rebuilds,
                   "T_end": temp, "prep_s": round(prep_s, 1), "autotune": tuned},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

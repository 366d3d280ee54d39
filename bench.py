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
oracle built for this machine with OpenMP, at one thread and at the process's
CPU share, on a bounded sample of the same workload), followed by the batched
FIRE, adjoint, active-adhesive and rigid-ring figures.
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
# HBM bytes per launch of the step kernel on this workload, from the FETCH_SIZE and WRITE_SIZE
# counter runs of `bench.py --traffic` (the line's own warm-up and timed steps, so the same state
# and mean neighbour count 4.443; each counter scaled by its reading for a 1 GiB device copy:
# 2.000 and 1.000): 2 x 3834.32 KiB + 608.66 KiB.  The summary with its commands is
# profiles/k_step_traffic_nvt1024_wca_smooth.md; re-measure after changing the step kernel.
TRAFFIC_BYTES_PER_LAUNCH = 8475959
TRAFFIC_MEAN_NEIGHBORS = 4.443


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
    if args.ring_order == "spatial":
        from cudaring_amd import packing
        s.close()
        st, _ = packing.spatial_order(st)
        s = DPMSet(st["nvp"])
        s.load_state(st)
        s.set_types("wca", "vertexSmooth")
        s.set_energy_costs(*COSTS)
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


def run_steps(s, steps):
    """`steps` Langevin steps as graph replays of CHUNK steps."""
    done = 0
    while done < steps:
        n = min(CHUNK, steps - done)
        s.langevin(n)
        done += n


def cpu_model():
    """`lscpu` "Model name" of this machine (from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """The CPU share of this process: OMP_NUM_THREADS when set, else the cores it may run on."""
    n = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return n if n > 0 else len(os.sched_getaffinity(0))


def native_oracle():
    """Build the oracle -O3 -march=native -fopenmp for this machine (outside the tree) and load it;
    returns the build description, or the in-tree portable build's when no compiler is found."""
    import subprocess
    import tempfile
    from oracle import oracle
    out = tempfile.mkdtemp(prefix="oracle_native_")
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native", f"OUT={out}"],
                       capture_output=True, text=True)
    if r.returncode == 0:
        oracle.use_library(os.path.join(out, "liboracle_native.so"))
        return "gcc -O3 -march=native -fopenmp -ffp-contract=off, built on this machine"
    print(f"# native oracle build failed, timing the portable build: {r.stderr.strip()[:200]}", file=sys.stderr)
    return "gcc -O3 -fopenmp -ffp-contract=off (portable in-tree build)"


def cpu_baseline(st, dt, steps, runs=5):
    """The oracle (the same algorithm: cell-list neighbour rows, rebuild on skin crossing, FP64) on
    `steps` Langevin steps of the same prepared state, timed at one thread and at every thread of
    this process's CPU share (OpenMP over vertices): one untimed run, then the median of `runs`."""
    from oracle import oracle
    build = native_oracle()
    nv = int(np.sum(st["nvp"]))
    size = 2 * float(np.mean(st["rad"]))

    def one_run(threads):
        o = oracle.Oracle(st["nvp"])
        o.load_state(st)
        o.set_types("wca", "vertexSmooth")
        o.set_energy_costs(*COSTS)
        o.set_dt(dt)
        o.set_rng(SEED, 0)
        o.set_threads(threads)
        cut = o.set_displacement_cutoff(SKIN, size)
        o.calc_neighbor_list(cut)
        o.calc_force_energy()
        o.init_langevin(T, GAMMA)
        t0 = time.perf_counter()
        o.langevin(steps)
        return time.perf_counter() - t0

    def median_rate(threads):
        one_run(threads)  # warm-up
        return nv * steps / float(np.median([one_run(threads) for _ in range(runs)]))

    n = cpu_threads()
    single = median_rate(1)
    multi = median_rate(n) if n > 1 else single
    return {"value": multi, "unit": "vertex-steps/s", "cores": n, "kind": "port",
            "single_thread": single, "cpu_model": cpu_model(), "build": build,
            "sample": f"{steps} Langevin steps of the same {NRINGS}-ring WCA vertexSmooth state (cell-list rows, "
                      f"rebuilt on skin crossing), median of {runs} after 1 warm-up, at {n} threads and at 1"}


CFG1_SEED, CFG1_STEPS = 20261202 + 1, 10000


def nve_small_bench(args):
    """The plain NVE configuration: 32 bidisperse rings (16/22 vertices, 608 vertices), harmonic
    vertex-vertex, phi 0.80, uniform random centres, per-ring Gaussian velocities at T 1e-3 with
    the mean removed, dt = 0.01 sigma / sqrt(ec), skin 0.3; 1e4 velocity-Verlet steps on the oracle
    (one thread and the process's CPU share; median of 3 after 1 warm-up) and on the GPU, with
    the waves per tile tuned on the loop (untimed chunks of it) first."""
    from cudaring_amd import DPMSet, autotune, packing
    from oracle import oracle
    st = packing.random_system(32, 0.80, CFG1_SEED)
    nv = int(np.sum(st["nvp"]))
    sigma = float(2 * np.sqrt(np.mean(st["a0"]) / np.pi))
    dt, size = 0.01 * sigma, 2 * float(np.mean(st["rad"]))

    def setup(obj):
        obj.load_state(st)
        obj.set_types("harmonic", "vertexVertex")
        obj.set_energy_costs(*COSTS)
        obj.set_dt(dt)
        cut = obj.set_displacement_cutoff(0.3, size)
        obj.calc_neighbor_list(cut)
        obj.calc_force_energy()
        obj.init_nve(1e-3, zero_momentum=True)

    def cpu_run(threads):
        o = oracle.Oracle(st["nvp"])
        o.set_rng(CFG1_SEED, 0)
        o.set_threads(threads)
        setup(o)
        t0 = time.perf_counter()
        o.nve(CFG1_STEPS)
        return time.perf_counter() - t0

    def cpu_rate(threads):
        cpu_run(threads)
        return nv * CFG1_STEPS / float(np.median([cpu_run(threads) for _ in range(3)]))

    g = DPMSet(st["nvp"])
    g.set_seed(CFG1_SEED)
    setup(g)
    chunk = CFG1_STEPS // 10
    # the list rebuilds about every third step here, so the launch shape is tuned on the loop itself
    # (chunks of it; the graph of a chunk is captured for the chosen shape, which the timed calls replay)
    waves, _ = autotune.tune_waves_per_tile_loop(g, lambda: g.nve(chunk))
    g.nve(chunk)  # setting the chosen shape dropped the graphs: capture the chunk's graph again
    g.synchronize()
    t0 = time.perf_counter()
    for _ in range(CFG1_STEPS // chunk):
        g.nve(chunk)
    g.synchronize()
    gpu = nv * CFG1_STEPS / (time.perf_counter() - t0)
    g.close()
    n = cpu_threads()
    out = {"workload": "NVE, 32 rings 16/22 (608 vertices), harmonic vertexVertex, phi 0.80, T 1e-3, "
                       "dt 0.01 sigma, 1e4 steps", "gpu_vertex_steps_per_s": round(gpu, 1),
           "waves_per_tile": waves}
    if not args.no_cpu:
        out.update({"cpu_vertex_steps_per_s_1_thread": round(cpu_rate(1), 1),
                    f"cpu_vertex_steps_per_s_{n}_threads": round(cpu_rate(n), 1)})
    return out


FIRE_SYSTEMS, FIRE_SEED = 512, 20261202 + 3
FIRE_PARAMS = (0.1, 0.5, 1.1, 0.99, 0.01, 0.1, 0.1)  # a_start, f_dec, f_inc, f_a, dt0, dt_max, a
FIRE_TOL, FIRE_DPHI, FIRE_MAXIT = 1e-12, 1e-3, 100000
# ka = kl = 1 as in the NVE configuration; kb = 0.1: with kb = 0.01 and calA0 > 1 the rings'
# soft bending modes keep FIRE (alpha decaying without bound, FIRE.cu:338-347) above 1e-12
# for more than 1e5 iterations, on the oracle as on the GPU
FIRE_COSTS = (1.0, 1.0, 0.1, 1.0)


def fire_systems():
    """512 dilute packings at preferred phi 0.5: N ~ U{16..256} bidisperse rings, calA0 ~ U[1.0, 1.2]
    per packing, centres on a jittered lattice (random, no ring crossings)."""
    from cudaring_amd import packing
    rng = np.random.default_rng(FIRE_SEED)
    sizes = rng.integers(16, 257, FIRE_SYSTEMS)
    calA0 = rng.uniform(1.0, 1.2, FIRE_SYSTEMS)
    return [packing.lattice_system(int(n), 0.5, FIRE_SEED + i, calA0=float(c), perturb=0.15)
            for i, (n, c) in enumerate(zip(sizes, calA0))]


def fire_setup(s, size, skin=0.3, cell_scale=1.0):
    """harmonic vertex-vertex, FIRE_COSTS, the neighbour skin (cutoff in units of `size`) and cell
    scale (the same for every path)."""
    s.set_types("harmonic", "vertexVertex")
    s.set_energy_costs(*FIRE_COSTS)
    cut = s.set_displacement_cutoff(skin, size)
    s.set_cell_scale(cell_scale)
    return cut


def fire_step(s, cut, mode):
    """One compression step of every system in `s` (compressFIRE.cpp:104-131, 153): grow by
    sqrt((phi + dphi) / phi), rebuild, FIRE to max|F_c| <= tol.  `mode` "engine": the set's own
    fire_minimize (every system to its own convergence, persistent kernel); "host": the reference
    driver's loop, one FIRE iteration and one host check of max|F_c| per call.  Returns
    iterations per system."""
    phi = s.phi()
    s.scale_vertices(np.sqrt((phi + FIRE_DPHI) / phi))
    s.calc_neighbor_list(cut)
    s.calc_force_energy()
    s.reset_last_positions()
    s.init_fire(FIRE_PARAMS, 20, 0)
    if mode == "engine":
        return s.fire_minimize(FIRE_TOL, FIRE_MAXIT)
    it, f = 0, s.max_unbalanced_force()[0]
    while f > FIRE_TOL and it < FIRE_MAXIT:  # the driver loop: one iteration, one host check
        s.fire(1)
        f = s.max_unbalanced_force()[0]
        it += 1
    return np.array([it])


def batched_fire(args):
    """cfg3: packings/s of the batched set against two one-system paths on a stratified sample of
    the same packings: the engine's own fire_minimize on a set holding one system, and the
    reference driver's loop (one iteration, one host check per call)."""
    from cudaring_amd import DPMSet, autotune, packing
    systems = fire_systems()
    st = packing.concat_systems(systems)
    s = DPMSet(st["nvp"], rings_per_system=st["rings_per_system"])
    s.load_state(st)
    size = 2 * float(np.mean(st["rad"]))
    cut = fire_setup(s, size)
    s.calc_neighbor_list(cut)
    s.calc_force_energy()
    s.init_fire(FIRE_PARAMS, 20, 0)
    t0 = time.perf_counter()
    it0 = s.fire_minimize(FIRE_TOL, FIRE_MAXIT)  # relax the dilute start (untimed)
    relax_s = time.perf_counter() - t0
    start = {k: s.get(k) for k in ("pos", "vel", "rad", "l0", "a0")}  # FIRE keeps its velocities
    nv_sys = np.array([int(np.sum(x["nvp"])) for x in systems])
    fire_kernel, fire_tune = autotune.tune_fire(s, FIRE_PARAMS, cut)
    (skin, cell_scale), cut, nbr_tune = autotune.tune_fire_neighbors(s, FIRE_PARAMS, size)

    s.synchronize()
    t0 = time.perf_counter()
    iters = np.zeros(FIRE_SYSTEMS, dtype=np.int64)
    for _ in range(args.fire_steps):
        iters += fire_step(s, cut, "engine")
    s.synchronize()
    t_batch = time.perf_counter() - t0
    done = FIRE_SYSTEMS * args.fire_steps
    converged = int(np.sum(s.max_unbalanced_force() <= FIRE_TOL))

    # one-system paths on a sample stratified by size (every k-th packing in order of N), from the
    # same relaxed start through the same phi steps
    order = np.argsort(nv_sys, kind="stable")
    pick = order[np.linspace(0, FIRE_SYSTEMS - 1, args.fire_plain).round().astype(int)]
    voff = np.concatenate([[0], np.cumsum(nv_sys)])
    roff = np.concatenate([[0], np.cumsum(st["rings_per_system"])])
    t_path = {"engine": 0.0, "host": 0.0}
    it_path = {"engine": 0, "host": 0}
    for i in pick:
        sl, rl = slice(voff[i], voff[i + 1]), slice(roff[i], roff[i + 1])
        sub = dict(systems[i])
        sub["pos"] = start["pos"][2 * voff[i]:2 * voff[i + 1]]
        sub["rad"], sub["l0"], sub["a0"] = start["rad"][sl], start["l0"][sl], start["a0"][rl]
        for mode in ("engine", "host"):
            p = DPMSet(sub["nvp"])
            p.load_state(sub)
            p.set("vel", start["vel"][2 * voff[i]:2 * voff[i + 1]])
            pcut = fire_setup(p, size, skin, cell_scale)
            p.set_fire_kernel(*fire_kernel)
            p.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.fire_steps):
                it_path[mode] += int(fire_step(p, pcut, mode)[0])
            p.synchronize()
            t_path[mode] += time.perf_counter() - t0
            p.close()
    n_sample = len(pick) * args.fire_steps
    engine_rate, host_rate = n_sample / t_path["engine"], n_sample / t_path["host"]
    cpu = None if args.no_cpu else cpu_fire(systems, pick, start, voff, roff, size, skin)
    batch_rate = done / t_batch
    vit = float(np.sum(iters * nv_sys))
    s.close()
    return {"packings_per_s": round(batch_rate, 2),
            "single_system_packings_per_s": round(engine_rate, 3),
            "speedup_vs_single_system": round(batch_rate / engine_rate, 2),
            "plain_packings_per_s": round(host_rate, 3),
            "speedup_vs_plain": round(batch_rate / host_rate, 2),
            "systems": FIRE_SYSTEMS, "vertices": int(nv_sys.sum()), "phi_steps": args.fire_steps,
            "dphi": FIRE_DPHI, "tol": FIRE_TOL, "converged_systems": converged,
            "batched_s": round(t_batch, 3), "iterations_mean": round(float(iters.mean()), 1),
            "iterations_max": int(iters.max()), "vertex_fire_iterations_per_s": round(vit / t_batch, 1),
            "sample": f"{len(pick)} of the {FIRE_SYSTEMS} packings, every {FIRE_SYSTEMS // len(pick)}-th in order of "
                      f"size (N {int(nv_sys[pick].min())}..{int(nv_sys[pick].max())} vertices), same start and phi "
                      f"steps, one set each: the engine's fire_minimize {it_path['engine']} iterations in "
                      f"{t_path['engine']:.2f} s; the driver loop with a host check per iteration "
                      f"{it_path['host']} iterations in {t_path['host']:.2f} s",
            "start_relax_iterations_max": int(it0.max()), "start_relax_s": round(relax_s, 2),
            "fire_kernel": {"persistent": bool(fire_kernel[0]), "waves": int(fire_kernel[1])},
            "fire_tune_s": fire_tune, "skin": skin, "cell_scale": cell_scale, "neighbor_tune_s": nbr_tune,
            "cpu": cpu}


def cpu_fire(systems, pick, start, voff, roff, size, skin):
    """The oracle on the sampled packings' first phi step (same start, same skin), one packing per
    thread across the process's CPU share: packings/s over the wall time, and at one thread from
    the sum of the per-packing times (each packing runs single-threaded)."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle

    def one(i):
        sl, rl = slice(voff[i], voff[i + 1]), slice(roff[i], roff[i + 1])
        sub = dict(systems[i])
        sub["pos"] = start["pos"][2 * voff[i]:2 * voff[i + 1]]
        sub["vel"] = start["vel"][2 * voff[i]:2 * voff[i + 1]]
        sub["rad"], sub["l0"], sub["a0"] = start["rad"][sl], start["l0"][sl], start["a0"][rl]
        o = oracle.Oracle(sub["nvp"])
        o.load_state(sub)
        o.set_threads(1)
        o.set_types("harmonic", "vertexVertex")
        o.set_energy_costs(*FIRE_COSTS)
        cut = o.set_displacement_cutoff(skin, size)
        t0 = time.perf_counter()
        o.calc_neighbor_list(cut)
        o.calc_force_energy()  # the shapes phi() reads
        phi = o.phi()
        o.scale_vertices(np.sqrt((phi + FIRE_DPHI) / phi))
        o.calc_neighbor_list(cut)
        o.calc_force_energy()
        o.init_fire(FIRE_PARAMS, 20, 0)
        it = o.fire_minimize(FIRE_TOL, FIRE_MAXIT)
        return time.perf_counter() - t0, int(it), bool(o.max_unbalanced_force() <= FIRE_TOL)

    n = cpu_threads()
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=n) as ex:
        res = list(ex.map(one, list(pick)))
    wall = time.perf_counter() - t0
    serial = sum(r[0] for r in res)
    return {"packings_per_s": round(len(pick) / wall, 3), "threads": n,
            "packings_per_s_1_thread": round(len(pick) / serial, 3),
            "iterations": int(sum(r[1] for r in res)), "converged": int(sum(r[2] for r in res)),
            "sample": f"the first phi step of the {len(pick)} sampled packings, one per thread"}


ADJ_RINGS, ADJ_PHI, ADJ_SEED, ADJ_T = 256, 0.85, 20261202 + 4, 1e-3


def adjoint_bench(args, torch):
    """cfg4: 256 rings at phi 0.85, harmonic vertex-vertex, NVE at dt = 0.01 sigma; the reverse pass
    (gradients of a linear functional of (x_N, v_N) with respect to x_0, v_0, per-ring ka, kl, kb
    and l0) against the checkpointed forward pass and against plain NVE steps."""
    from cudaring_amd import DPMSet, protocols
    s, st = protocols.prepared_state(lambda nvp: DPMSet(nvp), ADJ_RINGS, ADJ_PHI, ADJ_SEED, costs=COSTS,
                                     iters=args.prep_iters)
    sigma = float(2 * np.sqrt(np.mean(s.get("a0")) / np.pi))
    s.set_dt(0.01 * sigma)
    s.set_seed(ADJ_SEED)
    nr = s.np
    s.set_ring_costs(np.full(nr, COSTS[0]), np.full(nr, COSTS[1]), np.full(nr, COSTS[2]))
    size = 2 * float(np.mean(s.get("rad")))
    cut = s.set_displacement_cutoff(0.3, size)
    s.calc_neighbor_list(cut)
    s.calc_force_energy()
    s.init_nve(ADJ_T, zero_momentum=True)
    x0, v0 = s.get("pos"), s.get("vel")
    steps, every = args.adj_steps, 100

    def restart():
        s.set("pos", x0)
        s.set("vel", v0)
        s.calc_neighbor_list(cut)
        s.calc_force_energy()
        s.synchronize()

    restart()
    s.nve(steps)  # capture the graphs of both paths once
    restart()
    s.nve_forward(steps, every)
    restart()
    t0 = time.perf_counter()
    s.nve(steps)
    s.synchronize()
    t_plain = time.perf_counter() - t0
    restart()
    t0 = time.perf_counter()
    s.nve_forward(steps, every)
    s.synchronize()
    t_fwd = time.perf_counter() - t0
    rng = np.random.default_rng(ADJ_SEED)
    nv = s.nv
    ax = torch.as_tensor(rng.standard_normal(2 * nv), device="cuda")
    av = torch.as_tensor(rng.standard_normal(2 * nv), device="cuda")
    outs = [torch.empty(2 * nv, dtype=torch.float64, device="cuda") for _ in range(2)]
    gring = torch.empty(3 * nr, dtype=torch.float64, device="cuda")
    gl0 = torch.empty(nv, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    s.nve_backward(ax.data_ptr(), av.data_ptr(), outs[0].data_ptr(), outs[1].data_ptr(), gring.data_ptr(),
                   gl0.data_ptr())  # first call captures the reverse-pass graphs
    restart()
    s.nve_forward(steps, every)
    s.synchronize()
    t0 = time.perf_counter()
    s.nve_backward(ax.data_ptr(), av.data_ptr(), outs[0].data_ptr(), outs[1].data_ptr(), gring.data_ptr(),
                   gl0.data_ptr())
    s.synchronize()
    t_bwd = time.perf_counter() - t0
    finite = bool(torch.isfinite(outs[0]).all() and torch.isfinite(gring).all())
    s.close()
    return {"adjoint_over_forward": round(t_bwd / t_fwd, 3), "adjoint_over_plain_nve": round(t_bwd / t_plain, 3),
            "forward_s": round(t_fwd, 4), "backward_s": round(t_bwd, 4), "plain_nve_s": round(t_plain, 4),
            "rings": nr, "vertices": nv, "steps": steps, "checkpoint_every": every, "finite": finite,
            "workload": f"NVE harmonic vertexVertex, {nr} rings 16/22 at phi {ADJ_PHI}, dt 0.01 sigma, T {ADJ_T}; "
                        "gradients wrt x0, v0, per-ring ka/kl/kb and l0"}


CFG5_SEED = 20261202 + 5
CFG5_COSTS = (1e5, 20.0, 10.0, 1.0)  # ea, el, eb, ec of the active driver, runActive.cpp:39


def cfg5_bench(args, torch, batched=False):
    """Large active + adhesive: rings of 32 vertices in two sizes (ratio 1.4), vertex-segment
    interaction with the adhesive potential (l1 = 0.1, l2 = 0.2), active Brownian driving
    f0 = v0 gamma with v0 = 0.1 and taup = 1/Dr = 100, phi 0.90, shape costs of the reference's
    active driver.  One system of 16384 rings, or (batched) 64 systems with N ~ U{2048..8192} in
    one set.  Compressed with harmonic vertex-vertex FIRE from phi 0.5 (ka = kl = 1), then relaxed
    with the production interaction and costs, before timing."""
    from cudaring_amd import DPMSet, autotune, packing, protocols
    rng = np.random.default_rng(CFG5_SEED)
    sizes = rng.integers(2048, 8193, 64) if batched else [16384]
    systems = [packing.size_bidisperse_system(int(n), 0.5, CFG5_SEED + i) for i, n in enumerate(sizes)]
    if args.ring_order == "spatial":
        systems = [packing.spatial_order(x)[0] for x in systems]
    st = packing.concat_systems(systems)
    t0 = time.perf_counter()
    s = DPMSet(st["nvp"], rings_per_system=st["rings_per_system"])
    s.load_state(st)
    s.set_types("harmonic", "vertexVertex")
    s.set_energy_costs(*COSTS)
    protocols.compress(s, 0.5, 0.90, dphi=args.cfg5_dphi, iters=args.cfg5_iters,
                       progress=lambda phi: print(f"# compressed to phi {phi:.3f}", file=sys.stderr, flush=True))
    sigma = float(2 * np.sqrt(np.mean(s.get("a0")) / np.pi))
    # the active driver's shape costs and the production interaction, relaxed before the dynamics
    # (FIRE dt well inside the stability limit of the ea = 1e5 area mode)
    s.set_types("adhesive", "vertexSmooth")
    s.set_attraction(0.1, 0.2)
    s.set_energy_costs(*CFG5_COSTS)
    size = 2 * float(np.mean(s.get("rad")))
    cut = s.set_displacement_cutoff(SKIN, size)  # cutoff 0.5, runActive.cpp:37,130
    s.calc_neighbor_list(cut)
    s.calc_force_energy()
    protocols.fire_relax(s, 0.002 * size, 2 * args.cfg5_iters)
    s.set_types("adhesive", "vertexSmooth", particle="active")
    # units of runActive.cpp:113-123: driving in ec / sigma, taup and 1 / damping in sigma / sqrt(ec)
    s.set_self_propulsion(0.1 * GAMMA / sigma, 100.0 * sigma)
    s.set("pangle", rng.uniform(-np.pi, np.pi, s.np))
    s.set_dt(DT_UNITS * sigma)
    s.set_seed(CFG5_SEED)
    s.calc_neighbor_list(cut)
    s.calc_force_energy()
    s.init_langevin(1e-4, GAMMA / sigma)
    waves, wtimes = autotune.tune_waves_per_tile(s, reps=20)
    prep = time.perf_counter() - t0
    for _ in range(2):
        s.langevin(CHUNK)
    s.synchronize()
    s.reset_update_count()
    t0 = time.perf_counter()
    done = 0
    while done < args.cfg5_steps:
        n = min(CHUNK, args.cfg5_steps - done)
        s.langevin(n)
        done += n
    s.synchronize()
    el = time.perf_counter() - t0
    us, kbar = s.time_step_kernel(50)
    alg = alg_bytes_per_step(s.nv, s.np, kbar)
    out = {"vertex_steps_per_s": s.nv * args.cfg5_steps / el, "vertices": s.nv, "rings": s.np,
           "systems": len(sizes), "steps": args.cfg5_steps, "step_us": round(1e6 * el / args.cfg5_steps, 2),
           "kernel_us": round(us, 2), "mean_neighbors": round(kbar, 3),
           "achieved_GBps": round(alg / (us * 1e-6) / 1e9, 1),
           "frac_hbm": round(alg / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
           "rebuilds": [int(x) for x in s.update_count()[:4]], "T_end": float(np.mean(s.temperature())),
           "prep_s": round(prep, 1), "waves_per_tile": waves,
           "finite": bool(np.all(np.isfinite(s.get("pos"))))}
    if not args.no_cpu and not batched:
        out["cpu"] = cpu_cfg5(s, st, sigma, cut)
    s.close()
    return out


def cpu_cfg5(s, st, sigma, cut, steps=10):
    """The oracle on the active-adhesive state the GPU run ended in: `steps` Langevin steps at the
    process's CPU share and at one thread (median of 3 after 1 warm-up)."""
    from oracle import oracle
    state = {"nvp": st["nvp"], "theta0": st["theta0"], "box": s.get("box")}
    for k in ("pos", "vel", "rad", "l0", "a0", "pangle"):
        state[k] = s.get(k)

    def run(threads):
        o = oracle.Oracle(state["nvp"])
        o.load_state(state)
        o.set_threads(threads)
        o.set_types("adhesive", "vertexSmooth", "harmonic", "active")
        o.set_attraction(0.1, 0.2)
        o.set_energy_costs(*CFG5_COSTS)
        o.set_self_propulsion(0.1 * GAMMA / sigma, 100.0 * sigma)
        o.set_dt(DT_UNITS * sigma)
        o.set_rng(CFG5_SEED, 0)
        o.set_displacement_cutoff(SKIN, 2 * float(np.mean(state["rad"])))
        o.calc_neighbor_list(cut)
        o.calc_force_energy()
        o.init_langevin(1e-4, GAMMA / sigma, read_state=True)
        t0 = time.perf_counter()
        o.langevin(steps)
        return time.perf_counter() - t0

    def rate(threads):
        run(threads)
        return len(state["rad"]) * steps / float(np.median([run(threads) for _ in range(3)]))

    n = cpu_threads()
    return {"vertex_steps_per_s": round(rate(n), 1), "threads": n, "vertex_steps_per_s_1_thread": round(rate(1), 1),
            "sample": f"{steps} Langevin steps of the final state, median of 3 after 1 warm-up"}


def integrators_bench(s, dt, steps):
    """The benchmark workload under each integrator of the Langevin loop (Simulator.cu:23-30,
    112-203): vertex-steps/s over `steps` steps after 1,000 untimed ones, from the same state.
    The overdamped schemes run at dt/100: x += dt F / gamma is stable only for dt k / gamma < 2,
    and the WCA contact stiffness k ~ 456 ec / sigma_v^2 puts dt = 1e-3 sigma far above that."""
    out = {}
    for noise in ("langevin", "baoab", "brownian", "drivenBrownian"):
        s.set_noise_type(noise)
        s.set_dt(dt if noise in ("langevin", "baoab") else dt / 100)
        s.init_langevin(T, GAMMA, read_state=True)
        run_steps(s, 1000)
        s.synchronize()
        t0 = time.perf_counter()
        run_steps(s, steps)
        s.synchronize()
        el = time.perf_counter() - t0
        out[noise] = {"vertex_steps_per_s": round(s.nv * steps / el, 1), "step_us": round(1e6 * el / steps, 3),
                      "T_end": float(s.temperature()[0]), "finite": bool(np.all(np.isfinite(s.get("pos"))))}
    s.set_noise_type("langevin")
    s.set_dt(dt)
    return out


def rigid_bench(s, dt, steps):
    """The benchmark packing as rigid rings (RigidNVE, RigidLangevin and rigid FIRE; Simulator.cu:
    241-371, FIRE.cu:173-298): vertex-steps/s over `steps` steps (iterations), run as calls of
    CHUNK steps after one untimed call (which captures the graph the timed calls replay).  The rigid Langevin kick adds unit Gaussian draws to the ring velocities in both half
    kicks whatever T and dt (per step v -> c^2 v + (1 + c) z, c = (1 - h gamma)(1 + h^2 / 2)), so
    it runs at gamma = 1 / dt, where the ring speeds settle near 1.5 instead of ~sqrt(2 / (dt
    gamma)) = 45 at gamma = 1."""
    a0 = s.get("a0")
    s.set("prad", np.sqrt(a0 / np.pi))
    pos0 = s.get("pos")
    cut = s.set_displacement_cutoff(SKIN, 2 * s.get("rad").mean())
    out = {}
    for kind in ("nve", "langevin", "fire"):
        # every run from the benchmark packing at rest
        s.set("pos", pos0)
        s.set("pvel", np.zeros(2 * s.np))
        s.calc_neighbor_list(cut)
        s.calc_rigid_force_energy()
        if kind == "nve":
            s.init_rigid_nve(T)
            loop = s.rigid_nve
        elif kind == "langevin":
            s.init_rigid_langevin(T, 1.0 / dt)
            loop = s.rigid_langevin
        else:
            s.init_rigid_fire([0.2, 0.5, 1.1, 0.99, dt, 10 * dt, 0.2], 20, 0)
            loop = s.rigid_fire
        s.calc_rigid_force_energy()
        f_start = float(s.rigid_max_unbalanced_force()[0])
        try:
            loop(CHUNK)
            s.synchronize()
            s.reset_update_count()
            t0 = time.perf_counter()
            done = 0
            while done < steps:
                n = min(CHUNK, steps - done)
                loop(n)
                done += n
            s.synchronize()
            el = time.perf_counter() - t0
        except Exception as exc:  # a run the list width cannot hold: report it, time the others
            out[kind] = {"error": str(exc)}
            continue
        out[kind] = {"vertex_steps_per_s": round(s.nv * steps / el, 1), "step_us": round(1e6 * el / steps, 3),
                     "rebuilds": int(s.update_count()[0]), "steps": steps,
                     "finite": bool(np.all(np.isfinite(s.get("pos"))))}
        if kind == "fire":
            out[kind]["max_force_torque_start"] = f_start
            out[kind]["max_force_torque"] = float(s.rigid_max_unbalanced_force()[0])
    return out


COPY_ELEMS = 1 << 27  # float64 elements of the calibration copy: 1 GiB read + 1 GiB written


def traffic_run(s, torch):
    """Launches for `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE`, after the benchmark's own
    warm-up and timed steps: three copies of COPY_ELEMS doubles (the calibration kernel:
    8 * COPY_ELEMS bytes read and written each), then 50 back-to-back launches of the step kernel
    on the state the roofline figure is measured on."""
    src = torch.rand(COPY_ELEMS, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
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
    ap.add_argument("--equil", type=int, default=20000,
                    help="untimed Langevin steps from the prepared packing before warm-up and timing")
    ap.add_argument("--cpu-steps", type=int, default=300)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ring-order", choices=("random", "spatial"), default="random",
                    help="ring numbering of the prepared packing: as generated, or along a Morton curve")
    ap.add_argument("--fire-steps", type=int, default=5, help="phi steps of the batched FIRE compression")
    ap.add_argument("--fire-plain", type=int, default=16,
                    help="packings (stratified by size) timed on the one-system paths")
    ap.add_argument("--no-fire", action="store_true")
    ap.add_argument("--adj-steps", type=int, default=2000)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--fire-only", action="store_true")
    ap.add_argument("--adjoint-only", action="store_true")
    ap.add_argument("--cfg5", action="store_true", help="large active + adhesive system (16384 rings)")
    ap.add_argument("--cfg5-batched", action="store_true", help="64 active + adhesive systems in one set")
    ap.add_argument("--no-cfg5", action="store_true", help="skip the large active + adhesive run of the default line")
    ap.add_argument("--cfg5-steps", type=int, default=2000)
    ap.add_argument("--cfg5-dphi", type=float, default=0.02)
    ap.add_argument("--cfg5-iters", type=int, default=2000)
    ap.add_argument("--integrators", action="store_true",
                    help="the benchmark workload under each Langevin-loop integrator (5,000 steps each)")
    ap.add_argument("--no-rigid", action="store_true", help="skip the rigid-ring runs of the default line")
    ap.add_argument("--rigid", action="store_true",
                    help="the benchmark packing as rigid rings: rigid NVE, Langevin and FIRE throughput")
    ap.add_argument("--traffic", action="store_true",
                    help="counter-collection run: the line's warm-up and timed steps, then a copy of known "
                         "size and step-kernel launches only")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the engine measures one GPU's share (batched systems "
              "would shard across GPUs with no exchange).  Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch  # device plumbing: synchronisation around the timed region
    if args.fire_only:
        print(json.dumps({"batched_fire": batched_fire(args)}), flush=True)
        return
    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_bench(args, torch)}), flush=True)
        return
    if args.cfg5 or args.cfg5_batched:
        print(json.dumps({"active_adhesive": cfg5_bench(args, torch, batched=args.cfg5_batched)}), flush=True)
        return
    s, st, dt, prep_s, tuned = prepare(args)
    nv, np_ = s.nv, s.np
    # untimed: relax the prepared packing to the thermostat's steady state (its neighbour count and
    # rebuild interval), counting the rebuilds of the equilibration's second half
    run_steps(s, args.equil // 2)
    s.synchronize()
    s.reset_update_count()
    run_steps(s, args.equil - args.equil // 2)
    s.synchronize()
    equil_rebuilds = int(s.update_count()[0])
    if args.integrators:
        print(json.dumps({"integrators": integrators_bench(s, dt, 5000)}), flush=True)
        return
    if args.rigid:
        print(json.dumps({"rigid": rigid_bench(s, dt, 5000)}), flush=True)
        return

    # warm-up, then one replay of every chunk shape the timed loop uses (their graphs captured)
    run_steps(s, args.warmup)
    for n in sorted({min(CHUNK, args.steps), args.steps % CHUNK} - {0}):
        run_steps(s, n)
    s.synchronize()
    torch.cuda.synchronize()
    s.reset_update_count()
    t0 = time.perf_counter()
    run_steps(s, args.steps)
    s.synchronize()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    rebuilds = int(s.update_count()[0])
    temp = float(s.temperature()[0])
    if not np.isfinite(temp):
        raise RuntimeError("non-finite temperature after the timed run")
    if args.traffic:
        traffic_run(s, torch)
        return

    value = nv * args.steps / el
    us, kbar = s.time_step_kernel(200)
    alg = alg_bytes_per_step(nv, np_, kbar)
    achieved = alg / (us * 1e-6) / 1e9
    rigid = None if args.no_rigid else rigid_bench(s, dt, 2000)
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": TRAFFIC_BYTES_PER_LAUNCH,
            "traffic_mean_neighbors": TRAFFIC_MEAN_NEIGHBORS,
            "traffic_source": "profiles/k_step_traffic_nvt1024_wca_smooth.md",
            "kernel": "k_step<vertexSmooth,WCA,kick|langevin|start>", "kernel_us": round(us, 3),
            "alg_bytes_per_launch": int(alg), "mean_neighbors": round(kbar, 3),
            "step_us": round(1e6 * el / args.steps, 3)}
    s.close()
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(st, dt, args.cpu_steps)
    fire = None if args.no_fire else batched_fire(args)
    adjoint = None if args.no_adjoint else adjoint_bench(args, torch)
    active = None if args.no_cfg5 else cfg5_bench(args, torch)
    nve_small = nve_small_bench(args)
    out = {
        "metric": METRIC, "value": value, "unit": "vertex-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "NVT Langevin, 1024 bidisperse rings 16/22 (19456 vertices), WCA "
                               "vertex-segment, phi=0.85, T=1e-2, gamma=1, dt=1e-3 sigma, skin 0.5",
                   "vertices": nv, "rings": np_, "dt": dt, "rebuilds_in_timed_region": rebuilds,
                   "equilibration_steps": args.equil,
                   "steady_rebuild_interval": (round((args.equil - args.equil // 2) / equil_rebuilds, 1)
                                               if equil_rebuilds else None),
                   "T_end": temp, "prep_s": round(prep_s, 1), "autotune": tuned,
                   "vs_reference_driver": "shape costs ka, kl, kb = 1, 1, 0.01 (runNVT.cpp:38 uses "
                                          "1e5, 20, 10) and gamma = 1 in simulation units (runNVT.cpp:140 "
                                          "divides the damping by its time unit); neither changes the "
                                          "work per step"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "batched_fire": fire,
        "adjoint": adjoint,
        "active_adhesive": active,
        "nve_small": nve_small,
        "rigid": rigid,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

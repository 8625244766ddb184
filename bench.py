"""Benchmark: particle-substeps/s of the 401,856-particle dam break on one GPU.

One "step" is one solver substep (``Solver*.step``: grid build + every
neighbour pass of that solver) over the full dam-break scene
(``make_scene(400000)`` -> 401,856 particles), inputs resident in HBM.  The
headline ``value`` is WCSPH (the solver the reference publishes a substep rate
for); DFSPH and PBF are measured the same way and reported under
``solvers``.  Each solver first runs ``--settle-frames`` frames (default 30,
half a second of the collapse) with the driver's dt policy, untimed, so the
timed substeps see the flow the reference's published 60-frame runs average
over rather than the initial lattice at rest; the initial state's rate is
kept under ``initial_state``.  dt follows the dam-break driver's policy,
evaluated once on the device state before the timed region so the timed
substeps run back to back without host synchronisation.

``roofline`` covers the dominant kernel of the WCSPH substep: algorithmic bytes
per launch (the per-particle byte model below x 401,856) over its mean launch
duration, from HIP events on the launch stream.  ``cpu_baseline`` times the
oracle's FP32 restatement of the same WCSPH substep on the host cores,
starting from the state the timed GPU substeps start from.

The default run also adds, under ``extras``: the 64-scene batched substep
(graph-replayed, eager and 64 sequential solves), the differentiable WCSPH
rollout at 50,540 particles x 200 substeps (forward, forward + backward), and
whole WCSPH frames with the host CFL driver beside the device-resident one,
and the APIC particle-separation pass over the jittered dam-break lattice.
``roofline.traffic`` is the HBM bytes per launch of the dominant kernel, read
from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary for the launch
variant the autotuner selected.  ``valu_issue`` puts the same launch against
the roof that binds it: SQ_INSTS_VALU per launch (committed SQ summary) over
the live launch duration, against one wave64 VALU instruction per CU per
cycle.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--solvers wcsph,dfsph,pbf]
       [--settle-frames F] [--no-tune] [--no-extras] [--no-cpu-baseline]
Other modes (each prints its own JSON lines and exits):
  --micro 4000000,16000000   grid build / density / force at 4M and 16M points,
                             uniform and clustered (median of >= 20 substeps)
  --batched 64               the batched form alone
  --adjoint [--no-fd]        the differentiable rollout, with an FP64 central
                             finite-difference check unless --no-fd
  --frames N                 N whole WCSPH frames, host and device-resident drivers
  --separation               the APIC particle-separation pass alone
  --copy-calibration         a 256 MiB device copy (the PMC byte-counter calibration)
"""
from __future__ import annotations

import argparse
import json
import os
import re
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PUBLISHED_WCSPH = 62.78e6     # particle-substeps/s derived from the reference's WCSPH run (RTX 4090)

# Algorithmic bytes per particle of each kernel (one launch processes every
# particle once): the pass byte model of the hot path -- each pass reads every
# per-particle array it uses once and writes each result once; neighbour
# gathers are on-chip reuse and cost nothing.
PASS_BYTES = {
    "k_wcsph_density": 44, "k_wcsph_density_coop": 44, "k_wcsph_force": 96,
    "k_dfsph_density_factor": 40, "k_dfsph_nonpressure": 68, "k_dfsph_rate": 64, "k_dfsph_correction": 60,
    "k_pbf_density": 32, "k_pbf_nonpressure": 76, "k_pbf_lambda": 44, "k_pbf_delta": 72, "k_pbf_xsph": 80,
}
# bytes per particle-substep of each solver (grid build counted at 28 B)
SOLVER_BYTES = {"wcsph": 168, "pbf": 760}


PMC_TRAFFIC = os.path.join(ROOT, "profiles", "wcsph_dam_break_401856_pmc_traffic.md")


def variant_key(kernel: str, v: dict) -> str:
    """The kernel's name with the template arguments of a launch variant, as
    rocprofv3 prints it."""
    if kernel == "k_wcsph_density" and v.get("coop"):
        kernel = "k_wcsph_density_coop"
        args = (v["list_depth"], v["coop"], "false", v["lanes_per_particle"], v["unroll"])
    elif kernel.startswith("k_wcsph_density"):
        args = (v["block"], v["list_depth"], "false", v["lanes_per_particle"], v["unroll"])
    else:
        args = (v["block"], v["list_depth"], "false", v["unroll"], v["particles_per_thread"], v["lanes_per_particle"])
    return f"{kernel}<{', '.join(str(a) for a in args)}>"


def pmc_traffic(key: str):
    """HBM bytes per launch of one launch variant of a kernel (variant_key),
    from the committed rocprofv3 PMC summary (FETCH_SIZE and WRITE_SIZE runs
    scaled by the copy calibration); None when the variant was not measured."""
    try:
        with open(PMC_TRAFFIC) as f:
            for line in f:
                if key in line and line.startswith("|"):
                    return float(line.strip().strip("|").split("|")[-1])
    except OSError:
        pass
    return None


PMC_CACHE = os.path.join(ROOT, "profiles", "wcsph_dam_break_401856_pmc_cache.md")


def pmc_cache(key: str):
    """(L1 hit, L2 hit) fractions of one launch variant (variant_key), from
    the committed rocprofv3 cache summary (TCP_TCC_READ_REQ against
    TCP_TOTAL_CACHE_ACCESSES, TCC_HIT against TCC_MISS); None when the
    variant was not measured."""
    try:
        with open(PMC_CACHE) as f:
            for line in f:
                if key in line and line.startswith("|"):
                    cells = [c.strip() for c in line.strip().strip("|").split("|")]
                    return float(cells[4].rstrip("%")) / 100.0, float(cells[7].rstrip("%")) / 100.0
    except (OSError, ValueError, IndexError):
        pass
    return None


PMC_SQ = os.path.join(ROOT, "profiles", "wcsph_dam_break_401856_pmc_sq.md")
# wave64 FP32 VALU issue of the whole chip: one instruction per CU per cycle
# (4 SIMDs of 16 lanes, 4 cycles each), 256 CUs at 2.4 GHz
VALU_ISSUE_PEAK = 256 * 2.4e9


def pmc_valu_instructions(key: str):
    """SQ_INSTS_VALU per launch of one launch variant (variant_key; waves x
    VALU/wave), from the committed rocprofv3 SQ summary; None when the
    variant was not measured."""
    try:
        with open(PMC_SQ) as f:
            for line in f:
                if key in line and line.startswith("|"):
                    cells = [c.strip() for c in line.strip().strip("|").split("|")]
                    return float(cells[2].replace(",", "")) * float(cells[3].replace(",", ""))
    except (OSError, ValueError, IndexError):
        pass
    return None


def dfsph_bytes(rho_it, rho_corr, vel_it, vel_corr):
    return 260 + 64 * rho_it + 60 * rho_corr + 56 * vel_it + 60 * vel_corr


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--particles", type=int, default=400000)
    ap.add_argument("--solvers", default="wcsph,dfsph,pbf")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-substeps", type=int, default=40)
    ap.add_argument("--settle-frames", type=int, default=30,
                    help="dam-break frames each solver runs (untimed, with the driver's dt policy) before the "
                         "timed substeps; 0 times the initial state")
    ap.add_argument("--copy-calibration", action="store_true", help="run only the PMC copy calibration kernel")
    ap.add_argument("--micro", default="", help="comma-separated particle counts for the grid/density/force "
                                                "microbenchmark (e.g. 4000000,16000000)")
    ap.add_argument("--batched", type=int, default=0, help="run only the batched form over this many scenes "
                                                           "(64 for the batched workload)")
    ap.add_argument("--adjoint", action="store_true", help="run only the differentiable WCSPH form (50,540 "
                                                            "particles, 200 substeps) with the FP64 FD check")
    ap.add_argument("--no-fd", action="store_true", help="skip the FD check of --adjoint")
    ap.add_argument("--adjoint-no-clamp", action="store_true", help="--adjoint with p = B((rho/rho0)^g - 1) "
                                                                    "unclamped (no max(p, 0) branch)")
    ap.add_argument("--frames", type=int, default=0, help="run only the whole-frame comparison (host CFL driver "
                                                          "vs device-resident) over this many frames")
    ap.add_argument("--separation", action="store_true", help="run only the APIC particle separation "
                                                                "(401,856 particles, 2 iterations)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch variants (no autotuning)")
    ap.add_argument("--no-extras", action="store_true", help="skip the batched and differentiable forms in the "
                                                             "default run")
    return ap.parse_args()


def driver_dt(sim):
    if sim.method == "pbf":
        return sim.frame_dt / sim.substeps
    return min(sim.frame_dt / sim.substeps, sim.cfl_limit(*sim._cfl_maxima()))


def _time_substeps(sim, dt, steps):
    import torch

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim._substep(dt)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def _state_tensors(sim):
    """Every device array the next substep reads: the current state (positions,
    velocities, sph attributes) and the solver's stored acceleration."""
    from fluid_simulation_amd.array import Array

    st = sim.state_0
    arrays = [st.particle_q, st.particle_qd, sim.solver._particle_accel]
    sph = getattr(st, "sph", None)
    if sph is not None:
        arrays += [v for v in vars(sph).values() if isinstance(v, Array)]
    return [a.tensor for a in arrays if a is not None]


def bench_solver(method, scene, steps, warmup, tune=True, settle_frames=0):
    """Time `steps` substeps of `method` at the state `settle_frames` frames
    into the dam break (the driver advances the run untimed first); the
    initial state's rate is recorded too, under ``initial_state``."""
    import torch

    from fluid_simulation_amd import _lib
    from fluid_simulation_amd.examples.dam_break.simulation import DamBreakSimulation

    sim = DamBreakSimulation(scene, method, device="cuda:0")
    n = scene.count
    initial = None
    if settle_frames > 0:
        # time the initial state, then put it back so that the settling run
        # starts from the scene as given
        saved = [t.clone() for t in _state_tensors(sim)]
        substeps0 = sim.total_substeps
        dt0 = driver_dt(sim)
        for _ in range(warmup):
            sim._substep(dt0)
        el0 = _time_substeps(sim, dt0, steps)
        initial = {"particle_substeps_per_s": n * steps / el0, "ms_per_substep": 1e3 * el0 / steps, "dt": dt0}
        for t, c in zip(_state_tensors(sim), saved):
            t.copy_(c)
        sim.total_substeps = substeps0
        del saved
        t_settle = time.perf_counter()
        for _ in range(settle_frames):
            sim.step()
        torch.cuda.synchronize()
        print(f"# {method}: advanced {settle_frames} frames in {time.perf_counter() - t_settle:.1f} s",
              file=sys.stderr, flush=True)
    for _ in range(warmup):
        sim._substep(driver_dt(sim))
    dt = driver_dt(sim)
    tuned = None
    if tune:
        from fluid_simulation_amd.autotune import SOLVER_PASSES
        from fluid_simulation_amd.autotune import tune as tune_passes

        # variants give identical results: time them on scratch output state
        s_scratch = sim.model.state()
        tuned = tune_passes(lambda: sim.solver.step(sim.state_0, s_scratch, None, None, dt), SOLVER_PASSES[method])
        print(f"# {method} autotune: " + ", ".join(
            f"{k} -> block {v['block']} depth {v['list_depth']} unroll {v['unroll']} ppt {v['particles_per_thread']}"
            f" lpp {v['lanes_per_particle']}" + (f" coop {v['coop']}" if v.get("coop") else "")
            for k, v in tuned.items()), file=sys.stderr, flush=True)
    state = None
    if method == "wcsph":
        # the timed substeps start here: the CPU baseline restarts from this state
        state = (sim.state_0.particle_q.tensor.cpu().numpy().copy(),
                 sim.state_0.particle_qd.tensor.cpu().numpy().copy(), dt)
    elapsed = _time_substeps(sim, dt, steps)
    iters = sim.solver.last_iterations() if method == "dfsph" else None
    # per-kernel timing: the same substeps again with HIP events around each launch
    lib = _lib.load()
    lib.fs_profile_reset()
    lib.fs_profile_enable(1)
    for _ in range(steps):
        sim._substep(dt)
    lib.fs_profile_enable(0)
    prof = {}
    for name, (cnt, ms) in _lib.profile_report().items():
        key = re.sub(r"<.*>", "", name.strip("()"))
        c0, m0 = prof.get(key, (0, 0.0))
        prof[key] = (c0 + cnt, m0 + ms)
    sim.check_finite()
    rate = n * steps / elapsed
    out = {"particle_substeps_per_s": rate, "ms_per_substep": 1e3 * elapsed / steps, "dt": dt,
           "frames_before_timing": settle_frames, "sim_time_s": sim.sim_time}
    if initial:
        out["initial_state"] = initial
    if method == "wcsph":
        out["kernels_ms_per_substep"] = {k: v[1] / steps for k, v in sorted(prof.items())}
    else:
        # the event pairs around DFSPH's and PBF's many short launches measure
        # host enqueue gaps as well; their per-kernel split is the rocprof one
        out["kernel_times"] = f"profiles/{method}_dam_break_401856_kernel_stats.md"
    if tuned:
        out["autotune"] = {k: {"block": v["block"], "list_depth": v["list_depth"], "unroll": v["unroll"],
                               "particles_per_thread": v["particles_per_thread"],
                               "lanes_per_particle": v["lanes_per_particle"], "coop": v.get("coop", 0),
                               "us_per_step_by_variant": v["us_per_launch"]} for k, v in tuned.items()}
    if method == "dfsph":
        out["iterations_last_substep"] = iters
        b = dfsph_bytes(iters["density_iterations"], iters["density_corrections"], iters["divergence_iterations"],
                        iters["divergence_corrections"])
    else:
        b = SOLVER_BYTES[method]
    out["bytes_per_particle_substep"] = b
    out["effective_GBs"] = b * rate / 1e9
    out["hbm_frac"] = out["effective_GBs"] / HBM_PEAK_GBS
    return out, prof, state


def cpu_baseline(scene, substeps, threads=None, state=None):
    """The oracle's FP32 WCSPH substep on the host cores, from `state`
    ((x, v, dt) of the GPU run where its timed substeps start) or from the
    initial scene at dt = 1e-4."""
    import numpy as np

    from oracle import oracle as O
    from fluid_simulation_amd.examples.dam_break.simulation import material, solver_config

    threads = threads or min(os.cpu_count() or 1, 16)
    O.set_threads(threads)
    cfg = solver_config("wcsph", scene)
    mat = material("wcsph", scene)
    n = scene.count
    m = O.uniform_model(n, scene.mass, stiffness=mat["stiffness"], exponent=mat["exponent"],
                        viscosity=mat["viscosity"], surface_tension=mat["surface_tension"],
                        gravity=scene.spec.gravity, radius=scene.radius)
    m.world[:] = -1
    if state is not None:
        x0, v0, dt = state
        st = O.OState(np.ascontiguousarray(x0, dtype=np.float32), np.ascontiguousarray(v0, dtype=np.float32))
    else:
        st = O.OState(scene.positions.copy(), np.zeros_like(scene.positions))
        dt = 1.0e-4
    ocfg = O.WcsphConfig(**{k: getattr(cfg, k) for k in O.WcsphConfig.__dataclass_fields__})
    O.wcsph_step(m, st, ocfg, dt)   # warm-up (page-in, thread pool)
    t0 = time.perf_counter()
    for _ in range(substeps):
        st, _ = O.wcsph_step(m, st, ocfg, dt)
    el = time.perf_counter() - t0
    where = "from the GPU run's timed state" if state is not None else "from the initial state"
    return {"value": n * substeps / el, "unit": "particle-substeps/s", "cores": threads, "kind": "port",
            "sample": f"{substeps} WCSPH substeps of the {n}-particle dam break {where}, FP32 C oracle, OpenMP "
                      f"{threads} threads, {el:.1f} s"}


def cpu_baseline_detail(scene):
    """The oracle's WCSPH substep rate single-threaded (the analogue of the
    reference's CPU path, one host thread per launch) and on all host cores
    (up to 16), at the headline size and at the 10,164-particle column."""
    from fluid_simulation_amd.adjoint import column_scene
    from fluid_simulation_amd.examples.dam_break.scene import DEFAULT_SPEC, ParticleScene

    col = column_scene(10000)
    s = col.spacing
    column = ParticleScene(DEFAULT_SPEC, col.positions, 1000.0 * s ** 3, 0.5 * s, s ** 3, s)
    threads = min(os.cpu_count() or 1, 16)
    out = {}
    for name, sc, many, one in (("dam_break_401856", scene, 20, 2), ("column_10164", column, 200, 20)):
        out[name] = {"threads_%d" % threads: cpu_baseline(sc, many)["value"],
                     "threads_1": cpu_baseline(sc, one, threads=1)["value"]}
    out["unit"] = "particle-substeps/s"
    return out


def copy_calibration():
    """Plain device copy of 256 MiB, three times: the reference kernel for
    scaling rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes."""
    import torch

    a = torch.ones(64 * 1024 * 1024, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_per_launch": 2 * a.numel() * 4}), flush=True)


def micro(sizes, reps):
    """Grid build, density pass and force pass at 4M / 16M points, uniform and
    clustered (per-kernel HIP-event timing; median over ``reps`` >= 20
    substeps after a warm-up substep), with the default launch variants and
    with the variants the autotuner picks on each input."""
    import torch

    from fluid_simulation_amd import _lib
    from fluid_simulation_amd.autotune import select as select_variant
    from fluid_simulation_amd.autotune import tune as tune_passes
    from fluid_simulation_amd.examples import microbench as mb
    from fluid_simulation_amd.solver.sph.solver_wcsph import SolverWCSPH

    lib = _lib.load()
    for n in sizes:
        for dist in ("uniform", "clustered"):
            pts = mb.uniform_points(n) if dist == "uniform" else mb.clustered_points(n)
            model, cfg = mb.wcsph_model(pts)
            s0, s1 = model.state(), model.state()
            grid = model.particle_grid
            grid.reserve(n)
            grid.build(s0.particle_q, mb.smoothing_length(n))
            counts = grid.query_counts(s0.particle_q, mb.smoothing_length(n), "cubic")
            # lists hold every particle's neighbours (clustered inputs reach
            # several times the uniform count), so the force pass replays
            cap = max(96, -(-int(counts.max().item()) // 32) * 32)
            solver = SolverWCSPH(model, cfg, neighbor_list_capacity=cap)
            solver.step(s0, s1, None, None, 1e-5)          # warm-up
            torch.cuda.synchronize()

            def sample():
                samples = {"build": [], "density": [], "force": []}
                for _ in range(max(20, reps)):
                    lib.fs_profile_reset()
                    lib.fs_profile_enable(1)
                    solver.step(s0, s1, None, None, 1e-5)
                    lib.fs_profile_enable(0)
                    prof = {}
                    for name, (_cnt, ms) in _lib.profile_report().items():
                        key = re.sub(r"<.*>", "", name.strip("()"))
                        prof[key] = prof.get(key, 0.0) + ms
                    samples["build"].append(sum(prof[k] for k in ("k_hash_count", "k_bucket_scan", "k_scatter",
                                                                   "k_fix_gather")))
                    samples["density"].append(sum(v for k, v in prof.items() if k.startswith("k_wcsph_density")))
                    samples["force"].append(prof["k_wcsph_force"])
                return tuple(statistics.median(samples[k]) for k in ("build", "density", "force"))

            build, dens0, force0 = sample()                  # the default launch variants
            tuned = tune_passes(lambda: solver.step(s0, s1, None, None, 1e-5), ("wcsph_density", "wcsph_force"))
            build, dens, force = sample()                    # the variants the tuner picked on this input
            for name in ("wcsph_density", "wcsph_force"):
                select_variant(name, 0)
            nbr = float(counts.float().mean().item()) - 1.0
            print(json.dumps({"micro": dist, "particles": n, "h": solver.smoothing_length, "mean_neighbors": nbr,
                              "max_neighbors": int(counts.max().item()) - 1, "list_capacity": cap,
                              "grid_build_ms": build, "density_pass_ms": dens, "force_pass_ms": force,
                              "density_pass_ms_default_variant": dens0, "force_pass_ms_default_variant": force0,
                              "autotune": {k: {**{f: v[f] for f in ("block", "list_depth", "unroll",
                                                                     "particles_per_thread", "lanes_per_particle",
                                                                     "coop")},
                                               "us_per_launch_by_variant": [round(t, 1) for t in v["us_per_launch"]]}
                                           for k, v in tuned.items()},
                              "density_GBs_alg": PASS_BYTES["k_wcsph_density"] * n / dens / 1e6,
                              "force_GBs_alg": PASS_BYTES["k_wcsph_force"] * n / force / 1e6}), flush=True)
            del solver, model, s0, s1
            torch.cuda.empty_cache()


def batched(num_scenes, steps, warmup, chunk=10):
    """configs: 64 dam-break scenes of log-uniform size in one batch.  Timed:
    (a) batched substeps with the per-scene CFL dt on the device, replayed from
    a HIP graph of ``chunk`` substeps; (b) the same eagerly; (c) 64 sequential
    plain ``SolverWCSPH.step`` calls per substep, one per scene, with each
    scene's dt taken from the batch; (d) those 64 plain steps captured in one
    graph; (e) the batched substeps of (a) without the CFL, the work of (c)
    and (d).  Every variant advances every scene by one substep per step."""
    import torch

    from fluid_simulation_amd.batched import BatchedWCSPH, make_batch_scenes
    from fluid_simulation_amd.examples.dam_break.simulation import DamBreakSimulation

    steps = max(chunk, (steps + chunk - 1) // chunk * chunk)
    scenes = make_batch_scenes(num_scenes, seed=3)
    total = sum(s.count for s in scenes)
    batch = BatchedWCSPH(scenes, device="cuda:0")
    batch.begin_frame()
    for _ in range(warmup):
        batch.substep()
    batch.check_status()
    tuned = batch.autotune(use_cache=False)
    print("# batched autotune: " + ", ".join(f"{k} -> block {v['block']} depth {v['list_depth']}"
                                             for k, v in tuned.items()), file=sys.stderr, flush=True)
    batch.remaining.fill_(1.0e9)          # no scene reaches the end of its time inside the timed region
    torch.cuda.synchronize()

    t0 = time.perf_counter()
    for _ in range(steps):
        batch.substep()
    torch.cuda.synchronize()
    eager = (time.perf_counter() - t0) / steps

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(side):
        with torch.cuda.graph(graph, stream=side):
            for _ in range(chunk):
                batch.substep()
    torch.cuda.current_stream().wait_stream(side)
    graph.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps // chunk):
        graph.replay()
    torch.cuda.synchronize()
    graphed = (time.perf_counter() - t0) / steps
    batch.check_status()
    dts = batch.dts()
    del graph

    # the same substeps without the on-device CFL (each scene keeps its last
    # dt), the work the sequential plain solves below do: one graph of
    # ``chunk`` advances
    graph = torch.cuda.CUDAGraph()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(graph, stream=side):
            for _ in range(chunk):
                batch.advance()
    torch.cuda.current_stream().wait_stream(side)
    graph.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps // chunk):
        graph.replay()
    torch.cuda.synchronize()
    graphed_nocfl = (time.perf_counter() - t0) / steps
    del graph

    sims = [DamBreakSimulation(s, "wcsph", device="cuda:0") for s in scenes]
    for _ in range(2):
        for sim, dt in zip(sims, dts):
            sim._substep(float(dt))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for sim, dt in zip(sims, dts):
            sim._substep(float(dt))
    torch.cuda.synchronize()
    seq_eager = (time.perf_counter() - t0) / steps

    # the 64 plain solves of one substep, captured once; buffers swap inside
    # the capture, so capture two substeps to return to the starting buffers
    graph = torch.cuda.CUDAGraph()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(graph, stream=side):
            for _ in range(2):
                for sim, dt in zip(sims, dts):
                    sim._substep(float(dt))
    torch.cuda.current_stream().wait_stream(side)
    graph.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps // 2):
        graph.replay()
    torch.cuda.synchronize()
    seq_graph = (time.perf_counter() - t0) / (2 * (steps // 2))
    for sim in sims:
        sim.check_finite()
    out = {"batched": {"scenes": num_scenes, "particles": total, "min_scene": min(s.count for s in scenes),
                       "max_scene": max(s.count for s in scenes), "buckets": batch.num_buckets, "steps": steps,
                       "ms_per_substep_graph": graphed * 1e3, "ms_per_substep_eager": eager * 1e3,
                       "ms_per_substep_sequential_eager": seq_eager * 1e3,
                       "ms_per_substep_sequential_graph": seq_graph * 1e3,
                       "particle_substeps_per_s_graph": total / graphed,
                       "particle_substeps_per_s_sequential_eager": total / seq_eager,
                       "speedup_vs_sequential_eager": seq_eager / graphed,
                       "speedup_vs_sequential_graph": seq_graph / graphed,
                       "ms_per_substep_graph_without_cfl": graphed_nocfl * 1e3,
                       "speedup_same_work_vs_sequential_graph": seq_graph / graphed_nocfl,
                       "autotune": {k: {"block": v["block"], "list_depth": v["list_depth"], "unroll": v["unroll"],
                                        "particles_per_thread": v["particles_per_thread"],
                                        "lanes_per_particle": v["lanes_per_particle"], "coop": v.get("coop", 0),
                                        "us_per_launch_by_variant": v["us_per_launch"]} for k, v in tuned.items()}}}
    del graph, sims, batch
    torch.cuda.empty_cache()
    return out


def adjoint(target=50000, substeps=200, checkpoint_every=10, fd=True, reps=3, v_scale=0.05, clamp=True):
    """Differentiable WCSPH: the 50,540-particle column (x in [-1.2,-0.6],
    y in [-0.28,0.28], z in [0,0.6], jitter 0.01 s), seeded random initial
    velocities (``v_scale`` m/s rms per axis), 200 substeps at dt = 0.25 h/c0,
    loss |centroid - c*|^2 + (P95(x) - x*)^2 with seeded targets.  Times the
    plain rollout and the forward + adjoint rollout (checkpoints every 10).

    With ``fd``, on the host cores: (a) the FP64 oracle's adjoint against FP64
    central differences (steps 1e-4 and 1e-5 relative) of d/dstiffness,
    d/dviscosity and a random direction of d/dv0; (b) the GPU gradients against
    the FP32 oracle's adjoint of the same FP32 rollout; (c) the GPU gradients
    against the FP64 differences directly.  ``clamp`` is the EOS clamp
    p = max(p, 0) (the dam break's setting); the check also counts the
    particle-substeps whose clamp branch (rho > rho0) differs between the
    FP32 and the FP64 rollouts."""
    import numpy as np
    import torch

    from fluid_simulation_amd.adjoint import DifferentiableWCSPH, centroid_p95_loss, column_scene, dt_for

    B0, NU0 = 250000.0, 1.0e-6
    scene = column_scene(target)
    rng = np.random.default_rng(4)
    v0n = (rng.normal(size=(scene.count, 3)) * v_scale).astype(np.float32)
    x0n = scene.positions
    ctarget = x0n.mean(axis=0).astype(np.float64) + rng.normal(scale=0.02, size=3)
    xtarget = float(np.quantile(x0n[:, 0], 0.95)) + float(rng.normal(scale=0.02))
    dt = dt_for(scene.spacing)
    sim = DifferentiableWCSPH(x0n, scene.mass, scene.radius, scene.spacing, dt=dt, steps=substeps,
                              checkpoint_every=checkpoint_every, clamp_negative_pressure=clamp)
    x0 = torch.from_numpy(x0n).cuda()
    v0d = torch.from_numpy(v0n).cuda()
    ct, xt = torch.from_numpy(ctarget).cuda(), torch.tensor(xtarget, dtype=torch.float64).cuda()

    def run(grad):
        v0 = v0d.clone().requires_grad_(grad)
        B = torch.tensor(B0, dtype=torch.float64, device="cuda", requires_grad=grad)
        nu = torch.tensor(NU0, dtype=torch.float64, device="cuda", requires_grad=grad)
        xT, _ = sim.rollout(x0, v0, B, nu)
        loss = centroid_p95_loss(xT.double(), ct, xt)
        if grad:
            loss.backward()
            return loss.item(), float(B.grad), float(nu.grad), v0.grad.double().cpu().numpy()
        return loss.item(), None, None, None

    run(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        with torch.no_grad():
            run(False)
    torch.cuda.synchronize()
    t_fwd = (time.perf_counter() - t0) / reps
    run(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        loss, gB, gnu, gv = run(True)
    torch.cuda.synchronize()
    t_both = (time.perf_counter() - t0) / reps
    out = {"particles": scene.count, "substeps": substeps, "checkpoint_every": checkpoint_every, "dt": dt,
           "clamp_negative_pressure": clamp,
           "stiffness": B0, "viscosity": NU0, "v0_rms": v_scale, "loss": loss, "dL_dstiffness": gB,
           "dL_dviscosity": gnu, "ms_forward": t_fwd * 1e3, "ms_forward_backward": t_both * 1e3,
           "backward_over_forward": (t_both - t_fwd) / t_fwd,
           "particle_substeps_per_s_forward": scene.count * substeps / t_fwd,
           "particle_substeps_per_s_forward_backward": scene.count * substeps / t_both}
    del sim
    torch.cuda.empty_cache()
    if fd:
        from oracle import oracle

        threads = min(16, os.cpu_count() or 1)
        oracle.set_threads(threads)
        cfg = oracle.WcsphConfig(particle_radius=scene.radius, particle_length=scene.spacing,
                                 smoothing_length_coff=2.0, bound_width=1.2, bound_height=0.28, bound_length=0.6,
                                 boundary_damping=0.0, clamp_negative_pressure=clamp)
        dv = np.random.default_rng(12).normal(size=v0n.shape)
        rel = lambda a, b: abs(a - b) / max(abs(b), 1e-300)
        t0 = time.perf_counter()

        def model(prec, B=B0, nu=NU0):
            return oracle.uniform_model(scene.count, np.float32(scene.mass), stiffness=B, exponent=7.0,
                                        viscosity=nu, dtype=prec)

        def loss_and_grad(xT):
            n = xT.shape[0]
            c = xT.mean(axis=0)
            g = np.tile(2.0 * (c - ctarget) / n, (n, 1))
            q = 0.95 * (n - 1)
            kq = int(np.floor(q))
            fr = q - kq
            order = np.argsort(xT[:, 0], kind="stable")
            p95 = np.quantile(xT[:, 0], 0.95)
            g[order[kq], 0] += 2.0 * (p95 - xtarget) * (1.0 - fr)
            g[order[kq + 1], 0] += 2.0 * (p95 - xtarget) * fr
            return float(((c - ctarget) ** 2).sum() + (p95 - xtarget) ** 2), g

        rhos = {}
        kept = {}

        def oracle_adjoint(prec):
            npd = np.float64 if prec == "f64" else np.float32
            rhos[prec] = []
            states = oracle.wcsph_rollout(model(npd), x0n.astype(npd), v0n.astype(npd), cfg, dt, substeps, prec,
                                          densities=rhos[prec])
            _, g = loss_and_grad(states[-1][0].astype(np.float64))
            _, vb, gs, gvv = oracle.wcsph_rollout_adjoint(model(npd), states, cfg, dt, g, np.zeros_like(g), prec)
            if prec == "f32":
                kept["states"], kept["seed"] = states, g
            return float(gs.sum()), float(gvv.sum()), np.asarray(vb, np.float64)

        a64 = oracle_adjoint("f64")
        print(f"# adjoint check: fp64 oracle adjoint done ({time.perf_counter() - t0:.0f} s)", file=sys.stderr,
              flush=True)

        def loss64(B=B0, nu=NU0, v=v0n):
            xT, _ = oracle.wcsph_final_state(model(np.float64, B, nu), x0n.astype(np.float64),
                                             np.asarray(v, np.float64), cfg, dt, substeps, "f64")
            return loss_and_grad(xT)[0]

        fds = {}
        rms = float(np.sqrt((v0n.astype(np.float64) ** 2).mean()))
        for e in (1e-4, 1e-5):
            fds[e] = ((loss64(B=B0 * (1 + e)) - loss64(B=B0 * (1 - e))) / (2 * e * B0),
                      (loss64(nu=NU0 * (1 + e)) - loss64(nu=NU0 * (1 - e))) / (2 * e * NU0),
                      (loss64(v=v0n + e * rms * dv) - loss64(v=v0n - e * rms * dv)) / (2 * e * rms))
            print(f"# adjoint check: fp64 differences, step {e:g} done ({time.perf_counter() - t0:.0f} s)",
                  file=sys.stderr, flush=True)
        a32 = oracle_adjoint("f32")
        print(f"# adjoint check: fp32 oracle adjoint done ({time.perf_counter() - t0:.0f} s)", file=sys.stderr,
              flush=True)
        # the FP64 adjoint linearised along the FP32 trajectory (its stored
        # states, each substep recomputed in FP64): the derivative of the FP32
        # trajectory's branch pattern without FP32 arithmetic in the adjoint
        st64 = [(x.astype(np.float64), v.astype(np.float64)) for x, v in kept.pop("states")]
        _, vbm, gsm, gvm = oracle.wcsph_rollout_adjoint(model(np.float64), st64, cfg, dt, kept["seed"],
                                                        np.zeros_like(kept["seed"]), "f64")
        del st64
        mixed = (float(gsm.sum()), float(gvm.sum()), float((np.asarray(vbm, np.float64) * dv).sum()))
        print(f"# adjoint check: fp64 adjoint along the fp32 trajectory done ({time.perf_counter() - t0:.0f} s)",
              file=sys.stderr, flush=True)
        names = ("stiffness", "viscosity", "v0_direction")
        # particle-substeps on different sides of rho0 in the two precisions:
        # there p = max(B((rho/rho0)^g - 1), 0) takes different branches
        flips = sum(int(np.count_nonzero((r32 > np.float32(1000.0)) != (r64 > 1000.0)))
                    for r32, r64 in zip(rhos["f32"], rhos["f64"]))
        compressed = sum(int(np.count_nonzero(r64 > 1000.0)) for r64 in rhos["f64"])
        ad64 = (a64[0], a64[1], float((a64[2] * dv).sum()))
        gpu = (gB, gnu, float((gv * dv).sum()))
        check = {"step_relative": list(fds), "host_threads": threads,
                 "fp64_oracle_adjoint_vs_fp64_fd": {n: [ad64[k]] + [fds[e][k] for e in fds] +
                                                    [min(rel(ad64[k], fds[e][k]) for e in fds)]
                                                    for k, n in enumerate(names)},
                 "gpu_vs_fp32_oracle_adjoint": {"stiffness": rel(gB, a32[0]), "viscosity": rel(gnu, a32[1]),
                                                "v0_max_abs_over_max": float(np.abs(gv - a32[2]).max() /
                                                                             np.abs(a32[2]).max())},
                 "gpu_vs_fp64_fd": {n: [gpu[k]] + [fds[e][k] for e in fds] +
                                    [min(rel(gpu[k], fds[e][k]) for e in fds)] for k, n in enumerate(names)},
                 "fp64_adjoint_along_fp32_trajectory_vs_gpu": {n: [mixed[k], gpu[k], rel(gpu[k], mixed[k])]
                                                               for k, n in enumerate(names)},
                 "clamp_branch_disagreements_fp32_vs_fp64": flips,
                 "particle_substeps_above_rest_density_fp64": compressed,
                 "format": "[gradient, fd step 1e-4, fd step 1e-5, smaller relative error]",
                 "seconds": time.perf_counter() - t0}
        out["fd_check"] = check
    return {"adjoint": out}


def frames(particles=400000, count=3, chunk=16):
    """Whole display frames of the WCSPH dam break: the host CFL driver (one
    readback of the maxima per substep, simulation.py:155-190) against the
    device-resident driver (CFL on the device, graph-replayed chunks of
    substeps, one completion check per chunk)."""
    import torch

    from fluid_simulation_amd.examples.dam_break.device_driver import DeviceResidentDamBreak
    from fluid_simulation_amd.examples.dam_break.scene import make_scene
    from fluid_simulation_amd.examples.dam_break.simulation import DamBreakSimulation

    scene = make_scene(particles)
    host = DamBreakSimulation(scene, "wcsph", device="cuda:0")
    dev = DeviceResidentDamBreak(scene, "wcsph", device="cuda:0", chunk=chunk)
    host.step()
    dev.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    host_sub = 0
    for _ in range(count):
        host.step()
        host_sub += host.last_substeps
    torch.cuda.synchronize()
    t_host = (time.perf_counter() - t0) / count
    t0 = time.perf_counter()
    dev_sub = 0
    for _ in range(count):
        dev_sub += dev.step()
    torch.cuda.synchronize()
    t_dev = (time.perf_counter() - t0) / count
    out = {"particles": scene.count, "frames": count, "chunk": chunk,
           "ms_per_frame_host_driver": t_host * 1e3, "ms_per_frame_device_resident": t_dev * 1e3,
           "substeps_per_frame_host": host_sub / count,
           "substeps_per_frame_device_incl_idle": dev_sub / count,
           "speedup": t_host / t_dev}
    del host, dev
    torch.cuda.empty_cache()
    return {"frames": out}


# separation: per particle-iteration, grid 28 B + pass (perm 4, x 12, flags 4, x written 12) 32 B
SEPARATION_BYTES = 60


def separation(particles=400000, reps=10, cpu=True):
    """APIC particle separation over the dam-break lattice jittered by
    +-0.25 spacing (pairs closer than the rest distance everywhere), the APIC
    dam-break config (tank bounds, clamp_eps 1e-3, separation defaults: rest =
    2 x mean radius, 2 iterations).  One apply = 2 x (grid build + fused
    accumulate/apply pass).  Timed on positions reset before each apply."""
    import numpy as np
    import torch

    from fluid_simulation_amd import _lib
    from fluid_simulation_amd.examples.dam_break.scene import make_scene
    from fluid_simulation_amd.newton_compat import Axis, ModelBuilder
    from fluid_simulation_amd.separation import ParticleSeparation, SeparationConfig

    scene = make_scene(particles)
    rng = np.random.default_rng(11)
    pts = (scene.positions + rng.uniform(-0.25, 0.25, scene.positions.shape) * scene.spacing).astype(np.float32)
    n = pts.shape[0]
    b = ModelBuilder(up_axis=Axis.Z)
    b.add_particles(pos=pts, vel=np.zeros_like(pts), mass=[scene.mass] * n, radius=[scene.radius] * n)
    model = b.finalize(device="cuda:0")
    cfg = SeparationConfig(bound_lo=scene.spec.bound_lo, bound_hi=scene.spec.bound_hi, clamp_eps=1.0e-3)
    sep = ParticleSeparation(model, cfg)
    state = model.state()
    x0 = state.particle_q.tensor.clone()
    for _ in range(2):
        state.particle_q.tensor.copy_(x0)
        sep.apply(state)
    torch.cuda.synchronize()
    lib = _lib.load()
    lib.fs_profile_reset()
    total = 0.0
    for _ in range(reps):
        state.particle_q.tensor.copy_(x0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        lib.fs_profile_enable(1)
        sep.apply(state)
        lib.fs_profile_enable(0)
        total += time.perf_counter() - t0
    iters = sep.params().iterations
    prof = {}
    for name, (cnt, ms) in _lib.profile_report().items():
        key = re.sub(r"<.*>", "", name.strip("()"))
        c0, m0 = prof.get(key, (0, 0.0))
        prof[key] = (c0 + cnt, m0 + ms)
    build_ms = sum(prof[k][1] for k in ("k_hash_count", "k_bucket_scan", "k_scatter", "k_fix_gather")) / reps
    pass_cnt, pass_ms = prof["k_separation"]
    pass_us = pass_ms / pass_cnt * 1e3
    gpu_ms = build_ms + pass_ms / reps
    out = {"particles": n, "iterations": iters, "grid_dims": list(sep.dims), "rest_dist": sep.rest_dist,
           "ms_per_apply": total / reps * 1e3, "kernel_ms_per_apply": gpu_ms,
           "particle_iterations_per_s": n * iters / (gpu_ms / 1e3),
           "separation_pass_us": pass_us, "grid_build_us_per_iteration": build_ms / iters * 1e3,
           "pass_GBs_alg": 32 * n / (pass_us / 1e6) / 1e9,
           "GBs_alg": SEPARATION_BYTES * n * iters / (gpu_ms / 1e3) / 1e9}
    if cpu:
        from oracle import oracle as O
        O.set_threads(16)
        p = O.separation_params(cfg, model.particle_radius.numpy())
        flags = model.particle_flags.numpy()
        t0 = time.perf_counter()
        O.particle_separation(pts, flags, p, precision="f32")
        t_cpu = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n * iters / t_cpu, "unit": "particle-iterations/s", "cores": 16,
                               "kind": "port", "sample": "one apply (2 iterations) over all particles"}
    del sep, model, state
    torch.cuda.empty_cache()
    return {"separation": out}


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.micro:
        micro([int(x) for x in args.micro.split(",")], args.steps)
        return
    if args.frames:
        print(json.dumps(frames(count=args.frames)), flush=True)
        return
    if args.adjoint:
        print(json.dumps(adjoint(fd=not args.no_fd, clamp=not args.adjoint_no_clamp)), flush=True)
        return
    if args.batched:
        print(json.dumps(batched(args.batched, args.steps, args.warmup)), flush=True)
        return
    if args.separation:
        print(json.dumps(separation(reps=max(args.steps, 5))), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the path does not shard "
              "across GPUs here)", file=sys.stderr)
        sys.exit(2)
    import torch

    from fluid_simulation_amd.examples.dam_break.scene import make_scene

    scene = make_scene(args.particles)
    results = {}
    prof_wcsph = None
    wcsph_state = None
    for method in [s for s in args.solvers.split(",") if s]:
        print(f"# bench {method}: n={scene.count} steps={args.steps} warmup={args.warmup} "
              f"after {args.settle_frames} frames", file=sys.stderr, flush=True)
        res, prof, state = bench_solver(method, scene, args.steps, args.warmup, tune=not args.no_tune,
                                        settle_frames=args.settle_frames)
        results[method] = res
        if method == "wcsph":
            prof_wcsph = prof
            wcsph_state = state
        print(f"# {method}: {res['particle_substeps_per_s'] / 1e6:.1f} M particle-substeps/s, "
              f"{res['ms_per_substep']:.3f} ms/substep", file=sys.stderr, flush=True)
    head = "wcsph" if "wcsph" in results else next(iter(results))
    headline = results[head]
    prof = prof_wcsph if prof_wcsph is not None else prof
    dom = max((k for k in prof if k in PASS_BYTES), key=lambda k: prof[k][1])
    launches, total_ms = prof[dom]
    avg_s = total_ms / launches / 1e3
    achieved = PASS_BYTES[dom] * scene.count / avg_s / 1e9
    traffic = None
    tuned = headline.get("autotune", {})
    tkey = {"k_wcsph_density": "wcsph_density", "k_wcsph_density_coop": "wcsph_density",
            "k_wcsph_force": "wcsph_force"}.get(dom)
    if tkey in tuned:
        traffic = pmc_traffic(variant_key(dom, tuned[tkey]))
    roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "avg_launch_us": avg_s * 1e6,
                "bytes_per_launch": PASS_BYTES[dom] * scene.count,
                "traffic_source": "profiles/wcsph_dam_break_401856_pmc_traffic.md (HBM bytes per launch of the "
                                  "selected launch variant: FETCH_SIZE x2.0 + WRITE_SIZE x1.0)"}
    if tkey in tuned:
        hits = pmc_cache(variant_key(dom, tuned[tkey]))
        if hits:
            roofline.update(l1_hit=hits[0], l2_hit=hits[1],
                            cache_source="profiles/wcsph_dam_break_401856_pmc_cache.md (the selected launch "
                                         "variant's vector L1 and L2 hit rates)")
    valu = None
    if tkey in tuned:
        instr = pmc_valu_instructions(variant_key(dom, tuned[tkey]))
        if instr:
            rate = instr / avg_s
            valu = {"bound": "valu_issue", "kernel": dom, "instructions_per_launch": instr, "achieved": rate,
                    "peak": VALU_ISSUE_PEAK, "unit": "wave64 VALU instructions/s", "frac": rate / VALU_ISSUE_PEAK,
                    "source": "profiles/wcsph_dam_break_401856_pmc_sq.md (SQ_INSTS_VALU per launch of the selected "
                              "variant) over the live HIP-event launch duration"}
    extras = {}
    if not args.no_extras and args.particles == 400000:
        print("# batched form: 64 scenes", file=sys.stderr, flush=True)
        extras.update(batched(64, 10, 2))
        print("# differentiable form: 50,540 particles, 200 substeps", file=sys.stderr, flush=True)
        extras.update(adjoint(fd=False, reps=2))
        print("# whole frames: host CFL driver vs device-resident", file=sys.stderr, flush=True)
        extras.update(frames(count=2))
        print("# APIC particle separation: 401,856 particles, 2 iterations", file=sys.stderr, flush=True)
        extras.update(separation(cpu=not args.no_cpu_baseline))
    cpu = cpu_detail = None
    if not args.no_cpu_baseline:
        print("# cpu baseline (oracle)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(scene, args.cpu_substeps, state=wcsph_state)
        if args.particles == 400000:
            cpu_detail = cpu_baseline_detail(scene)
    value = headline["particle_substeps_per_s"]
    line = {
        "metric": "particle-substeps/s @402k dam-break, % HBM peak; batched & adjoint vs plain",
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": headline["ms_per_substep"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": value / PUBLISHED_WCSPH if head == "wcsph" else None,
        "dtype": "f32",
        "data": "synthetic: make_scene(400000) dam-break lattice (401,856 particles), zero initial velocity, "
                f"advanced {args.settle_frames} frames by the solver before the timed substeps",
        "config": {"workload": f"{head} dam break, {scene.count} particles, one solver substep per step, "
                               f"timed at the state {args.settle_frames} frames "
                               f"({args.settle_frames / 60:.2f} s) into the run",
                   "particles": scene.count, "solver": head, "frames_before_timing": args.settle_frames,
                   "dt": headline["dt"]},
        "hbm_frac_substep": headline["hbm_frac"],
        "roofline": roofline,
        "valu_issue": valu,
        "cpu_baseline": cpu,
        "cpu_baseline_detail": cpu_detail,
        "solvers": results,
        **extras,
        "device": torch.cuda.get_device_name(0),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

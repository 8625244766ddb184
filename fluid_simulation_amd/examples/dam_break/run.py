"""Headless dam-break run that writes the reference's run records.

Counterpart of the reference's ``examples/dam_break/run.py`` (:22-234) for the
SPH solvers.  The command line keeps the reference's options that apply to
them (:43-94).  The frame loop is the same (:183-210): the finite check after
every frame, and ``DamBreakSimulation.metrics`` at the initial state, every
``--metrics-every`` frames, at report frames and at the final state.  The
records are the same (:211-227):
* ``metrics.csv``: one row per metrics call, header = the union of the row
  keys in first-seen order;
* ``run.json``: scene report (with the SHA-256 of the initial positions),
  simulation report, completed frames, simulated and elapsed seconds,
  diagnostics, environment, surface, final metrics.

The viewer and the surface extraction (Newton's ``ParticleSurface``) are
display consumers of the positions and are not part of this package:
``--viewer gl`` is refused, and ``--usd`` needs ``--no-surface`` (the
reference then writes the particles alone, surface.py:82-86).  ``--usd PATH``
(.usd or .usda) writes the particles, time-sampled every ``--usd-every``
frames plus the final frame, and the tank and height guides as a USD text
layer (``usd.py``); its record goes into run.json's ``surface`` entry.
``--device-resident`` runs frames with the device-resident driver (no host
synchronisation per substep).

    python -m fluid_simulation_amd.examples.dam_break.run --solver wcsph --frames 60
"""
from __future__ import annotations

import argparse
import csv
import json
import sys
import time
from pathlib import Path

from .scene import DEFAULT_SPEC, make_scene

METHODS = ("wcsph", "dfsph", "pbf")


def positive_int(value):
    value = int(value)
    if value < 1:
        raise argparse.ArgumentTypeError("Must be a positive integer")
    return value


def positive_float(value):
    value = float(value)
    if not 0.0 < value < float("inf"):
        raise argparse.ArgumentTypeError("Must be a finite positive number")
    return value


def nonnegative_float(value):
    value = float(value)
    if not 0.0 <= value < float("inf"):
        raise argparse.ArgumentTypeError("Must be a finite non-negative number")
    return value


def parse_args(method=None, argv=None):
    parser = argparse.ArgumentParser(description="Dam-break run with metrics.csv / run.json records")
    if method is None:
        parser.add_argument("--solver", choices=METHODS, default="wcsph")
    else:
        parser.set_defaults(solver=method)
    parser.add_argument("--device", default="cuda:0")
    parser.add_argument("--particles", type=positive_int, default=DEFAULT_SPEC.target_count,
                        help="Target count; the lattice determines the actual count")
    parser.add_argument("--frames", type=positive_int, default=300)
    parser.add_argument("--substeps", type=positive_int, default=4,
                        help="Base substeps per 1/60 s frame; WCSPH and DFSPH may use more for CFL")
    parser.add_argument("--viewer", choices=["gl", "none"], default="none")
    parser.add_argument("--no-viewer", dest="viewer", action="store_const", const="none")
    parser.add_argument("--no-surface", action="store_true")
    parser.add_argument("--show-particles", action="store_true")
    parser.add_argument("--usd", help="Particle animation as a USD text layer (.usd or .usda); needs --no-surface")
    parser.add_argument("--usd-every", type=positive_int, default=1)
    parser.add_argument("--output-dir", type=Path, default=None, help="Metrics and run metadata directory")
    parser.add_argument("--wcsph-stiffness", type=positive_float, default=None,
                        help="WCSPH Tait pressure constant B in Pa (default: 250000)")
    parser.add_argument("--wcsph-viscosity", type=nonnegative_float, default=None,
                        help="WCSPH kinematic viscosity in m^2/s (default: 1e-6)")
    parser.add_argument("--wcsph-smoothing-length-coff", type=positive_float, default=None,
                        help="WCSPH kernel support radius / particle spacing (default: 2)")
    parser.add_argument("--wcsph-allow-negative-pressure", action="store_true",
                        help="Disable the pressure clamp for WCSPH comparisons")
    parser.add_argument("--report-every", type=positive_int, default=10)
    parser.add_argument("--metrics-every", type=positive_int, default=10,
                        help="Full statistics interval; finite-state checks still run every frame")
    parser.add_argument("--device-resident", action="store_true",
                        help="device-side CFL and graph-replayed substeps")
    args = parser.parse_args(argv)
    if args.viewer != "none":
        parser.error("the viewer is not part of this package; use --no-viewer")
    if args.usd:
        if Path(args.usd).suffix.lower() not in (".usd", ".usda"):
            parser.error("--usd writes a USD text layer: use a .usd or .usda path (.usdc is binary crate)")
        if not args.no_surface:
            parser.error("surface meshing (Newton's ParticleSurface) is not part of this package: add --no-surface "
                         "to write the particles")
    if args.solver != "wcsph" and (
        args.wcsph_stiffness is not None or args.wcsph_viscosity is not None
        or args.wcsph_smoothing_length_coff is not None or args.wcsph_allow_negative_pressure
    ):
        parser.error("--wcsph-* options apply to WCSPH")
    args.output_dir = args.output_dir or Path.cwd() / "output" / args.solver
    return args


def write_records(output_dir: Path, rows: list[dict], report: dict) -> None:
    """metrics.csv and run.json as the reference writes them (run.py:215-227)."""
    output_dir.mkdir(parents=True, exist_ok=True)
    fieldnames = list(dict.fromkeys(key for row in rows for key in row))
    with (output_dir / "metrics.csv").open("w", newline="") as stream:
        writer = csv.DictWriter(stream, fieldnames=fieldnames)
        writer.writeheader()
        writer.writerows(rows)
    (output_dir / "run.json").write_text(json.dumps(report, indent=2) + "\n")


def environment(device) -> dict:
    import torch

    from ... import _lib

    return dict(python=sys.version.split()[0], torch=torch.__version__, hip=torch.version.hip,
                fluidsim=int(_lib.load().fs_version()), device=str(device),
                device_name=torch.cuda.get_device_name(device))


def main(method=None, argv=None):
    from .device_driver import DeviceResidentDamBreak
    from .simulation import DamBreakSimulation

    args = parse_args(method, argv)
    scene = make_scene(args.particles)
    print(f"{args.solver.upper()} dam_break: {scene.count} particles, "
          f"radius={scene.radius:.6g} m, spacing={scene.spacing:.6g} m", flush=True)
    print(f"tank={scene.spec.bound_lo} .. {scene.spec.bound_hi}; "
          f"{args.frames} frames at {scene.spec.fps:g} fps", flush=True)
    print(f"ideal flat height={scene.ideal_height:.6f} m", flush=True)
    options = {key: getattr(args, key) for key in ("wcsph_stiffness", "wcsph_viscosity", "wcsph_smoothing_length_coff")
               if getattr(args, key) is not None}
    sim_kwargs = dict(wcsph_clamp_negative_pressure=not args.wcsph_allow_negative_pressure, **options)
    if args.device_resident:
        driver = DeviceResidentDamBreak(scene, args.solver, device=args.device, substeps=args.substeps,
                                        **sim_kwargs)
        sim = driver.sim
        advance = driver.step
    else:
        sim = DamBreakSimulation(scene, args.solver, args.device, args.substeps, **sim_kwargs)
        advance = sim.step
    if args.solver == "wcsph":
        print(f"WCSPH B={sim.material['stiffness']:g} Pa, viscosity={sim.material['viscosity']:g} m^2/s, "
              f"h={sim.solver.smoothing_length:.6g} m, "
              f"non-negative pressure={sim.solver.config.clamp_negative_pressure}", flush=True)
    usd = None
    if args.usd:
        from .usd import ParticleUsdWriter

        usd = ParticleUsdWriter(args.usd, scene, args.solver)

    def record():
        if usd is not None:
            usd.record(sim.frame, sim.state_0.particle_q.numpy())

    rows = [sim.metrics()]
    started = time.perf_counter()
    try:
        record()
        while sim.frame < args.frames:
            advance()
            sim.check_finite()
            report_now = sim.frame % args.report_every == 0 or sim.frame == args.frames
            if sim.frame % args.metrics_every == 0 or report_now:
                metrics = sim.metrics()
                rows.append(metrics)
            if sim.frame % args.usd_every == 0:
                record()
            if report_now:
                print(f"frame {sim.frame}/{args.frames}, t={sim.sim_time:.4f} s, "
                      f"substeps={sim.last_substeps}, vmax={metrics['speed_max']:.4g}, finite=True", flush=True)
        record()          # the final state is always in the layer
        if rows[-1]["frame"] != sim.frame:
            rows.append(sim.metrics())
    finally:
        if usd is not None:
            usd.close()
    elapsed = time.perf_counter() - started
    simulation = sim.report()
    if args.device_resident:
        simulation["solver_execution"] += "; device-resident frames (graph-replayed substeps)"
    report = dict(scene=scene.report(), simulation=simulation, completed_frames=sim.frame,
                  simulation_seconds=sim.sim_time, elapsed_seconds=elapsed,
                  diagnostics=dict(metrics_every=args.metrics_every, finite_check_every=1),
                  environment=environment(sim.model.device),
                  surface=usd.report() if usd is not None else dict(enabled=False), final_metrics=rows[-1])
    write_records(args.output_dir, rows, report)
    print(f"Finished in {elapsed:.2f} s. Metrics: {args.output_dir}", flush=True)
    if usd is not None:
        print(f"USD: {usd.path} ({usd.samples} samples)", flush=True)
    return report


if __name__ == "__main__":
    main()

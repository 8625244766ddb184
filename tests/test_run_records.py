"""Run records of the headless dam-break runner (host side).

The reference's run.py writes metrics.csv (header = union of row keys in
first-seen order) and run.json (examples/dam_break/run.py:215-227); its CLI
refuses --wcsph-* options for other solvers (:88-92).
"""
import csv
import json

import pytest

from fluid_simulation_amd.examples.dam_break import run


def test_cli_defaults_and_validation(tmp_path):
    args = run.parse_args(argv=["--output-dir", str(tmp_path)])
    assert (args.solver, args.frames, args.substeps, args.particles) == ("wcsph", 300, 4, 400000)
    assert (args.report_every, args.metrics_every) == (10, 10)
    assert run.parse_args("pbf", argv=["--no-viewer", "--no-surface"]).solver == "pbf"
    a = run.parse_args(argv=["--usd", "x.usda", "--no-surface", "--usd-every", "5", "--show-particles"])
    assert (a.usd, a.usd_every) == ("x.usda", 5)
    for bad in (["--solver", "pbf", "--wcsph-stiffness", "1e5"], ["--solver", "dfsph", "--wcsph-allow-negative-pressure"],
                ["--viewer", "gl"], ["--usd", "x.usd"], ["--usd", "x.usdc", "--no-surface"],
                ["--frames", "0"], ["--wcsph-viscosity", "-1"], ["--solver", "apic"]):
        with pytest.raises(SystemExit):
            run.parse_args(argv=bad)


def test_records_have_the_reference_layout(tmp_path):
    rows = [dict(frame=0, simulation_time=0.0, substeps=0, total_substeps=0, finite=True, speed_max=0.0),
            dict(frame=10, simulation_time=1 / 6, substeps=5, total_substeps=50, finite=True, speed_max=1.5,
                 density_mean=1000.0, density_p99=1010.0)]
    report = dict(scene={"initial_position_sha256": "ab"}, simulation={"method": "wcsph"}, completed_frames=10,
                  simulation_seconds=1 / 6, elapsed_seconds=0.5, diagnostics={}, environment={},
                  surface=dict(enabled=False), final_metrics=rows[-1])
    run.write_records(tmp_path / "out", rows, report)
    with (tmp_path / "out" / "metrics.csv").open() as f:
        reader = csv.DictReader(f)
        assert reader.fieldnames == ["frame", "simulation_time", "substeps", "total_substeps", "finite", "speed_max",
                                     "density_mean", "density_p99"]
        got = list(reader)
    assert [r["frame"] for r in got] == ["0", "10"] and got[0]["density_mean"] == ""
    loaded = json.loads((tmp_path / "out" / "run.json").read_text())
    assert list(loaded) == ["scene", "simulation", "completed_frames", "simulation_seconds", "elapsed_seconds",
                            "diagnostics", "environment", "surface", "final_metrics"]

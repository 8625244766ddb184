"""The headless dam-break runner on the GPU: a few frames of each solver, the
records it writes, and the device-resident path (same rows and keys)."""
import csv
import hashlib
import json

import numpy as np
import pytest

from fluid_simulation_amd.examples.dam_break import run
from fluid_simulation_amd.examples.dam_break.scene import make_scene

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method,extra", [("wcsph", []), ("dfsph", []), ("pbf", []), ("wcsph", ["--device-resident"]),
                                          ("pbf", ["--device-resident"]),
                                          ("dfsph", ["--device-resident"])])
def test_run_writes_records(tmp_path, method, extra):
    out = tmp_path / method
    report = run.main(method, ["--particles", "2000", "--frames", "4", "--metrics-every", "2", "--report-every", "4",
                               "--output-dir", str(out), *extra])
    with (out / "metrics.csv").open() as f:
        rows = list(csv.DictReader(f))
    assert [int(r["frame"]) for r in rows] == [0, 2, 4]
    assert all(r["finite"] == "True" for r in rows)
    substeps = [int(r["substeps"]) for r in rows[1:]]
    assert all(s >= 4 for s in substeps)
    if method == "pbf" or not extra:
        assert int(rows[-1]["total_substeps"]) >= 2 * substeps[-1]
    assert float(rows[-1]["density_mean"]) > 0.0
    loaded = json.loads((out / "run.json").read_text())
    assert loaded["completed_frames"] == 4
    assert abs(loaded["simulation_seconds"] - 4 / 60) < 1e-12
    scene = make_scene(2000)
    assert loaded["scene"]["initial_position_sha256"] == hashlib.sha256(scene.positions.tobytes()).hexdigest()
    assert loaded["scene"]["actual_count"] == scene.count
    assert loaded["simulation"]["method"] == method
    assert loaded["final_metrics"]["frame"] == 4
    assert report["completed_frames"] == 4
    assert np.isfinite(float(rows[-1]["height_mean"]))


def test_run_writes_particle_usd(tmp_path):
    """--usd with --no-surface: the particles every --usd-every frames and at
    the final frame, in the layer's time samples (surface.py:156-160)."""
    from fluid_simulation_amd.examples.dam_break.usd import read_points_samples

    out = tmp_path / "wcsph"
    usd = tmp_path / "anim" / "dam.usd"
    report = run.main("wcsph", ["--particles", "2000", "--frames", "5", "--usd", str(usd), "--no-surface",
                                "--usd-every", "2", "--output-dir", str(out)])
    samples = read_points_samples(usd)
    assert sorted(samples) == [0, 2, 4, 5]
    np.testing.assert_array_equal(samples[0], make_scene(2000).positions)
    assert all(np.isfinite(x).all() and x.shape == (make_scene(2000).count, 3) for x in samples.values())
    assert not np.array_equal(samples[5], samples[0])
    surf = json.loads((out / "run.json").read_text())["surface"]
    assert surf["usd_samples"] == 4 and surf["usd_path"] == str(usd.resolve())
    assert report["surface"]["frame_samples"][-1]["frame"] == 5

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
                                          ("pbf", ["--device-resident"])])
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
    if method != "wcsph" or not extra:
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

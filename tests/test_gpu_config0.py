"""The CPU-scale WCSPH case over 60 display frames: the GPU driver against
the FP32 oracle's trajectory (tests/golden/config0_column_wcsph_60frames.npz,
written by tests/golden/make_config0.py): the 10,164-particle jittered water
column in the dam-break tank, dam-break WCSPH material, the driver's CFL
policy with a base of 4 substeps.

Each side chooses its own substeps from its own state.  The kernels evaluate
the oracle's float32 expressions in the same order and both drivers evaluate
the CFL policy in float64 from float32 maxima, so the whole trajectory is
bit-identical: the same 3,563 dt values and the same final state.  The
per-frame statistics the survey prescribes for long horizons (front = max x,
p99 of rho, mean height, mean x, substeps per frame) are compared as well,
and the first differing substep is reported if there is one.
"""
import os

import numpy as np
import pytest

from fluid_simulation_amd.examples.dam_break.scene import DEFAULT_SPEC, ParticleScene
from fluid_simulation_amd.examples.dam_break.simulation import DamBreakSimulation

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "config0_column_wcsph_60frames.npz")


def test_column_60_frames_match_the_oracle_trajectory():
    g = np.load(FIXTURE)
    s = float(g["spacing"])
    scene = ParticleScene(DEFAULT_SPEC, g["positions"], float(g["mass"]), float(g["radius"]), s ** 3, s)
    sim = DamBreakSimulation(scene, "wcsph", device="cuda:0")
    assert sim.solver.smoothing_length == pytest.approx(2.0 * s)
    dts = []
    advance = sim._substep

    def record(dt):
        dts.append(dt)
        advance(dt)

    sim._substep = record
    per_frame = []
    for _ in range(g["per_frame"].shape[0]):
        sim.step()
        x = sim.state_0.particle_q.numpy()
        rho = sim.state_0.sph.rho.numpy()
        per_frame.append([sim.last_substeps, float(x[:, 0].max()), float(np.percentile(rho, 99)),
                          float(x[:, 2].mean()), float(x[:, 0].mean())])
    per_frame = np.asarray(per_frame)
    dts = np.asarray(dts)
    ref = g["dts"]
    n = min(len(dts), len(ref))
    diff = np.nonzero(dts[:n] != ref[:n])[0]
    print(f"substeps gpu {len(dts)} oracle {len(ref)}; first dt mismatch "
          f"{int(diff[0]) if diff.size else None}", flush=True)
    x, v, rho = (sim.state_0.particle_q.numpy(), sim.state_0.particle_qd.numpy(), sim.state_0.sph.rho.numpy())
    for name, a, b in (("x", x, g["x"]), ("v", v, g["v"]), ("rho", rho, g["rho"])):
        print(f"  {name}: exact fraction {np.mean(a.view(np.uint32) == b.view(np.uint32)):.6f}, "
              f"max |diff| {np.abs(a.astype(np.float64) - b).max():.3e}", flush=True)
    np.testing.assert_array_equal(per_frame[:, 0], g["per_frame"][:, 0])
    np.testing.assert_array_equal(per_frame, g["per_frame"])
    np.testing.assert_array_equal(dts, ref)
    np.testing.assert_array_equal(x.view(np.uint32), g["x"].view(np.uint32))
    np.testing.assert_array_equal(v.view(np.uint32), g["v"].view(np.uint32))
    np.testing.assert_array_equal(rho.view(np.uint32), g["rho"].view(np.uint32))

"""Write the golden trajectory of the CPU-scale WCSPH case (run from the
repository root: ``python tests/golden/make_config0.py``).

The case is the 10,164-particle water column: x in [-1.2,-0.6], y in
[-0.28,0.28], z in [0,0.6], a 22 x 21 x 22 lattice of spacing s = 0.56/21
jittered by default_rng(0).uniform(+-0.01 s).  It uses the dam-break WCSPH
material and config: B 250000, gamma 7, nu 1e-6, no surface tension, the
pressure clamp on, no damping, the tank bounds and h = 2 s.  It runs 60
display frames under the driver's CFL policy with a base of 4 substeps
(examples/dam_break/simulation.py:155-190): the oracle evaluates the policy
on its own state before every substep, as the driver does on the device
state.

The FP32 oracle is the checker here; it is pinned by tests/test_oracle_pins.py
and the golden substep fixtures.  Stored: the inputs, every substep's dt,
per-frame statistics (substeps, front = max x, p99 of rho, mean height,
mean x) and the final x, v, rho.
"""
from __future__ import annotations

import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from fluid_simulation_amd.adjoint import column_scene  # noqa: E402
from fluid_simulation_amd.examples.dam_break.scene import DEFAULT_SPEC  # noqa: E402

FRAMES = 60
SUBSTEPS = 4
STIFFNESS, EXPONENT, VISCOSITY = 250000.0, 7.0, 1.0e-6


def frame_stats(x, rho):
    return [float(x[:, 0].max()), float(np.percentile(rho, 99)), float(x[:, 2].mean()), float(x[:, 0].mean())]


def main(threads: int = 8):
    O.set_threads(threads)
    col = column_scene(10000)
    s = col.spacing
    n = col.count
    spec = DEFAULT_SPEC
    mass, radius = 1000.0 * s ** 3, 0.5 * s
    om = O.uniform_model(n, mass, rest_density=spec.density, surface_tension=0.0, viscosity=VISCOSITY,
                         stiffness=STIFFNESS, exponent=EXPONENT, gravity=spec.gravity, radius=radius)
    cfg = O.WcsphConfig(particle_radius=radius, particle_length=s, smoothing_length_coff=2.0,
                        bound_width=spec.bound_hi[0], bound_height=spec.bound_hi[1],
                        bound_length=spec.bound_hi[2] * 0.5, boundary_damping=0.0, clamp_negative_pressure=True)
    h = O.smoothing_length(cfg)
    frame_dt = 1.0 / spec.fps
    st = O.OState(col.positions.copy(), np.zeros_like(col.positions),
                  {"rho": np.zeros(n, np.float32), "pressure": np.zeros(n, np.float32)})
    accel = None
    dts, per_frame = [], []
    for frame in range(FRAMES):
        remaining, steps = frame_dt, 0
        while remaining > 1.0e-12:
            speed2, accel2 = O.norm_maxima(st.v, accel)
            dt = min(remaining, O.cfl_limit("wcsph", frame_dt, SUBSTEPS, math.sqrt(speed2), math.sqrt(accel2),
                                            h=h, spacing=s, stiffness=STIFFNESS, exponent=EXPONENT,
                                            rest_density=spec.density, viscosity=VISCOSITY))
            st, extras = O.wcsph_step(om, st, cfg, dt)
            accel = extras["accel"]
            dts.append(dt)
            remaining -= dt
            steps += 1
        per_frame.append([steps] + frame_stats(st.x, st.sph["rho"]))
        print(f"frame {frame + 1}: {steps} substeps, front {per_frame[-1][1]:.5f}", flush=True)
    np.savez_compressed(os.path.join(HERE, "config0_column_wcsph_60frames.npz"),
                        positions=col.positions, spacing=np.float64(s), mass=np.float64(mass),
                        radius=np.float64(radius), dts=np.asarray(dts, np.float64),
                        per_frame=np.asarray(per_frame, np.float64), x=st.x, v=st.v, rho=st.sph["rho"])


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 8)

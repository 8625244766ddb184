"""The oracle's own CFL-driven run of the CPU-scale WCSPH column reproduces
the committed trajectory fixture (first 2 of its 60 frames), so a drift of
the oracle is caught on the CPU (the GPU test compares all 60 frames)."""
import math
import os

import numpy as np

from oracle import oracle as O
from tests.golden import make_config0 as M

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "config0_column_wcsph_60frames.npz")


def test_first_frames_reproduce_the_fixture():
    g = np.load(FIXTURE)
    s = float(g["spacing"])
    spec = M.DEFAULT_SPEC
    n = g["positions"].shape[0]
    assert n == 10164 and abs(s - 0.56 / 21) < 1e-15   # s = min(0.6/22, 0.56/21, 0.6/22)
    om = O.uniform_model(n, float(g["mass"]), rest_density=spec.density, surface_tension=0.0,
                         viscosity=M.VISCOSITY, stiffness=M.STIFFNESS, exponent=M.EXPONENT, gravity=spec.gravity,
                         radius=float(g["radius"]))
    cfg = O.WcsphConfig(particle_radius=float(g["radius"]), particle_length=s, smoothing_length_coff=2.0,
                        bound_width=spec.bound_hi[0], bound_height=spec.bound_hi[1],
                        bound_length=spec.bound_hi[2] * 0.5, boundary_damping=0.0, clamp_negative_pressure=True)
    h = O.smoothing_length(cfg)
    st = O.OState(g["positions"].copy(), np.zeros_like(g["positions"]),
                  {"rho": np.zeros(n, np.float32), "pressure": np.zeros(n, np.float32)})
    accel, k = None, 0
    for frame in range(2):
        remaining, steps = 1.0 / spec.fps, 0
        while remaining > 1.0e-12:
            speed2, accel2 = O.norm_maxima(st.v, accel)
            dt = min(remaining, O.cfl_limit("wcsph", 1.0 / spec.fps, M.SUBSTEPS, math.sqrt(speed2),
                                            math.sqrt(accel2), h=h, spacing=s, stiffness=M.STIFFNESS,
                                            exponent=M.EXPONENT, rest_density=spec.density, viscosity=M.VISCOSITY))
            assert dt == g["dts"][k], (frame, k)
            st, extras = O.wcsph_step(om, st, cfg, dt)
            accel = extras["accel"]
            remaining -= dt
            steps += 1
            k += 1
        np.testing.assert_array_equal([steps] + M.frame_stats(st.x, st.sph["rho"]), g["per_frame"][frame])

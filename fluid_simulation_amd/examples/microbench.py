"""Neighbour-search and density micro-benchmark workloads.

Two point distributions in the unit box, sized so that a particle sees about
40 neighbours within the smoothing length h (cell width = h):
  * uniform  -- ``default_rng(1).random((N, 3))`` in [0,1)^3;
                h = (3 * 40 / (4 pi N))^(1/3)
  * clustered -- 64 Gaussian blobs (sigma 0.05 m, centres uniform in
                [0.15, 0.85]^3, ``default_rng(2)``), N/64 points each,
                points outside [0,1]^3 resampled; same h (blob centres are
                ~8x denser than the uniform case)
Mass is 1000/N (unit box of water).  ``wcsph_model`` wraps the points in a
Newton-style model for ``SolverWCSPH`` so one substep times the grid build, the
density/EOS pass and the force pass separately (per-kernel HIP events).
"""
from __future__ import annotations

import math

import numpy as np

from ..newton_compat import Axis, ModelBuilder


def smoothing_length(n: int, neighbours: float = 40.0) -> float:
    return (3.0 * neighbours / (4.0 * math.pi * n)) ** (1.0 / 3.0)


def uniform_points(n: int, seed: int = 1) -> np.ndarray:
    return np.random.default_rng(seed).random((n, 3), dtype=np.float32)


def clustered_points(n: int, blobs: int = 64, sigma: float = 0.05, seed: int = 2) -> np.ndarray:
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.15, 0.85, (blobs, 3))
    per = [n // blobs + (1 if b < n % blobs else 0) for b in range(blobs)]
    out = []
    for c, k in zip(centres, per):
        pts = rng.normal(c, sigma, (k, 3))
        bad = np.any((pts < 0.0) | (pts >= 1.0), axis=1)
        while bad.any():
            pts[bad] = rng.normal(c, sigma, (int(bad.sum()), 3))
            bad = np.any((pts < 0.0) | (pts >= 1.0), axis=1)
        out.append(pts)
    return np.concatenate(out).astype(np.float32)


def wcsph_model(points: np.ndarray, device="cuda:0"):
    """Model + SolverWCSPH config for one substep over ``points`` (water,
    B=250000, gamma=7, nu=1e-6, box walls at the unit cube)."""
    from ..solver.sph.solver_wcsph import SolverWCSPH

    n = points.shape[0]
    h = smoothing_length(n)
    b = ModelBuilder(up_axis=Axis.Z)
    SolverWCSPH.register_custom_attributes(b)
    b.add_particles(pos=points, vel=np.zeros_like(points), mass=np.full(n, 1000.0 / n, np.float32),
                    radius=np.full(n, 0.25 * h, np.float32),
                    custom_attributes={"sph:rest_density": np.full(n, 1000.0, np.float32),
                                       "sph:surface_tension": np.zeros(n, np.float32),
                                       "sph:viscosity": np.full(n, 1.0e-6, np.float32),
                                       "sph:stiffness": np.full(n, 250000.0, np.float32),
                                       "sph:exponent": np.full(n, 7.0, np.float32),
                                       "sph:particle_mask": np.zeros(n, np.int32)})
    model = b.finalize(device=device)
    model.set_gravity((0.0, 0.0, -10.0))
    cfg = SolverWCSPH.SphConfig(particle_radius=0.25 * h, particle_length=0.5 * h, smoothing_length_coff=2.0,
                                bound_width=2.0, bound_height=2.0, bound_length=1.0, boundary_damping=0.0)
    return model, cfg

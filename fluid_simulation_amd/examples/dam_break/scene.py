"""Canonical dam-break particle lattice.

Counterpart of the reference's ``examples/dam_break/scene.py`` (same spec
fields, same endpoint-lattice and resolution-rounding rules, same SHA-256 input
fingerprint), so that the same target count yields bit-identical positions.
"""
from __future__ import annotations

from dataclasses import asdict, dataclass
import hashlib

import numpy as np


@dataclass(frozen=True)
class SceneSpec:
    """Scene specification  (reference scene.py:9-19)."""
    bound_lo: tuple = (-1.20, -0.28, 0.0)
    bound_hi: tuple = (1.20, 0.28, 1.20)
    fill_x_frac: float = 0.30
    fill_z_frac: float = 0.72
    margin: float = 0.03
    density: float = 1000.0
    gravity: tuple = (0.0, 0.0, -10.0)
    fps: float = 60.0
    target_count: int = 400000


DEFAULT_SPEC = SceneSpec()


def spawn_particle_block(lo, hi, res, density: float = 1000.0):
    """Endpoint lattice of ``res`` points per axis  (reference scene.py:25-41)."""
    axes = [np.linspace(lo[a], hi[a], res[a]) for a in range(3)]
    grid = np.meshgrid(*axes, indexing="ij")
    positions = np.stack([g.ravel() for g in grid], axis=1).astype(np.float32)
    spacing = float(min((hi[a] - lo[a]) / max(res[a] - 1, 1) for a in range(3)))
    volume = spacing ** 3
    return positions, density * volume, 0.5 * spacing, volume, spacing


def spawn_wave_tank(tank_lo, tank_hi, target_count: int = 400000, fill_x_frac: float = 0.30,
                    fill_z_frac: float = 0.70, margin: float = 0.03, density: float = 1000.0):
    """Water column with rounded per-axis resolution  (reference scene.py:44-75)."""
    if target_count < 1:
        raise ValueError("Target particle count must be positive")
    tank_lo = np.asarray(tank_lo, dtype=np.float32)
    tank_hi = np.asarray(tank_hi, dtype=np.float32)
    extent = tank_hi - tank_lo
    water_lo = np.array([tank_lo[0] + margin, tank_lo[1] + margin, tank_lo[2] + margin], dtype=np.float32)
    water_hi = np.array([
        tank_lo[0] + margin + fill_x_frac * max(extent[0] - 2.0 * margin, 1.0e-3),
        tank_hi[1] - margin,
        tank_lo[2] + margin + fill_z_frac * max(extent[2] - 2.0 * margin, 1.0e-3),
    ], dtype=np.float32)
    water_ext = np.maximum(water_hi - water_lo, 1.0e-3)
    water_volume = float(np.prod(water_ext))
    spacing = float((water_volume / max(target_count, 1)) ** (1.0 / 3.0))
    res = tuple(max(2, int(round(e / spacing))) for e in water_ext)
    return spawn_particle_block(tuple(water_lo), tuple(water_hi), res, density=density)


@dataclass
class ParticleScene:
    """Positions plus per-particle mass/radius  (reference scene.py:78-115)."""
    spec: SceneSpec
    positions: np.ndarray
    mass: float
    radius: float
    volume: float
    spacing: float

    @property
    def count(self):
        return len(self.positions)

    @property
    def rest_volume(self):
        return self.count * self.mass / self.spec.density

    @property
    def ideal_height(self):
        lo, hi = self.spec.bound_lo, self.spec.bound_hi
        return self.rest_volume / ((hi[0] - lo[0]) * (hi[1] - lo[1]))

    def report(self):
        return {
            **asdict(self.spec),
            "actual_count": self.count,
            "lattice_resolution": [len(np.unique(self.positions[:, i])) for i in range(3)],
            "particle_spacing": self.spacing,
            "particle_radius": self.radius,
            "particle_volume": self.volume,
            "particle_mass": self.mass,
            "rest_volume": self.rest_volume,
            "ideal_flat_height": self.ideal_height,
            "initial_position_sha256": hashlib.sha256(self.positions.tobytes()).hexdigest(),
            "water_lo": self.positions.min(axis=0).tolist(),
            "water_hi": self.positions.max(axis=0).tolist(),
            "units": "meters, kilograms, seconds; Z up",
        }


def make_scene(particle_count=DEFAULT_SPEC.target_count):
    """Dam-break scene for a target count  (reference scene.py:118-123)."""
    spec = SceneSpec(target_count=int(particle_count))
    return ParticleScene(spec, *spawn_wave_tank(spec.bound_lo, spec.bound_hi, spec.target_count, spec.fill_x_frac,
                                                spec.fill_z_frac, spec.margin, spec.density))

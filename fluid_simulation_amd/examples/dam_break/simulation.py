"""Dam-break driver for the gfx950 solvers.

Counterpart of the reference's ``examples/dam_break/simulation.py``
(``DamBreakSimulation``, :56-241): the same model construction and material,
the same per-method solver configs, the same CFL sub-stepping policy
(:155-190), device reductions for the CFL maxima and the non-finite check
(:19-53, :140-153, :227-235) and the same ``metrics``/``report`` dicts.
The APIC grid solver is not part of this path.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import asdict

import numpy as np
import torch

from ... import _lib
from ...array import Array
from ...newton_compat import Axis, ModelBuilder, ShapeConfig
from ...solver.pbf import SolverPBF
from ...solver.sph.solver_dfsph import SolverDFSPH
from ...solver.sph.solver_wcsph import SolverWCSPH

SOLVERS = {"wcsph": SolverWCSPH, "dfsph": SolverDFSPH, "pbf": SolverPBF}


def solver_config(method: str, scene, *, wcsph_smoothing_length_coff=2.0, wcsph_clamp_negative_pressure=True):
    """Per-method solver config of the dam break (simulation.py:110-132)."""
    bounds = dict(bound_width=scene.spec.bound_hi[0], bound_height=scene.spec.bound_hi[1],
                  bound_length=scene.spec.bound_hi[2] * 0.5)
    common = dict(particle_radius=scene.radius, particle_length=scene.spacing, smoothing_length_coff=2.0,
                  boundary_damping=0.0, **bounds)
    if method == "wcsph":
        common["smoothing_length_coff"] = wcsph_smoothing_length_coff
        return SolverWCSPH.SphConfig(clamp_negative_pressure=wcsph_clamp_negative_pressure, **common)
    if method == "dfsph":
        return SolverDFSPH.SphConfig(max_rho_iterations=5, max_vel_iterations=12, warm_start=False, max_kappa=1.0,
                                     max_kappa_v=100.0, **common)
    if method == "pbf":
        return SolverPBF.PbfConfig(max_iterations=4, lambda_epsilon=100.0, min_neighbors_for_lambda=8, scorr_k=0.0,
                                   max_delta_position=scene.radius * 0.5, xsph_c=0.05, **common)
    raise ValueError(f"unknown method {method!r}; expected one of {sorted(SOLVERS)}")


def material(method: str, scene, *, wcsph_stiffness=250000.0, wcsph_viscosity=1.0e-6):
    """sph: material attributes of the dam break (simulation.py:80-92)."""
    mat = {"rest_density": scene.spec.density, "surface_tension": 0.0, "viscosity": 1.0e-6}
    if method == "wcsph":
        mat.update(stiffness=wcsph_stiffness, exponent=7.0, viscosity=wcsph_viscosity)
    return mat


class DamBreakSimulation:
    def __init__(self, scene, method, device="cuda:0", substeps=4, voxel_size=None, enable_projection=True, *,
                 wcsph_stiffness=250000.0, wcsph_viscosity=1.0e-6, wcsph_smoothing_length_coff=2.0,
                 wcsph_clamp_negative_pressure=True):
        del voxel_size, enable_projection   # APIC-only options
        if method not in SOLVERS:
            raise ValueError(f"unknown method {method!r}; expected one of {sorted(SOLVERS)}")
        if method == "wcsph":
            for name, value in (("stiffness", wcsph_stiffness),
                                ("smoothing_length_coff", wcsph_smoothing_length_coff)):
                if not math.isfinite(value) or value <= 0.0:
                    raise ValueError(f"WCSPH {name} must be finite and positive")
            if not math.isfinite(wcsph_viscosity) or wcsph_viscosity < 0.0:
                raise ValueError("WCSPH viscosity must be finite and non-negative")
        self.scene = scene
        self.method = method
        self.frame_dt = 1.0 / scene.spec.fps
        self.substeps = substeps
        self.frame = 0
        self.sim_time = 0.0
        self.last_substeps = 0
        self.total_substeps = 0
        solver_cls = SOLVERS[method]
        builder = ModelBuilder(up_axis=Axis.Z)
        builder.default_particle_radius = scene.radius
        solver_cls.register_custom_attributes(builder)
        self.material = material(method, scene, wcsph_stiffness=wcsph_stiffness, wcsph_viscosity=wcsph_viscosity)
        attributes = {f"sph:{key}": [value] * scene.count for key, value in self.material.items()}
        attributes["sph:particle_mask"] = [0] * scene.count
        builder.add_particles(pos=scene.positions, vel=np.zeros_like(scene.positions),
                              mass=[scene.mass] * scene.count, radius=[scene.radius] * scene.count,
                              custom_attributes=attributes)
        builder.add_ground_plane(cfg=ShapeConfig(density=0.0, mu=0.0))
        self.model = builder.finalize(device=device)
        self.model.set_gravity(scene.spec.gravity)
        self.state_0, self.state_1 = self.model.state(), self.model.state()
        config = solver_config(method, scene, wcsph_smoothing_length_coff=wcsph_smoothing_length_coff,
                               wcsph_clamp_negative_pressure=wcsph_clamp_negative_pressure)
        self.solver = solver_cls(self.model, config)
        dev = self.model.device
        self._invalid_state = torch.zeros(1, dtype=torch.int32, device=dev)
        self._norm_maxima = torch.zeros(2, dtype=torch.float32, device=dev)

    def _stream(self):
        return ctypes.c_void_p(_lib.stream_handle(self.model.device))

    def _cfl_maxima(self):
        """Reduce on the device; each CFL substep reads back only two scalars (simulation.py:140-153)."""
        v = self.state_0.particle_qd
        use_accel = self.method == "wcsph" and self.total_substeps
        a = self.solver._particle_accel
        _lib.check(_lib.load().fs_norm_maxima(ctypes.c_void_p(v.ptr), v.size,
                                              ctypes.c_void_p(a.ptr if use_accel else 0),
                                              a.size if use_accel else 0,
                                              ctypes.c_void_p(self._norm_maxima.data_ptr()), self._stream()),
                   "fs_norm_maxima")
        speed2, accel2 = (float(x) for x in self._norm_maxima.cpu().tolist())
        if not math.isfinite(speed2):
            raise RuntimeError(f"Non-finite velocity at frame {self.frame}")
        if not math.isfinite(accel2):
            raise RuntimeError("Non-finite WCSPH acceleration")
        return math.sqrt(speed2), math.sqrt(accel2)

    def cfl_limit(self, speed: float, accel_max: float) -> float:
        """Per-substep dt limit (simulation.py:167-181)."""
        limit = self.frame_dt / self.substeps
        if self.method == "wcsph":
            sound_speed = math.sqrt(self.material["stiffness"] * self.material["exponent"] / self.scene.spec.density)
            limit = min(limit, 0.25 * self.solver.smoothing_length / (sound_speed + speed))
            if accel_max > 0.0:
                limit = min(limit, 0.25 * math.sqrt(self.solver.smoothing_length / accel_max))
            viscosity = self.material["viscosity"]
            if viscosity > 0.0:
                limit = min(limit, 0.025 * self.solver.smoothing_length ** 2 / viscosity)
        elif speed > 0.0:
            limit = min(limit, 0.4 * self.scene.spacing / speed)
        return limit

    def step(self):
        """Advance a complete display frame; never discard unused CFL time (simulation.py:155-190)."""
        if self.method == "pbf":
            steps = self.substeps
            dt = self.frame_dt / steps
            for _ in range(steps):
                self._substep(dt)
        else:
            remaining = self.frame_dt
            steps = 0
            while remaining > 1.0e-12:
                speed, accel_max = self._cfl_maxima()
                dt = min(remaining, self.cfl_limit(speed, accel_max))
                if dt < 1.0e-10 or steps >= 4096:
                    raise RuntimeError("CFL step cannot safely complete this frame")
                self._substep(dt)
                remaining -= dt
                steps += 1
        self.frame += 1
        self.sim_time = self.frame * self.frame_dt
        self.last_substeps = steps

    def _substep(self, dt):
        self.solver.step(self.state_0, self.state_1, control=None, contacts=None, dt=dt)
        self.state_0, self.state_1 = self.state_1, self.state_0
        self.total_substeps += 1

    def metrics(self):
        """simulation.py:198-225"""
        x, v = self.state_0.particle_q.numpy(), self.state_0.particle_qd.numpy()
        finite = bool(np.isfinite(x).all() and np.isfinite(v).all())
        if not finite:
            raise RuntimeError(f"Non-finite particle state at frame {self.frame}")
        result = dict(frame=self.frame, simulation_time=self.sim_time, substeps=self.last_substeps,
                      total_substeps=self.total_substeps, finite=finite,
                      speed_max=float(np.linalg.norm(v, axis=1).max(initial=0.0)),
                      height_mean=float(x[:, 2].mean()), height_min=float(x[:, 2].min()),
                      height_max=float(x[:, 2].max()))
        if self.frame:
            rho = self.state_0.sph.rho.numpy()
            if not np.isfinite(rho).all():
                raise RuntimeError("Non-finite density")
            result.update(density_mean=float(rho.mean()), density_min=float(rho.min()),
                          density_max=float(rho.max()))
            if self.method == "wcsph":
                pressure = self.state_0.sph.pressure.numpy()
                if not np.isfinite(pressure).all():
                    raise RuntimeError("Non-finite pressure")
                density_p99 = float(np.percentile(rho, 99))
                result.update(density_p99=density_p99,
                              density_compression_p99=max(0.0, density_p99 / self.scene.spec.density - 1.0),
                              pressure_min=float(pressure.min()), pressure_max=float(pressure.max()))
        return result

    def check_finite(self):
        """Validate on the device and read back a single integer (simulation.py:227-235)."""
        x, v, rho = self.state_0.particle_q, self.state_0.particle_qd, self.state_0.sph.rho
        _lib.check(_lib.load().fs_check_finite(ctypes.c_void_p(x.ptr), ctypes.c_void_p(v.ptr), x.size,
                                               ctypes.c_void_p(rho.ptr), rho.size,
                                               ctypes.c_void_p(self._invalid_state.data_ptr()), self._stream()),
                   "fs_check_finite")
        if int(self._invalid_state.item()):
            raise RuntimeError(f"Non-finite particle state at frame {self.frame}")

    def report(self):
        """simulation.py:237-241"""
        return dict(method=self.method, config=asdict(self.solver.config), material=self.material,
                    frame_dt=self.frame_dt, base_substeps=self.substeps,
                    time_step="fixed" if self.method == "pbf" else "CFL, complete frame",
                    solver_execution="fused HIP passes (libfluidsim_hip)")

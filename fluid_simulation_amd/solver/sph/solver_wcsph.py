"""Weakly compressible SPH on the gfx950 path.

Drop-in for the reference's ``solver/sph/solver_wcsph.py`` ``SolverWCSPH``:
same class name, ``SphConfig`` fields and defaults (:309-318), ``sph:``
attributes (:320-392), ``smoothing_length`` (:408-410), ``_particle_accel``
and ``step(state_in, state_out, control, contacts, dt)`` (:431-535).  One call
of ``step`` enqueues, on the current stream, the grid build and two fused
neighbour passes (density+EOS; force+integrate+bounds) of
``libfluidsim_hip.so`` (``fs_wcsph_step``).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from ... import _lib
from ...array import float32
from ...newton_compat import AttributeAssignment
from .._base import ParticleSolver, ptr, register, register_common_model_attributes


class SolverWCSPH(ParticleSolver):
    _name = "SolverWCSPH"

    @dataclass
    class SphConfig:
        particle_radius: float = 0.1
        particle_length: float = 0.2
        smoothing_length_coff: float = 1.35
        bound_width: float = 10.0
        bound_height: float = 10.0
        bound_length: float = 10.0
        boundary_damping: float = -0.1
        clamp_negative_pressure: bool = True

    @classmethod
    def register_custom_attributes(cls, builder) -> None:
        """solver_wcsph.py:320-392"""
        register_common_model_attributes(builder)
        register(builder, "stiffness", float32, 85000.0, AttributeAssignment.MODEL)
        register(builder, "exponent", float32, 7.0, AttributeAssignment.MODEL)
        for name in ("rho", "pressure"):
            register(builder, name, float32, 0.0, AttributeAssignment.STATE)

    def __init__(self, model, config: SphConfig, **kwargs):
        super().__init__(model, config, **kwargs)

    def step(self, state_in, state_out, control, contacts, dt: float) -> None:
        del control, contacts
        if not self._check_step(state_in, state_out):
            return
        cfg = self.config
        p = _lib.FsWcsphParams()
        p.smoothing_length = float(self.smoothing_length)
        p.particle_radius = self._particle_radius()
        p.bound[:] = self._bounds
        p.boundary_damping = float(cfg.boundary_damping)
        if isinstance(dt, torch.Tensor):
            # a device scalar (float32, one element): no host value needed, so a
            # device-side dt policy and graph capture can drive whole frames
            if dt.dtype != torch.float32 or dt.numel() != 1 or dt.device.type != "cuda":
                raise ValueError("a tensor dt must be one float32 on the solver's device")
            p.dt = 0.0
            p.dt_device = dt.data_ptr()
        else:
            p.dt = float(dt)
        p.clamp_negative_pressure = 1 if cfg.clamp_negative_pressure else 0
        io = _lib.FsWcsphIO()
        io.x_in = ptr(state_in.particle_q)
        io.v_in = ptr(state_in.particle_qd)
        io.x_out = ptr(state_out.particle_q)
        io.v_out = ptr(state_out.particle_qd)
        io.rho = ptr(state_out.sph.rho)
        io.pressure = ptr(state_out.sph.pressure)
        io.accel = ptr(self._particle_accel)
        io.rho_in = ptr(state_in.sph.rho)
        io.pressure_in = ptr(state_in.sph.pressure)
        m = self._model_struct(with_eos=True)
        lib = _lib.load()
        _lib.check(lib.fs_wcsph_step(ctypes.byref(self.model.particle_grid._c), ctypes.byref(self._work),
                                     ctypes.byref(m), ctypes.byref(p), ctypes.byref(io),
                                     ctypes.c_void_p(self._stream())), "fs_wcsph_step")
        self._lists_after_step()

"""Position-based fluids on the gfx950 path.

Drop-in for the reference's ``solver/pbf/solver_pbf.py`` ``SolverPBF``: same
``PbfConfig`` fields and defaults (:521-539), ``sph:`` attributes (:541-628),
``_w_delta_q`` (:671-674) and ``step`` (:676-954): one grid build on the
gravity-predicted positions, then density, non-pressure, ``max_iterations`` x
(lambda, delta-position) and final density/XSPH passes, all enqueued by one
``fs_pbf_step`` call.  No host synchronisation, so the step can be captured in
a HIP graph.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

from ... import _lib
from ...array import float32, int32, vec3
from ...newton_compat import AttributeAssignment
from .._base import ParticleSolver, ptr, register, register_common_model_attributes


def poly6(r_norm: float, h: float) -> float:
    """The PBF ``get_cubic`` (poly6, solver_pbf.py:23-32) evaluated in Python
    floats, as the reference's host-side call at :671-674 evaluates it."""
    if h > 0.0 and r_norm < h:
        h2 = h * h
        rhs = h2 - r_norm * r_norm
        c = h * h * h
        k = 315.0 / (64.0 * math.pi * (c * c * c))
        return k * (rhs * rhs * rhs)
    return 0.0


class SolverPBF(ParticleSolver):
    _name = "SolverPBF"
    #: the PBF dam break reaches ~80 neighbours within h in its splash
    neighbor_list_capacity = 128

    @dataclass
    class PbfConfig:
        max_iterations: int = 10
        lambda_epsilon: float = 100.0
        min_neighbors_for_lambda: int = 0
        scorr_k: float = 1.0e-4
        scorr_n: float = 4.0
        scorr_q: float = 0.3
        max_delta_position: float = 0.0
        xsph_c: float = 0.01
        particle_radius: float = 0.1
        particle_length: float = 0.2
        smoothing_length_coff: float = 1.35
        bound_width: float = 10.0
        bound_height: float = 10.0
        bound_length: float = 10.0
        boundary_damping: float = 0.0
        boundary_epsilon: float = 0.0
        rebuild_grid_each_iteration: bool = False

    @classmethod
    def register_custom_attributes(cls, builder) -> None:
        """solver_pbf.py:541-628"""
        register_common_model_attributes(builder)
        register(builder, "rho", float32, 0.0, AttributeAssignment.STATE)
        register(builder, "pbf_lambda", float32, 0.0, AttributeAssignment.STATE)
        register(builder, "neighbor_count", int32, 0, AttributeAssignment.STATE)
        register(builder, "delta_p", vec3, (0.0, 0.0, 0.0), AttributeAssignment.STATE)
        register(builder, "step_displacement", vec3, (0.0, 0.0, 0.0), AttributeAssignment.STATE)

    def __init__(self, model, config: PbfConfig, **kwargs):
        super().__init__(model, config, **kwargs)
        self._boundary_epsilon = float(config.boundary_epsilon)

    def _w_delta_q(self, smoothing_length: float) -> float:
        q = float(self.config.scorr_q)
        return float(poly6(q * smoothing_length, smoothing_length))

    def step(self, state_in, state_out, control, contacts, dt: float) -> None:
        del control, contacts
        if not self._check_step(state_in, state_out):
            return
        cfg = self.config
        h = float(self.smoothing_length)
        p = _lib.FsPbfParams()
        p.smoothing_length = h
        p.particle_radius = self._particle_radius()
        p.bound[:] = self._bounds
        p.boundary_damping = float(cfg.boundary_damping)
        p.boundary_epsilon = self._boundary_epsilon
        p.dt = float(dt)
        p.lambda_epsilon = float(cfg.lambda_epsilon)
        p.scorr_k = float(cfg.scorr_k)
        p.scorr_n = float(cfg.scorr_n)
        p.scorr_w_delta_q = self._w_delta_q(h)
        p.max_delta_position = float(cfg.max_delta_position)
        p.xsph_c = float(cfg.xsph_c)
        p.max_iterations = int(cfg.max_iterations)
        p.min_neighbors_for_lambda = int(cfg.min_neighbors_for_lambda)
        p.rebuild_grid_each_iteration = 1 if cfg.rebuild_grid_each_iteration else 0
        so = state_out.sph
        io = _lib.FsPbfIO()
        io.x_in = ptr(state_in.particle_q)
        io.v_in = ptr(state_in.particle_qd)
        io.x_out = ptr(state_out.particle_q)
        io.v_out = ptr(state_out.particle_qd)
        io.rho = ptr(so.rho)
        io.pbf_lambda = ptr(so.pbf_lambda)
        io.neighbor_count = ptr(so.neighbor_count)
        io.delta_p = ptr(so.delta_p)
        io.accel = ptr(self._particle_accel)
        m = self._model_struct()
        _lib.check(_lib.load().fs_pbf_step(ctypes.byref(self.model.particle_grid._c), ctypes.byref(self._work),
                                           ctypes.byref(m), ctypes.byref(p), ctypes.byref(io),
                                           ctypes.c_void_p(self._stream())), "fs_pbf_step")
        self._lists_after_step()

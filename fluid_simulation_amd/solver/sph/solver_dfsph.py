"""Divergence-free SPH on the gfx950 path.

Drop-in for the reference's ``solver/sph/solver_dfsph.py`` ``SolverDFSPH``:
same ``SphConfig`` fields and defaults (:446-463), ``sph:`` attributes
(:465-535) and ``step`` (:598-898).  The constant-density and divergence-free
loops keep the reference's early exit (``iter > 0 and residual <= eta``,
:733-735, :849-851) but decide it on the device, so ``step`` never
synchronises the host.  The residual is the reference's numpy float32 mean
over the dynamic particles (:579-596): a float64 mean from the rate pass
decides whenever it lies farther from eta than float32 rounding reaches, and
the float32 mean is formed in numpy's summation order otherwise (or always,
with ``exact_residual = True``).  ``last_iterations()`` reads back the loop
counts of the previous step.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from ... import _lib
from ...array import float32, int32, vec3
from ...newton_compat import AttributeAssignment
from .._base import ParticleSolver, ptr, register, register_common_model_attributes


class SolverDFSPH(ParticleSolver):
    _name = "SolverDFSPH"
    #: the dam break's DFSPH config (max kappa 1.0) lets the flow compress
    #: into clumps of up to ~270 neighbours within h; an overflowing particle
    #: rescans the grid in every replaying pass, so the lists are deep
    neighbor_list_capacity = 384

    @dataclass
    class SphConfig:
        max_rho_iterations: int = 2
        max_vel_iterations: int = 1
        warm_start: bool = True
        eta_density: float = 1.0e-1
        eta_divergence: float = 1.0e-4
        max_kappa: float = 1.0e4
        max_kappa_v: float = 1.0e6
        min_neighbors_for_factor: int = 15
        particle_radius: float = 0.1
        particle_length: float = 0.2
        smoothing_length_coff: float = 1.35
        bound_width: float = 10.0
        bound_height: float = 10.0
        bound_length: float = 10.0
        boundary_damping: float = -0.1
        voxel_size: float = 0.3

    @classmethod
    def register_custom_attributes(cls, builder) -> None:
        """solver_dfsph.py:465-535"""
        register_common_model_attributes(builder)
        for name in ("rho", "factor", "rho_star", "d_rho_dt", "kappa", "kappa_v"):
            register(builder, name, float32, 0.0, AttributeAssignment.STATE)
        register(builder, "neighbor_count", int32, 0, AttributeAssignment.STATE)
        register(builder, "v_star", vec3, (0.0, 0.0, 0.0), AttributeAssignment.STATE)

    #: form numpy's float32 residual for every break test, not only for
    #: residuals within float32 rounding of eta (same decisions, slower)
    exact_residual = False

    def __init__(self, model, config: SphConfig, **kwargs):
        super().__init__(model, config, **kwargs)
        self._iterations = torch.zeros(4, dtype=torch.int32, device=model.device)

    def last_iterations(self) -> dict:
        """Loop counts of the last step (host readback)."""
        it = self._iterations.cpu().tolist()
        return {"density_iterations": it[0], "density_corrections": it[1], "divergence_iterations": it[2],
                "divergence_corrections": it[3]}

    def step(self, state_in, state_out, control, contacts, dt: float) -> None:
        del control, contacts
        if not self._check_step(state_in, state_out):
            return
        cfg = self.config
        p = _lib.FsDfsphParams()
        p.eta_density = float(cfg.eta_density)
        p.eta_divergence = float(cfg.eta_divergence)
        p.smoothing_length = float(self.smoothing_length)
        p.particle_radius = self._particle_radius()
        p.bound[:] = self._bounds
        p.boundary_damping = float(cfg.boundary_damping)
        if isinstance(dt, torch.Tensor):
            # a device scalar (float32, one element), as for SolverWCSPH: a
            # device-side dt policy and graph capture drive whole frames; a
            # device dt of 0 is an idle substep that carries the state over
            if dt.dtype != torch.float32 or dt.numel() != 1 or dt.device.type != "cuda":
                raise ValueError("a tensor dt must be one float32 on the solver's device")
            p.dt = 0.0
            p.dt_device = dt.data_ptr()
        else:
            p.dt = float(dt)
        p.max_kappa = float(cfg.max_kappa)
        p.max_kappa_v = float(cfg.max_kappa_v)
        p.max_rho_iterations = int(cfg.max_rho_iterations)
        p.max_vel_iterations = int(cfg.max_vel_iterations)
        p.warm_start = 1 if cfg.warm_start else 0
        p.min_neighbors_for_factor = int(cfg.min_neighbors_for_factor)
        p.exact_residual = 1 if self.exact_residual else 0
        si, so = state_in.sph, state_out.sph
        io = _lib.FsDfsphIO()
        io.x_in = ptr(state_in.particle_q)
        io.v_in = ptr(state_in.particle_qd)
        io.kappa_in = ptr(si.kappa)
        io.kappa_v_in = ptr(si.kappa_v)
        io.x_out = ptr(state_out.particle_q)
        io.v_out = ptr(state_out.particle_qd)
        io.rho = ptr(so.rho)
        io.factor = ptr(so.factor)
        io.rho_star = ptr(so.rho_star)
        io.d_rho_dt = ptr(so.d_rho_dt)
        io.kappa = ptr(so.kappa)
        io.kappa_v = ptr(so.kappa_v)
        io.neighbor_count = ptr(so.neighbor_count)
        io.v_star = ptr(so.v_star)
        io.accel = ptr(self._particle_accel)
        io.iterations = self._iterations.data_ptr()
        io.rho_in = ptr(si.rho)
        io.factor_in = ptr(si.factor)
        io.rho_star_in = ptr(si.rho_star)
        io.d_rho_dt_in = ptr(si.d_rho_dt)
        io.neighbor_count_in = ptr(si.neighbor_count)
        io.v_star_in = ptr(si.v_star)
        m = self._model_struct()
        _lib.check(_lib.load().fs_dfsph_step(ctypes.byref(self.model.particle_grid._c), ctypes.byref(self._work),
                                             ctypes.byref(m), ctypes.byref(p), ctypes.byref(io),
                                             ctypes.c_void_p(self._stream())), "fs_dfsph_step")
        self._lists_after_step()

"""APIC particle separation on the GPU.

The reference's APIC solver pushes apart particles closer than a rest
distance after each substep, to counter clumping
(``solver/apic/solver_apic.py:193-268``, ``SolverAPIC._apply_particle_separation``
``:1003-1066``).  The pass runs on the same hash grid as the SPH passes.

``ParticleSeparation(model, config)`` holds what ``SolverAPIC`` keeps for the
pass:
* its own ``HashGrid`` of the box dims;
* the wall margin and push band (``SolverAPIC.__init__``, ``:640-650``);
* the rest distance (``:1009-1016``).

``apply(state)`` is ``_apply_particle_separation``: ``iterations`` x (grid
build at the current positions, accumulate, apply), in place on
``state.particle_q``.  One fused kernel per iteration does both of the
reference's launches.  ``config`` is any object with the ``SolverAPIC.Config``
fields the pass reads; ``SeparationConfig`` lists them with the reference's
defaults, and a ``SolverAPIC.Config`` instance can be passed as is.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .array import Array
from .hashgrid import HashGrid


@dataclass
class SeparationConfig:
    """``SolverAPIC.Config`` fields read by the separation pass, with the
    reference's defaults (solver_apic.py:557-582)."""
    bound_lo: tuple[float, float, float] = (-1.0, -1.0, 0.0)
    bound_hi: tuple[float, float, float] = (1.0, 1.0, 2.0)
    clamp_eps: float = 1.0e-4
    wall_margin: float = -1.0            # <0 => mean particle radius
    wall_push_band: float = -1.0         # <0 => 1.5 * wall_margin
    enable_particle_separation: bool = True
    separation_rest_dist: float = -1.0   # <0 => 2 * mean particle radius
    separation_strength: float = 0.55
    separation_iterations: int = 2


class ParticleSeparation:
    def __init__(self, model, config):
        self.model = model
        self.config = config
        self.device = torch.device(model.device)
        n = int(model.particle_count)
        _lib.require_device(self.device)
        radii = model.particle_radius.numpy() if n > 0 else np.zeros(0, np.float32)
        # SolverAPIC.__init__ (:640-650)
        if n > 0 and float(config.wall_margin) < 0.0:
            self.wall_margin = float(np.mean(radii))
        elif float(config.wall_margin) < 0.0:
            self.wall_margin = float(config.clamp_eps)
        else:
            self.wall_margin = float(config.wall_margin)
        if float(config.wall_push_band) < 0.0:
            self.wall_push_band = 1.5 * self.wall_margin
        else:
            self.wall_push_band = float(config.wall_push_band)
        # _apply_particle_separation (:1009-1027)
        if float(config.separation_rest_dist) > 0.0:
            self.rest_dist = float(config.separation_rest_dist)
        else:
            self.rest_dist = float(2.0 * float(np.mean(radii))) if n > 0 else 0.0
        self.query_radius = max(self.rest_dist, 1.0e-5)
        lo = np.asarray(config.bound_lo, dtype=np.float64)
        hi = np.asarray(config.bound_hi, dtype=np.float64)
        extent = np.maximum(hi - lo, self.query_radius)
        dims = np.maximum(4, np.ceil(extent / self.query_radius).astype(int))
        self.dims = (int(dims[0]), int(dims[1]), int(dims[2]))
        self.grid = HashGrid(*self.dims, device=self.device)
        self.delta = Array(torch.zeros((n, 3), dtype=torch.float32, device=self.device))
        self._n = n
        if n > 0:
            self.grid.reserve(n)
            lib = _lib.load()
            nbytes = lib.fs_work_bytes(n, 0)
            self._work_mem = torch.empty(int(nbytes), dtype=torch.uint8, device=self.device)
            self._work = _lib.FsWork()
            _lib.check(lib.fs_work_init(ctypes.byref(self._work), ctypes.c_void_p(self._work_mem.data_ptr()), nbytes,
                                        n, 0, ctypes.c_void_p(_lib.stream_handle(self.device))), "fs_work_init")

    def params(self) -> _lib.FsSeparationParams:
        cfg = self.config
        p = _lib.FsSeparationParams()
        p.rest_dist = self.rest_dist
        p.query_radius = self.query_radius
        p.strength = float(cfg.separation_strength)
        p.wall_band = max(float(self.wall_push_band), float(self.wall_margin) * 2.0)
        p.wall_margin = float(self.wall_margin)
        for d in range(3):
            p.bound_lo[d] = float(cfg.bound_lo[d])
            p.bound_hi[d] = float(cfg.bound_hi[d])
        p.iterations = max(1, int(cfg.separation_iterations))
        return p

    def apply(self, state) -> None:
        """SolverAPIC._apply_particle_separation(state)  (solver_apic.py:1003-1066)."""
        n = self._n
        if n <= 0:
            return
        x = state.particle_q
        t = x.tensor if isinstance(x, Array) else x
        if t.dtype != torch.float32 or not t.is_cuda or tuple(t.shape) != (n, 3) or not t.is_contiguous():
            raise ValueError(f"particle_q must be a contiguous float32 ({n}, 3) tensor on {self.device}")
        p = self.params()
        if not math.isfinite(p.query_radius):
            raise ValueError("separation rest distance must be finite")
        _lib.check(_lib.load().fs_particle_separation(
            ctypes.byref(self.grid._c), ctypes.byref(self._work), ctypes.byref(p), ctypes.c_void_p(t.data_ptr()),
            ctypes.c_void_p(self.model.particle_flags.ptr), ctypes.c_void_p(self.delta.ptr), n,
            ctypes.c_void_p(_lib.stream_handle(self.device))), "fs_particle_separation")


def apply_particle_separation(model, state, config) -> ParticleSeparation:
    """One-shot form: build the pass for ``model`` and apply it to ``state``."""
    sep = ParticleSeparation(model, config)
    sep.apply(state)
    return sep

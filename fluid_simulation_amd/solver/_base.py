"""Shared host plumbing of the particle solvers: model descriptor, workspace,
passthrough copies and the ``sph:`` attribute registration helpers."""
from __future__ import annotations

import ctypes

import torch

from .. import _lib
from ..array import Array, copy, float32, int32, vec3
from ..hashgrid import HashGrid
from ..newton_compat import AttributeAssignment, AttributeFrequency, CustomAttribute, SolverBase


def register(builder, name, dtype, default, assignment):
    builder.add_custom_attribute(CustomAttribute(name=name, frequency=AttributeFrequency.PARTICLE,
                                                 assignment=assignment, dtype=dtype, default=default,
                                                 namespace="sph"))


def register_common_model_attributes(builder) -> None:
    """particle_mask, rest_density, surface_tension, viscosity (solver_wcsph.py:322-361)."""
    register(builder, "particle_mask", int32, 0, AttributeAssignment.MODEL)
    register(builder, "rest_density", float32, 1000.0, AttributeAssignment.MODEL)
    register(builder, "surface_tension", float32, 0.01, AttributeAssignment.MODEL)
    register(builder, "viscosity", float32, 0.005, AttributeAssignment.MODEL)


def ptr(a) -> int | None:
    if a is None:
        return None
    return a.ptr if isinstance(a, Array) else a.data_ptr()


class ParticleSolver(SolverBase):
    """Common setup: reserve the shared grid, allocate the sorted workspace."""

    _name = "Solver"

    #: neighbour-list entries reserved per particle (lists replayed by the
    #: passes that follow a grid build; overflowing particles scan the grid).
    #: 96 covers the WCSPH dam break's 33 neighbours at rest and ~40 in its
    #: splash; DFSPH and PBF set their own.  Denser inputs pass a larger
    #: ``neighbor_list_capacity`` (keyword-only, after the reference's
    #: (model, config)).
    neighbor_list_capacity = 96

    def __init__(self, model, config, *, neighbor_list_capacity: int | None = None):
        super().__init__(model=model)
        self.config = config
        if neighbor_list_capacity is not None:
            if int(neighbor_list_capacity) < 0:
                raise ValueError("neighbor_list_capacity must be non-negative")
            self.neighbor_list_capacity = int(neighbor_list_capacity)
        n = model.particle_count
        _lib.require_device(model.device)
        if n > 1:
            if model.particle_grid is None:
                model.particle_grid = HashGrid(128, 128, 128, device=model.device)
            model.particle_grid.reserve(n)
        self._particle_accel = Array.full(n, 0.0, vec3, model.device)
        self._particle_volume = Array.full(n, 0.0, float32, model.device)
        self._bounds = (float(config.bound_width), float(config.bound_height), float(config.bound_length))
        lib = _lib.load()
        cap_list = int(self.neighbor_list_capacity)
        nbytes = lib.fs_work_bytes(max(n, 1), cap_list)
        self._work_mem = torch.empty(int(nbytes), dtype=torch.uint8, device=model.device)
        self._work = _lib.FsWork()
        _lib.check(lib.fs_work_init(ctypes.byref(self._work), ctypes.c_void_p(self._work_mem.data_ptr()), nbytes,
                                    max(n, 1), cap_list, ctypes.c_void_p(self._stream())), "fs_work_init")

    @property
    def smoothing_length(self) -> float:
        return self.config.particle_length * self.config.smoothing_length_coff

    def _stream(self) -> int:
        return _lib.stream_handle(self.model.device)

    def _particle_radius(self) -> float:
        """solver_wcsph.py:451-453"""
        r = self.config.particle_radius
        return float(r if r > 0.0 else self.model.particle_max_radius)

    def _copy_passthrough_state(self, state_in, state_out) -> None:
        """solver_wcsph.py:412-420"""
        if state_in.body_count:
            copy(state_out.body_q, state_in.body_q)
            copy(state_out.body_qd, state_in.body_qd)
            copy(state_out.body_f, state_in.body_f)
        if state_in.joint_q is not None and state_out.joint_q is not None:
            copy(state_out.joint_q, state_in.joint_q)
        if state_in.joint_qd is not None and state_out.joint_qd is not None:
            copy(state_out.joint_qd, state_in.joint_qd)

    def _model_struct(self, with_eos: bool = False) -> _lib.FsModel:
        m = self.model
        sph = m.sph
        s = _lib.FsModel()
        s.n = m.particle_count
        s.num_worlds = int(m.gravity.size)
        s.flags = ptr(m.particle_flags)
        s.mask = ptr(sph.particle_mask)
        s.world = ptr(m.particle_world)
        s.mass = ptr(m.particle_mass)
        s.gravity = ptr(m.gravity)
        s.rest_density = ptr(sph.rest_density)
        s.surface_tension = ptr(sph.surface_tension)
        s.viscosity = ptr(sph.viscosity)
        if with_eos:
            s.stiffness = ptr(sph.stiffness)
            s.exponent = ptr(sph.exponent)
        return s

    def _check_step(self, state_in, state_out) -> bool:
        """Common step preamble; False when there is nothing to do."""
        if state_in is state_out:
            raise ValueError(f"{self._name}.step requires distinct state_in and state_out")
        self._copy_passthrough_state(state_in, state_out)
        if self.model.particle_count == 0:
            return False
        if self.model.particle_grid is None:
            raise RuntimeError(f"{self._name} requires model.particle_grid to be available for particle "
                               "neighborhoods.")
        return True

    def update_contacts(self, contacts) -> None:
        del contacts

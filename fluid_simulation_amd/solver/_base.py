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
    #: splash; DFSPH and PBF set their own.  With this default the lists grow
    #: when a recording pass finds a longer list (``grow_lists``); an explicit
    #: ``neighbor_list_capacity`` (keyword-only, after the reference's
    #: (model, config)) stays fixed.
    neighbor_list_capacity = 96
    #: the neighbour lists may take at most this fraction of the device memory
    list_memory_fraction = 0.25

    def __init__(self, model, config, *, neighbor_list_capacity: int | None = None):
        super().__init__(model=model)
        self.config = config
        self.adaptive_lists = neighbor_list_capacity is None
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
        self._need_host = torch.zeros(1, dtype=torch.int32, pin_memory=True)
        self._need_event = None
        self._alloc_work(int(self.neighbor_list_capacity))

    def _alloc_work(self, cap_list: int) -> None:
        """(Re)initialise the sorted workspace with ``cap_list`` list entries
        per particle (per-step scratch only: nothing in it outlives a step)."""
        n = self.model.particle_count
        lib = _lib.load()
        nbytes = lib.fs_work_bytes(max(n, 1), cap_list)
        self._work_mem = None
        self._work_mem = torch.empty(int(nbytes), dtype=torch.uint8, device=self.model.device)
        self._work = _lib.FsWork()
        _lib.check(lib.fs_work_init(ctypes.byref(self._work), ctypes.c_void_p(self._work_mem.data_ptr()), nbytes,
                                    max(n, 1), cap_list, ctypes.c_void_p(self._stream())), "fs_work_init")
        self.neighbor_list_capacity = cap_list
        off = int(self._work.control) - self._work_mem.data_ptr() + 4 * _lib.FS_CONTROL_LIST_NEED
        self._need_dev = self._work_mem[off:off + 4].view(torch.int32)
        self._need_event = None

    # ---- neighbour-list capacity ---------------------------------------------
    def list_need(self) -> int:
        """The longest neighbour list a particle needed whose list overflowed
        in the last step (0 if none did); synchronises."""
        return int(self._need_dev.item())

    def grow_lists(self, need: int) -> bool:
        """Re-initialise the workspace with lists of at least ``need`` entries
        (1.25 need rounded up to 64, at least twice the current capacity,
        within ``list_memory_fraction`` of the device memory).  Results do not
        depend on the capacity; overflowing particles only scan the grid in
        every pass.  Returns whether the capacity changed."""
        cap = int(self.neighbor_list_capacity)
        n = max(self.model.particle_count, 1)
        total = torch.cuda.get_device_properties(self.model.device).total_memory
        limit = int(total * self.list_memory_fraction) // (4 * n) // 64 * 64
        target = min(max(2 * cap, -(-int(need) * 5 // 4 // 64) * 64), limit)
        if target <= cap:
            return False
        self._alloc_work(target)
        return True

    def adapt_lists(self) -> bool:
        """Grow the lists if the last step overflowed them (default capacity
        only); synchronises.  For drivers that replay captured steps, between
        replays; returns whether the workspace changed."""
        if not self.adaptive_lists or self.neighbor_list_capacity <= 0:
            return False
        need = self.list_need()
        return need > self.neighbor_list_capacity and self.grow_lists(need)

    def _lists_before_step(self) -> None:
        """Grow the lists if an earlier step's readback (queued without a
        sync) reports an overflow; nothing under graph capture."""
        ev = self._need_event
        if ev is None or torch.cuda.is_current_stream_capturing() or not ev.query():
            return
        self._need_event = None
        need = int(self._need_host[0])
        if need > self.neighbor_list_capacity:
            self.grow_lists(need)

    def _lists_after_step(self) -> None:
        """Queue the copy of this step's overflow word to pinned host memory
        (read by a later step once it has landed); nothing under capture."""
        if (not self.adaptive_lists or self.neighbor_list_capacity <= 0 or self._need_event is not None
                or torch.cuda.is_current_stream_capturing()):
            return
        self._need_host.copy_(self._need_dev, non_blocking=True)
        ev = torch.cuda.Event()
        ev.record()
        self._need_event = ev

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
        if self.adaptive_lists:
            self._lists_before_step()
        return True

    def update_contacts(self, contacts) -> None:
        del contacts

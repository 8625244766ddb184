"""Newton-compatible containers for the particle path.

The reference's solvers sit behind Newton's ``SolverBase`` and read Newton's
``Model``/``State`` arrays plus ``sph:`` custom attributes.  Newton is not part
of this stack, so this module supplies the same surface over torch tensors:

* ``ModelBuilder`` -- ``add_custom_attribute``, ``add_particles``,
  ``add_ground_plane`` (accepted, no particle effect), ``finalize``
* ``Model`` -- ``particle_count``, ``particle_q/qd`` (initial state),
  ``particle_mass``, ``particle_radius``, ``particle_flags``,
  ``particle_world``, ``gravity`` (one vec3 per world), ``particle_max_radius``,
  ``particle_grid``, ``set_gravity``, ``state()``, ``control()``,
  ``Model.AttributeFrequency`` / ``Model.AttributeAssignment``, and the
  namespaced MODEL attributes (``model.sph.rest_density`` ...)
* ``State`` -- ``particle_q``, ``particle_qd``, ``particle_f``, ``body_count``,
  ``joint_q``/``joint_qd`` and the namespaced STATE attributes
* ``SolverBase``, ``ParticleFlags``, ``Axis``

Arrays are :class:`fluid_simulation_amd.array.Array` (Warp-array-shaped).
"""
from __future__ import annotations

import enum
from dataclasses import dataclass
from types import SimpleNamespace
from typing import Any

import numpy as np
import torch

from .array import Array, dtype_token, float32, int32, vec3


class ParticleFlags(enum.IntFlag):
    """Newton ``ParticleFlags``: ACTIVE is bit 0."""
    ACTIVE = 1


class Axis(enum.IntEnum):
    X = 0
    Y = 1
    Z = 2


class AttributeFrequency(enum.Enum):
    PARTICLE = "particle"


class AttributeAssignment(enum.Enum):
    MODEL = "model"
    STATE = "state"


class _Namespace(SimpleNamespace):
    def names(self):
        return list(vars(self))


class Control:
    """Placeholder for Newton's Control (the particle solvers ignore it)."""


class Contacts:
    """Placeholder for Newton's Contacts (the particle solvers ignore it)."""


class State:
    """Per-step mutable arrays (Newton State)."""

    def __init__(self):
        self.particle_q: Array | None = None
        self.particle_qd: Array | None = None
        self.particle_f: Array | None = None
        self.body_count = 0
        self.body_q = self.body_qd = self.body_f = None
        self.joint_q = None
        self.joint_qd = None

    @property
    def particle_count(self):
        return 0 if self.particle_q is None else self.particle_q.size


class Model:
    """Static particle arrays and attribute specs (Newton Model)."""

    AttributeFrequency = AttributeFrequency
    AttributeAssignment = AttributeAssignment

    def __init__(self, device):
        self.device = torch.device(device)
        self.particle_count = 0
        self.particle_q = self.particle_qd = None
        self.particle_mass = self.particle_radius = None
        self.particle_flags = self.particle_world = None
        self.particle_max_radius = 0.0
        self.gravity: Array | None = None
        self.num_worlds = 1
        self.up_axis = Axis.Z
        self.particle_grid = None
        self.body_count = 0
        self._attr_specs: list = []
        self._attr_values: dict = {}

    def set_gravity(self, gravity, world: int | None = None) -> None:
        g = np.asarray(gravity, dtype=np.float32).reshape(3)
        t = self.gravity.tensor
        if world is None:
            t[:] = torch.as_tensor(g, device=t.device)
        else:
            t[world] = torch.as_tensor(g, device=t.device)

    def _attribute_array(self, spec, count):
        key = f"{spec.namespace}:{spec.name}" if spec.namespace else spec.name
        values = self._attr_values.get(key)
        if values is not None:
            return values.clone()
        return Array.full(count, spec.default, spec.dtype, self.device)

    def state(self, requires_grad: bool = False) -> State:
        del requires_grad
        s = State()
        if self.particle_count:
            s.particle_q = self.particle_q.clone()
            s.particle_qd = self.particle_qd.clone()
            s.particle_f = Array.full(self.particle_count, 0.0, vec3, self.device)
        for spec in self._attr_specs:
            if spec.assignment == AttributeAssignment.STATE:
                _set_namespaced(s, spec, self._attribute_array(spec, self.particle_count))
        return s

    def control(self) -> Control:
        return Control()


def _set_namespaced(obj, spec, array):
    if spec.namespace:
        ns = getattr(obj, spec.namespace, None)
        if ns is None:
            ns = _Namespace()
            setattr(obj, spec.namespace, ns)
        setattr(ns, spec.name, array)
    else:
        setattr(obj, spec.name, array)


@dataclass
class CustomAttribute:
    """Newton ``ModelBuilder.CustomAttribute``."""
    name: str
    frequency: Any = AttributeFrequency.PARTICLE
    assignment: Any = AttributeAssignment.MODEL
    dtype: Any = float32
    default: Any = 0.0
    namespace: str | None = None

    def __post_init__(self):
        self.dtype = dtype_token(self.dtype)
        if self.frequency != AttributeFrequency.PARTICLE:
            raise ValueError("only PARTICLE-frequency custom attributes are supported")

    @property
    def key(self):
        return f"{self.namespace}:{self.name}" if self.namespace else self.name


@dataclass
class ShapeConfig:
    density: float = 1000.0
    mu: float = 0.5


class ModelBuilder:
    """Particle subset of Newton's ``ModelBuilder``."""

    CustomAttribute = CustomAttribute
    ShapeConfig = ShapeConfig

    def __init__(self, up_axis=Axis.Z, gravity: float = -9.81):
        self.up_axis = Axis(up_axis)
        g = [0.0, 0.0, 0.0]
        g[int(self.up_axis)] = float(gravity)
        self.gravity = tuple(g)
        self.default_particle_radius = 0.1
        # per-particle columns kept as lists of numpy chunks
        self._cols: dict[str, list] = {k: [] for k in ("pos", "vel", "mass", "radius", "flags", "world")}
        self._count = 0
        self._attrs: dict[str, CustomAttribute] = {}
        self._attr_chunks: dict[str, list] = {}
        self.ground_planes: list = []

    @property
    def particle_count(self):
        return self._count

    def add_custom_attribute(self, attribute: CustomAttribute) -> None:
        if attribute.key in self._attrs:
            old = self._attrs[attribute.key]
            if (old.dtype, old.assignment) != (attribute.dtype, attribute.assignment):
                raise ValueError(f"custom attribute {attribute.key} registered twice with different specs")
            return
        self._attrs[attribute.key] = attribute

    def add_particles(self, pos, vel, mass, radius=None, flags=None, custom_attributes=None, world: int = -1):
        pos = np.asarray(pos, dtype=np.float32).reshape(-1, 3)
        n = pos.shape[0]
        vel = np.asarray(vel, dtype=np.float32).reshape(n, 3)
        mass = np.asarray(mass, dtype=np.float32).reshape(n)
        radius = np.full(n, self.default_particle_radius, np.float32) if radius is None else \
            np.asarray(radius, dtype=np.float32).reshape(n)
        flags = np.full(n, int(ParticleFlags.ACTIVE), np.int32) if flags is None else \
            np.asarray(flags, dtype=np.int32).reshape(n)
        start = self._count
        for key, values in (custom_attributes or {}).items():
            if key not in self._attrs:
                raise KeyError(f"custom attribute {key!r} is not registered")
            spec = self._attrs[key]
            vals = np.asarray(values, dtype=np.int32 if spec.dtype == int32 else np.float32)
            if spec.dtype == vec3:
                vals = vals.reshape(-1, 3)
            if vals.shape[0] != n:
                raise ValueError(f"custom attribute {key!r}: {vals.shape[0]} values for {n} particles")
            chunks = self._attr_chunks.setdefault(key, [])
            have = sum(c.shape[0] for c in chunks)
            if have < start:
                chunks.append(_default_block(spec, start - have))
            chunks.append(vals)
        for k, v in (("pos", pos), ("vel", vel), ("mass", mass), ("radius", radius), ("flags", flags),
                     ("world", np.full(n, int(world), np.int32))):
            self._cols[k].append(v)
        self._count += n

    def add_particle(self, pos, vel, mass, radius=None, flags=None, custom_attributes=None):
        attrs = {k: [v] for k, v in (custom_attributes or {}).items()}
        self.add_particles([pos], [vel], [mass], None if radius is None else [radius],
                           None if flags is None else [flags], attrs)

    def add_ground_plane(self, cfg: ShapeConfig | None = None, **kwargs):
        self.ground_planes.append((cfg, kwargs))
        return len(self.ground_planes) - 1

    def finalize(self, device="cuda:0", requires_grad: bool = False) -> Model:
        del requires_grad
        from .hashgrid import HashGrid

        m = Model(device)
        n = self._count
        m.particle_count = n
        m.up_axis = self.up_axis
        dev = m.device

        def col(k, shape, dt):
            return np.concatenate(self._cols[k]) if self._cols[k] else np.zeros(shape, dt)

        m.particle_q = Array.from_values(col("pos", (0, 3), np.float32), vec3, dev)
        m.particle_qd = Array.from_values(col("vel", (0, 3), np.float32), vec3, dev)
        m.particle_mass = Array.from_values(col("mass", (0,), np.float32), float32, dev)
        radius = col("radius", (0,), np.float32)
        m.particle_radius = Array.from_values(radius, float32, dev)
        m.particle_flags = Array.from_values(col("flags", (0,), np.int32), int32, dev)
        m.particle_world = Array.from_values(col("world", (0,), np.int32), int32, dev)
        m.particle_max_radius = float(radius.max()) if n else 0.0
        m.gravity = Array.from_values(np.asarray([self.gravity], np.float32), vec3, dev)
        m.num_worlds = 1
        m._attr_specs = list(self._attrs.values())
        for spec in m._attr_specs:
            chunks = self._attr_chunks.get(spec.key)
            if chunks is not None:
                have = sum(c.shape[0] for c in chunks)
                if have < n:
                    chunks = chunks + [_default_block(spec, n - have)]
                m._attr_values[spec.key] = Array.from_values(np.concatenate(chunks), spec.dtype, dev, count=n)
            if spec.assignment == AttributeAssignment.MODEL:
                _set_namespaced(m, spec, m._attribute_array(spec, n))
        m.particle_grid = HashGrid(128, 128, 128, device=dev) if n > 1 else None
        return m


def _default_block(spec, count):
    if spec.dtype == vec3:
        return np.tile(np.asarray(spec.default, np.float32).reshape(1, 3), (count, 1))
    return np.full(count, spec.default, np.int32 if spec.dtype == int32 else np.float32)


class SolverBase:
    """Newton ``SolverBase``: holds the model; subclasses implement ``step``."""

    def __init__(self, model: Model):
        self.model = model

    @property
    def device(self):
        return self.model.device

    def step(self, state_in: State, state_out: State, control, contacts, dt: float) -> None:
        raise NotImplementedError

    def update_contacts(self, contacts) -> None:
        del contacts

"""Batched WCSPH: many independent dam-break scenes advanced together.

The reference solves one scene per ``SolverWCSPH`` (``solver_wcsph.py:308-538``);
a training loop over many scenes calls it once per scene.  Here every scene of
a batch lives in one set of concatenated particle arrays and one segmented hash
grid, so a substep of all scenes is one grid build plus the two neighbour
passes -- six launches whatever the number of scenes -- and the per-scene CFL
substep (``simulation.py:166-187``) is computed on the device, so a sequence
of substeps captures into one HIP graph with no host synchronisation.

Scene ``s`` owns particles ``[offset_s, offset_s + N_s)`` and buckets
``[base_s, base_s + dx*dy*dz)`` of the segmented table, where each per-axis
bucket count is 128 (the plain grid's) or at least the scene's cell span + 4.
Either way two cells that a query can meet alias in the segmented table
exactly when they alias in the plain 128^3 table, so candidate sets and visit
order -- and therefore every float result -- equal those of a plain solve of
the scene with the same dt.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .examples.dam_break.scene import ParticleScene, SceneSpec, spawn_wave_tank
from .hashgrid import HashGrid
from .newton_compat import ParticleFlags

PLAIN_GRID_DIM = 128


def make_batch_scenes(num_scenes: int = 64, seed: int = 3, count_range=(5_000, 200_000),
                      length_range=(1.2, 3.0), fill_range=(0.2, 0.5)) -> list[ParticleScene]:
    """Dam-break scenes of uneven size: per scene, in this draw order, a
    particle target log-uniform in ``count_range``, a tank length uniform in
    ``length_range`` (centred on x = 0; y and z as the reference tank) and a
    column fill fraction uniform in ``fill_range`` (the reference's
    ``fill_x_frac``), spawned by the reference's wave-tank rule
    (``scene.py:44-75``)."""
    rng = np.random.default_rng(seed)
    scenes = []
    lo_c, hi_c = math.log(count_range[0]), math.log(count_range[1])
    for _ in range(int(num_scenes)):
        target = int(round(math.exp(rng.uniform(lo_c, hi_c))))
        half = 0.5 * float(rng.uniform(*length_range))
        fill = float(rng.uniform(*fill_range))
        spec = SceneSpec(bound_lo=(-half, -0.28, 0.0), bound_hi=(half, 0.28, 1.20), fill_x_frac=fill,
                         target_count=target)
        scenes.append(ParticleScene(spec, *spawn_wave_tank(spec.bound_lo, spec.bound_hi, spec.target_count,
                                                           spec.fill_x_frac, spec.fill_z_frac, spec.margin,
                                                           spec.density)))
    return scenes


@dataclass
class SceneParams:
    """Per-scene solver parameters, as the dam-break driver derives them
    (``simulation.py:110-132``): h = 2 d, radius d/2, tank half-extents."""
    smoothing_length: float
    particle_radius: float
    bound: tuple
    boundary_damping: float = 0.0
    clamp_negative_pressure: bool = True

    @classmethod
    def from_scene(cls, scene: ParticleScene, smoothing_length_coff: float = 2.0):
        hi = scene.spec.bound_hi
        return cls(smoothing_length=scene.spacing * smoothing_length_coff, particle_radius=scene.radius,
                   bound=(hi[0], hi[1], hi[2] * 0.5))


def _cell(x: np.float32, inv: np.float32) -> int:
    return int(math.floor(float(np.float32(x) * inv)))


def scene_dims(params: SceneParams) -> tuple:
    """Per-axis bucket counts of a scene's segment: min(128, span + 4), span =
    the cells covered by the clamp box [-W,W] x [-H,H] x [0,2L] that the
    solver's bounds keep every particle in.  A query reaches at most two cells
    beyond the box on each side, so any two cells it compares differ by less
    than span + 4 along an axis and only alias when equal -- as on the plain
    128^3 grid when span + 4 <= 128."""
    h = np.float32(params.smoothing_length)
    inv = np.float32(1.0) / h
    w, hh, ll = (np.float32(b) for b in params.bound)
    box = ((-w, w), (-hh, hh), (np.float32(0.0), np.float32(2.0) * ll))
    dims = []
    for lo, hi in box:
        span = _cell(hi, inv) - _cell(lo, inv) + 1
        dims.append(min(PLAIN_GRID_DIM, span + 4))
    return tuple(dims)


class BatchedWCSPH:
    """WCSPH over a batch of scenes sharing one material (the dam break's:
    B = 250000, gamma = 7, nu = 1e-6, no surface tension; ``simulation.py:86-92``).

    ``x``/``v`` hold two buffers each (``cur`` indexes the live one); ``rho``,
    ``pressure`` and ``accel`` are per particle, in scene order.  ``substep()``
    computes every scene's CFL dt on the device and advances all scenes;
    ``set_dt()`` fixes the dts instead (scenes with dt 0 stay where they are).
    """

    list_capacity = 96

    def __init__(self, scenes: list[ParticleScene], device="cuda:0", *, stiffness=250000.0, exponent=7.0,
                 viscosity=1.0e-6, surface_tension=0.0, substeps=4, params: list[SceneParams] | None = None):
        _lib.require_device(device)
        if not scenes:
            raise ValueError("a batch needs at least one scene")
        self.device = torch.device(device)
        self.scenes = list(scenes)
        self.params = params or [SceneParams.from_scene(s) for s in self.scenes]
        self.material = dict(stiffness=float(stiffness), exponent=float(exponent), viscosity=float(viscosity),
                             surface_tension=float(surface_tension))
        self.substeps = int(substeps)
        self.frame_dt = 1.0 / self.scenes[0].spec.fps
        counts = np.array([s.count for s in self.scenes], dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        n = int(self.offsets[-1])
        if n >= 2 ** 31:
            raise ValueError("batch exceeds 2^31 particles")
        self.n = n
        S = len(self.scenes)
        dev = self.device

        def dev_f32(a):
            return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)

        scene_of = np.repeat(np.arange(S, dtype=np.int32), counts)
        x0 = np.concatenate([s.positions for s in self.scenes]).astype(np.float32)
        mass = np.repeat(np.array([np.float32(s.mass) for s in self.scenes], np.float32), counts)
        rest = np.repeat(np.array([np.float32(s.spec.density) for s in self.scenes], np.float32), counts)
        self.x = [dev_f32(x0), torch.empty((n, 3), dtype=torch.float32, device=dev)]
        self.v = [torch.zeros((n, 3), dtype=torch.float32, device=dev),
                  torch.empty((n, 3), dtype=torch.float32, device=dev)]
        self.cur = 0
        self.rho = torch.zeros(n, dtype=torch.float32, device=dev)
        self.pressure = torch.zeros(n, dtype=torch.float32, device=dev)
        self.accel = torch.zeros((n, 3), dtype=torch.float32, device=dev)
        self.scene_of = torch.from_numpy(scene_of).to(dev)
        self.scene_sorted = torch.empty(n, dtype=torch.int32, device=dev)
        self.mass = dev_f32(mass)
        self.rest_density = dev_f32(rest)
        self.flags = torch.full((n,), int(ParticleFlags.ACTIVE), dtype=torch.int32, device=dev)
        self.mask = torch.zeros(n, dtype=torch.int32, device=dev)
        self.world = torch.from_numpy(scene_of).to(dev)          # world s = scene s: per-scene gravity
        self.gravity = dev_f32(np.array([s.spec.gravity for s in self.scenes], np.float32))
        m = self.material
        self.stiffness = torch.full((n,), m["stiffness"], dtype=torch.float32, device=dev)
        self.exponent = torch.full((n,), m["exponent"], dtype=torch.float32, device=dev)
        self.viscosity = torch.full((n,), m["viscosity"], dtype=torch.float32, device=dev)
        self.surface_tension = torch.full((n,), m["surface_tension"], dtype=torch.float32, device=dev)

        # scene table
        table = (_lib.FsScene * S)()
        base = 0
        for k, (sc, prm) in enumerate(zip(self.scenes, self.params)):
            t = table[k]
            t.particle_offset = int(self.offsets[k])
            t.particle_count = int(counts[k])
            t.bucket_base = base
            dims = scene_dims(prm)
            t.dims[:] = dims
            h = np.float32(prm.smoothing_length)
            t.smoothing_length = float(h)
            t.cfl_smoothing_length = float(prm.smoothing_length)
            t.cell_width_inv = float(np.float32(1.0) / h)
            t.particle_radius = float(np.float32(prm.particle_radius))
            t.bound[:] = [float(np.float32(b)) for b in prm.bound]
            t.boundary_damping = float(prm.boundary_damping)
            t.dt = 0.0
            t.clamp_negative_pressure = 1 if prm.clamp_negative_pressure else 0
            base += dims[0] * dims[1] * dims[2]
        self.num_buckets = base
        self._table_host = table
        self.table = torch.frombuffer(bytearray(bytes(table)), dtype=torch.uint8).to(dev)

        self.grid = HashGrid(base, 1, 1, device=dev)
        self.grid.reserve(n)
        lib = _lib.load()
        nbytes = lib.fs_work_bytes(n, self.list_capacity)
        self._work_mem = torch.empty(int(nbytes), dtype=torch.uint8, device=dev)
        self._work = _lib.FsWork()
        _lib.check(lib.fs_work_init(ctypes.byref(self._work), ctypes.c_void_p(self._work_mem.data_ptr()), nbytes, n,
                                    self.list_capacity, ctypes.c_void_p(self._stream())), "fs_work_init")
        self._model = _lib.FsModel()
        mm = self._model
        mm.n = n
        mm.num_worlds = S
        for field, t in (("flags", self.flags), ("mask", self.mask), ("world", self.world), ("mass", self.mass),
                         ("gravity", self.gravity), ("rest_density", self.rest_density),
                         ("surface_tension", self.surface_tension), ("viscosity", self.viscosity),
                         ("stiffness", self.stiffness), ("exponent", self.exponent)):
            setattr(mm, field, t.data_ptr())
        self.remaining = torch.zeros(S, dtype=torch.float64, device=dev)
        self._cfl_scratch = torch.zeros(2 * S, dtype=torch.int32, device=dev)
        self.status = torch.zeros(1, dtype=torch.int32, device=dev)
        self.total_substeps = 0

    # ---- helpers -------------------------------------------------------------
    @property
    def num_scenes(self) -> int:
        return len(self.scenes)

    def _stream(self) -> int:
        return _lib.stream_handle(self.device)

    def scene_slice(self, k: int) -> slice:
        return slice(int(self.offsets[k]), int(self.offsets[k + 1]))

    @property
    def positions(self) -> torch.Tensor:
        return self.x[self.cur]

    @property
    def velocities(self) -> torch.Tensor:
        return self.v[self.cur]

    def scene_table(self) -> list[dict]:
        """The device scene table read back (synchronises)."""
        raw = self.table.cpu().numpy().tobytes()
        table = (_lib.FsScene * self.num_scenes).from_buffer_copy(raw)
        return [{f: (list(getattr(t, f)) if f in ("dims", "bound") else getattr(t, f)) for f, _ in t._fields_}
                for t in table]

    def dts(self) -> np.ndarray:
        """float32 dt of every scene's last substep (synchronises)."""
        return np.array([t["dt"] for t in self.scene_table()], dtype=np.float32)

    def set_dt(self, dts) -> None:
        """Write per-scene dts into the device table (host copy, not graph-safe)."""
        dts = np.broadcast_to(np.asarray(dts, dtype=np.float32), (self.num_scenes,))
        for k in range(self.num_scenes):
            self._table_host[k].dt = float(dts[k])
        self.table.copy_(torch.frombuffer(bytearray(bytes(self._table_host)), dtype=torch.uint8))

    # ---- stepping ------------------------------------------------------------
    def begin_frame(self) -> None:
        """Every scene has one display frame of time left (simulation.py:155-165)."""
        self.remaining.fill_(self.frame_dt)

    def cfl(self) -> None:
        """Each scene's next dt from its own CFL limit and remaining frame time,
        on the device (no host sync).  The first substep of a run has no
        acceleration yet (simulation.py:142-143)."""
        m = self.material
        v = self.v[self.cur]
        spec = self.scenes[0].spec
        _lib.check(_lib.load().fs_batch_cfl_dt(
            ctypes.c_void_p(self.table.data_ptr()), self.num_scenes, int(max(s.count for s in self.scenes)),
            ctypes.c_void_p(v.data_ptr()),
            ctypes.c_void_p(self.accel.data_ptr()), 1 if self.total_substeps else 0, m["stiffness"], m["exponent"],
            float(spec.density), m["viscosity"], self.frame_dt, self.substeps,
            ctypes.c_void_p(self.remaining.data_ptr()), ctypes.c_void_p(self._cfl_scratch.data_ptr()),
            ctypes.c_void_p(self.status.data_ptr()), ctypes.c_void_p(self._stream())), "fs_batch_cfl_dt")

    def advance(self) -> None:
        """One WCSPH substep of every scene with the dts in the scene table."""
        src, dst = self.cur, 1 - self.cur
        io = _lib.FsBatchIO()
        io.x_in = self.x[src].data_ptr()
        io.v_in = self.v[src].data_ptr()
        io.x_out = self.x[dst].data_ptr()
        io.v_out = self.v[dst].data_ptr()
        io.rho = self.rho.data_ptr()
        io.pressure = self.pressure.data_ptr()
        io.accel = self.accel.data_ptr()
        io.scene_of = self.scene_of.data_ptr()
        io.scene_sorted = self.scene_sorted.data_ptr()
        _lib.check(_lib.load().fs_batch_wcsph_step(ctypes.byref(self.grid._c), ctypes.byref(self._work),
                                                   ctypes.byref(self._model), ctypes.c_void_p(self.table.data_ptr()),
                                                   self.num_scenes, ctypes.byref(io),
                                                   ctypes.c_void_p(self._stream())), "fs_batch_wcsph_step")
        self.cur = dst
        self.total_substeps += 1

    def substep(self) -> None:
        self.cfl()
        self.advance()

    def autotune(self, use_cache: bool = True) -> dict:
        """Pick the fastest launch variant of the batched density and force
        passes on this batch (results are identical under every variant);
        the simulation state is restored afterwards."""
        from .autotune import autotune

        cur, total = self.cur, self.total_substeps
        x, v = self.x[cur].clone(), self.v[cur].clone()
        rho, p, a = self.rho.clone(), self.pressure.clone(), self.accel.clone()
        dts = self.dts()
        if not np.any(dts > 0):
            self.set_dt(np.full(self.num_scenes, 1.0e-5, np.float32))
        res = autotune(self.advance, ("batch_density", "batch_force"), f"batched-{self.num_scenes}-{self.n}",
                       use_cache=use_cache)
        self.cur, self.total_substeps = cur, total
        self.x[cur].copy_(x)
        self.v[cur].copy_(v)
        self.rho.copy_(rho)
        self.pressure.copy_(p)
        self.accel.copy_(a)
        self.set_dt(dts)
        return res

    def check_status(self) -> None:
        """Raise the reference driver's errors for what the device CFL flagged
        (simulation.py:148-153, :182-184); synchronises."""
        code = int(self.status.item())
        if code == 1:
            raise RuntimeError("Non-finite velocity or WCSPH acceleration in a batched scene")
        if code == 2:
            raise RuntimeError("CFL step cannot safely complete this frame")

    def frame_done(self) -> bool:
        """True when every scene has used up its frame time (synchronises)."""
        return not bool((self.remaining > 1.0e-12).any().item())

    def step_frame(self, chunk: int = 8, max_substeps: int = 4096) -> int:
        """Advance every scene by one display frame; each scene takes its own
        CFL substeps and idles (dt = 0) once its frame time is used.  The host
        checks for completion every ``chunk`` substeps.  Returns the number of
        batched substeps taken."""
        self.begin_frame()
        taken = 0
        while True:
            for _ in range(chunk):
                self.substep()
            taken += chunk
            self.check_status()
            if self.frame_done():
                return taken
            if taken >= max_substeps:
                raise RuntimeError("CFL step cannot safely complete this frame")

"""Device-resident dam-break frames: no host synchronisation per substep.

The reference driver (``simulation.py:140-190``) reads the CFL maxima back to
the host before every WCSPH/DFSPH substep.  Here:

* WCSPH keeps the plain solver but takes its dt from device memory.  The
  driver's CFL policy runs on the device (``fs_batch_cfl_dt`` on a one-scene
  table, in float64 as the host formula).  It writes the substep's dt and the
  frame time still remaining.  A chunk of substeps replays from one HIP
  graph.  Once the frame time is used up, the rest of the chunk runs with
  dt = 0, which leaves the state, density, pressure and acceleration as they
  are.  The host reads one flag per chunk to learn whether the frame is
  complete.
* DFSPH does the same with its own policy (``fs_dfsph_cfl_dt``: frame/4,
  lowered to 0.4 spacing / vmax).  Its passes read dt from device memory too,
  and a substep with dt = 0 is idle: it copies the whole input state (x, v,
  and the ``sph`` arrays) to the output.  Its substeps are long and few per
  frame, so its chunks are short (4 substeps).
* PBF has a fixed substep (frame/substeps), so a whole frame is one graph.

Before each frame the driver lets the solver grow its neighbour lists if the
last replayed substep overflowed them (``adapt_lists``), and captures the
graph again when it did: the graph holds the workspace's addresses.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ... import _lib
from .simulation import DamBreakSimulation


class DeviceResidentDamBreak:
    def __init__(self, scene, method: str = "wcsph", device="cuda:0", substeps: int = 4, chunk: int | None = None,
                 **sim_kwargs):
        """``sim_kwargs`` go to ``DamBreakSimulation`` (the WCSPH material and
        pressure-clamp options).  ``chunk``: substeps per graph replay
        (default 16 for WCSPH, 4 for DFSPH)."""
        if method not in ("wcsph", "dfsph", "pbf"):
            raise ValueError(f"unknown method {method!r}")
        if chunk is None:
            chunk = 4 if method == "dfsph" else 16
        if chunk < 2 or chunk % 2:
            raise ValueError("chunk must be even (the substeps swap two state buffers)")
        self.method = method
        self.scene = scene
        self.device = torch.device(device)
        self.frame = 0
        self.substeps = int(substeps)
        self.chunk = int(chunk)
        self.total_substeps = 0
        self._graph = None
        self.sim = DamBreakSimulation(scene, method, device=device, substeps=substeps, **sim_kwargs)
        self.frame_dt = self.sim.frame_dt
        if method in ("wcsph", "dfsph"):
            self._cfl_setup()

    # ---- state views ----------------------------------------------------------
    @property
    def positions(self) -> torch.Tensor:
        return self.sim.state_0.particle_q.tensor

    @property
    def velocities(self) -> torch.Tensor:
        return self.sim.state_0.particle_qd.tensor

    @property
    def density(self) -> torch.Tensor:
        return self.sim.state_0.sph.rho.tensor

    # ---- device CFL --------------------------------------------------------------
    def _cfl_setup(self):
        sol = self.sim.solver
        t = _lib.FsScene()
        t.particle_offset = 0
        t.particle_count = self.scene.count
        t.smoothing_length = float(np.float32(sol.smoothing_length))
        t.cfl_smoothing_length = float(sol.smoothing_length)   # the host policy's float64 h
        t.dt = 0.0
        self._table = torch.frombuffer(bytearray(bytes(t)), dtype=torch.uint8).to(self.device)
        off = _lib.FsScene.dt.offset
        self.dt = self._table[off:off + 4].view(torch.float32)   # the substep's dt, on the device
        self.remaining = torch.zeros(1, dtype=torch.float64, device=self.device)
        self._scratch = torch.zeros(2, dtype=torch.int32, device=self.device)
        self.status = torch.zeros(1, dtype=torch.int32, device=self.device)

    def _cfl(self):
        sim = self.sim
        mat = sim.material
        if self.method == "dfsph":
            _lib.check(_lib.load().fs_dfsph_cfl_dt(
                ctypes.c_void_p(self._table.data_ptr()), 1, self.scene.count,
                ctypes.c_void_p(sim.state_0.particle_qd.ptr), float(self.scene.spacing), self.frame_dt,
                self.substeps, ctypes.c_void_p(self.remaining.data_ptr()),
                ctypes.c_void_p(self._scratch.data_ptr()), ctypes.c_void_p(self.status.data_ptr()),
                ctypes.c_void_p(_lib.stream_handle(self.device))), "fs_dfsph_cfl_dt")
            return
        _lib.check(_lib.load().fs_batch_cfl_dt(
            ctypes.c_void_p(self._table.data_ptr()), 1, self.scene.count,
            ctypes.c_void_p(sim.state_0.particle_qd.ptr), ctypes.c_void_p(sim.solver._particle_accel.ptr),
            1 if sim.total_substeps else 0, float(mat["stiffness"]), float(mat["exponent"]),
            float(self.scene.spec.density), float(mat["viscosity"]), self.frame_dt, self.substeps,
            ctypes.c_void_p(self.remaining.data_ptr()), ctypes.c_void_p(self._scratch.data_ptr()),
            ctypes.c_void_p(self.status.data_ptr()), ctypes.c_void_p(_lib.stream_handle(self.device))),
            "fs_batch_cfl_dt")

    def _cfl_substep(self):
        self._cfl()
        self.sim._substep(self.dt)

    def _cfl_chunk(self):
        for _ in range(self.chunk):
            self._cfl_substep()

    def _pbf_frame(self):
        dt = self.frame_dt / self.substeps
        for _ in range(self.substeps):
            self.sim._substep(dt)

    # ---- capture -----------------------------------------------------------------
    def _capture(self, work):
        stream = torch.cuda.Stream(self.device)
        stream.wait_stream(torch.cuda.current_stream(self.device))
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.stream(stream):
            with torch.cuda.graph(graph, stream=stream):
                work()
        torch.cuda.current_stream(self.device).wait_stream(stream)
        return graph

    # ---- stepping ----------------------------------------------------------------
    def step(self) -> int:
        """Advance one display frame; returns the substeps run (WCSPH: the
        substeps replayed, including idle ones after the frame completed)."""
        sim = self.sim
        if sim.solver.adapt_lists():
            self._graph = None                 # the graph holds the old workspace's lists
        if self.method == "pbf":
            if self.substeps % 2:
                self._pbf_frame()              # odd substep counts: eager frames
            else:
                if self._graph is None:
                    total = sim.total_substeps
                    self._graph = self._capture(self._pbf_frame)
                    sim.total_substeps = total
                self._graph.replay()
                sim.total_substeps += self.substeps
            self._end_frame(self.substeps)
            return self.substeps
        self.remaining.fill_(self.frame_dt)
        taken = 0
        if sim.total_substeps == 0 and self.method == "wcsph":
            self._cfl_substep()                # the first substep has no acceleration term yet
            taken += 1
        if self._graph is None:
            total = sim.total_substeps
            self._graph = self._capture(self._cfl_chunk)
            sim.total_substeps = total             # capture only recorded the chunk
        while bool((self.remaining > 1.0e-12).item()):
            self._graph.replay()
            sim.total_substeps += self.chunk
            taken += self.chunk
            code = int(self.status.item())
            if code == 1:
                raise RuntimeError("Non-finite velocity or WCSPH acceleration" if self.method == "wcsph"
                                   else f"Non-finite velocity at frame {self.frame}")
            if code == 2 or taken > 4096:
                raise RuntimeError("CFL step cannot safely complete this frame")
        self._end_frame(taken)
        return taken

    def _end_frame(self, taken: int):
        sim = self.sim
        self.frame += 1
        self.total_substeps += taken
        sim.frame = self.frame
        sim.sim_time = self.frame * self.frame_dt
        sim.last_substeps = taken

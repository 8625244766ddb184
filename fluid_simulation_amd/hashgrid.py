"""Device hash grid with Warp ``HashGrid`` semantics, backed by libfluidsim_hip.

``HashGrid(dim_x, dim_y, dim_z)`` / ``reserve(n)`` / ``build(points, radius)``
mirror the Warp calls the reference makes (``solver_wcsph.py:398-402,456``).
After a build, ``point_ids`` (sorted slot -> original index), ``point_cells``
(sorted bucket keys) and the bucket prefix ``cell_start`` are zero-copy torch
views; ``cell_of`` (bucket of each original point) and ``cell_tables()``
(Warp-layout start/end tables) are derived from them.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .array import Array


class HashGrid:
    def __init__(self, dim_x: int = 128, dim_y: int = 128, dim_z: int = 128, device="cuda:0", round_mode: int = 0):
        self.dims = (int(dim_x), int(dim_y), int(dim_z))
        self.device = torch.device(device)
        self.round_mode = int(round_mode)
        self.capacity = 0
        self._c = _lib.FsGrid()
        self._ws: torch.Tensor | None = None

    # Warp exposes an integer handle; here it is the descriptor's address
    @property
    def id(self) -> int:
        return ctypes.addressof(self._c)

    @property
    def num_points(self) -> int:
        return int(self._c.num_points)

    def reserve(self, num_points: int) -> None:
        num_points = int(num_points)
        if num_points <= self.capacity and self._ws is not None:
            return
        _lib.require_device(self.device)
        lib = _lib.load()
        nbytes = lib.fs_grid_workspace_bytes(num_points, *self.dims)
        if nbytes == 0:
            raise ValueError("invalid grid dimensions or capacity")
        self._ws = torch.empty(int(nbytes), dtype=torch.uint8, device=self.device)
        _lib.check(lib.fs_grid_init(ctypes.byref(self._c), ctypes.c_void_p(self._ws.data_ptr()), nbytes, num_points,
                                    *self.dims, self.round_mode, ctypes.c_void_p(_lib.stream_handle(self.device))),
                   "fs_grid_init")
        self.capacity = num_points

    def build(self, points, radius: float) -> None:
        t = points.tensor if isinstance(points, Array) else points
        n = int(t.shape[0])
        if n > self.capacity:
            self.reserve(n)
        lib = _lib.load()
        _lib.check(lib.fs_grid_build(ctypes.byref(self._c), ctypes.c_void_p(t.data_ptr()), n, float(radius),
                                     ctypes.c_void_p(_lib.stream_handle(self.device))), "fs_grid_build")

    # ---- zero-copy views of the device arrays -------------------------------
    def _view(self, addr: int, count: int) -> torch.Tensor:
        base = self._ws.data_ptr()
        off = int(addr) - base
        return self._ws[off:off + 4 * count].view(torch.int32)

    @property
    def point_ids(self) -> torch.Tensor:
        return self._view(self._c.point_ids, self.num_points)

    @property
    def point_cells(self) -> torch.Tensor:
        return self._view(self._c.point_cells, self.num_points)

    @property
    def cell_of(self) -> torch.Tensor:
        """Bucket of each original point, from the sorted arrays (the build
        itself keeps buckets only in its processing order)."""
        out = torch.empty(self.num_points, dtype=torch.int32, device=self.device)
        out[self.point_ids.long()] = self.point_cells
        return out

    @property
    def cell_start(self) -> torch.Tensor:
        m = self.dims[0] * self.dims[1] * self.dims[2]
        return self._view(self._c.cell_start, m + 1)

    def cell_tables(self):
        """Warp-layout (cell_starts, cell_ends), 0/0 for empty buckets."""
        m = self.dims[0] * self.dims[1] * self.dims[2]
        starts = torch.empty(m, dtype=torch.int32, device=self.device)
        ends = torch.empty(m, dtype=torch.int32, device=self.device)
        _lib.check(_lib.load().fs_grid_export_tables(ctypes.byref(self._c), ctypes.c_void_p(starts.data_ptr()),
                                                     ctypes.c_void_p(ends.data_ptr()),
                                                     ctypes.c_void_p(_lib.stream_handle(self.device))),
                   "fs_grid_export_tables")
        return starts, ends

    def query_counts(self, points, radius: float, within: str | None = None) -> torch.Tensor:
        """Candidate count of hash_grid_query(points[i], radius) per point."""
        t = points.tensor if isinstance(points, Array) else points
        mode = {None: -1, "cubic": 0, "pbf": 1}[within]
        counts = torch.empty(self.num_points, dtype=torch.int32, device=self.device)
        _lib.check(_lib.load().fs_grid_count_candidates(ctypes.byref(self._c), ctypes.c_void_p(t.data_ptr()),
                                                        float(radius), mode, ctypes.c_void_p(counts.data_ptr()),
                                                        ctypes.c_void_p(_lib.stream_handle(self.device))),
                   "fs_grid_count_candidates")
        return counts

    def query_lists(self, points, radius: float, within: str | None = None):
        """CSR (offsets int64[n+1], indices int32) of the query candidates in
        visit order, original numbering."""
        t = points.tensor if isinstance(points, Array) else points
        counts = self.query_counts(t, radius, within)
        offsets = torch.zeros(self.num_points + 1, dtype=torch.int64, device=self.device)
        offsets[1:] = torch.cumsum(counts.to(torch.int64), 0)
        total = int(offsets[-1].item())
        indices = torch.empty(max(total, 1), dtype=torch.int32, device=self.device)
        mode = {None: -1, "cubic": 0, "pbf": 1}[within]
        _lib.check(_lib.load().fs_grid_list_candidates(ctypes.byref(self._c), ctypes.c_void_p(t.data_ptr()),
                                                       float(radius), mode, ctypes.c_void_p(offsets.data_ptr()),
                                                       ctypes.c_void_p(indices.data_ptr()),
                                                       ctypes.c_void_p(_lib.stream_handle(self.device))),
                   "fs_grid_list_candidates")
        return offsets, indices[:total]

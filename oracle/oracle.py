"""CPU oracle for the particle-fluid hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg import this module, as the checker; the product package never does.

The kernels live in ``fs_oracle.c`` (built twice: FP32 and FP64).  This module
composes them the way the reference's ``step`` methods compose their Warp
launches, line by line:

* ``wcsph_step``  -- ``solver/sph/solver_wcsph.py:431-535``
* ``dfsph_step``  -- ``solver/sph/solver_dfsph.py:598-898`` (host residuals
  ``:574-596`` evaluated with numpy exactly as the reference does)
* ``pbf_step``    -- ``solver/pbf/solver_pbf.py:676-954``
* ``cfl_dt`` / ``frame_substeps`` -- ``examples/dam_break/simulation.py:140-190``
* ``particle_separation`` -- ``solver/apic/solver_apic.py:193-268`` composed as
  ``SolverAPIC._apply_particle_separation`` (``:1003-1066``)

Every array is a numpy array in original particle order: positions and
velocities ``(N, 3)``, scalars ``(N,)``.  ``precision`` is ``"f32"`` (the
restatement the GPU path is compared with) or ``"f64"`` (the error reference).
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from dataclasses import dataclass, field, replace

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS: dict = {}

GRID_DIM = 128


def build() -> None:
    """Compile the FP32 and FP64 oracle libraries with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _lib(precision: str):
    if precision in _LIBS:
        return _LIBS[precision]
    path = os.path.join(HERE, f"libfs_oracle_{precision}.so")
    if not os.path.exists(path):
        build()
    lib = load_library(path, precision)
    _LIBS[precision] = lib
    return lib


def load_library(path: str, precision: str):
    """Load an oracle build (``fs_oracle.c`` compiled with ORACLE_REAL float
    for "f32", double for "f64") and declare its scalar signatures."""
    lib = ctypes.CDLL(path)
    real = ctypes.c_float if precision == "f32" else ctypes.c_double
    lib._real = real
    lib._np = np.float32 if precision == "f32" else np.float64

    class Grid(ctypes.Structure):
        _fields_ = [
            ("dim_x", ctypes.c_int), ("dim_y", ctypes.c_int), ("dim_z", ctypes.c_int),
            ("round_mode", ctypes.c_int), ("num_points", ctypes.c_int),
            ("cell_width", real), ("cell_width_inv", real),
            ("point_ids", ctypes.c_void_p), ("cell_starts", ctypes.c_void_p), ("cell_ends", ctypes.c_void_p),
        ]

    lib._Grid = Grid
    lib.ora_cubic_w.restype = real
    lib.ora_cubic_w.argtypes = [real, real]
    lib.ora_poly6_w.restype = real
    lib.ora_poly6_w.argtypes = [real, real]
    lib.ora_grid_candidates.restype = ctypes.c_longlong
    assert lib.ora_real_bytes() == ctypes.sizeof(real)
    return lib


def set_threads(n: int) -> None:
    for p in ("f32", "f64"):
        _lib(p).ora_set_threads(int(n))


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else ctypes.c_void_p(0)


# --------------------------------------------------------------------------
# Hash grid (Warp HashGrid semantics)
# --------------------------------------------------------------------------


class OracleGrid:
    """Warp ``HashGrid(dim, dim, dim)`` restated: ``build`` and the query order.

    Arrays after ``build``: ``cell_of`` (bucket of particle i, original order),
    ``point_ids`` (sorted slot -> original index), ``point_cells`` (sorted
    bucket keys), ``cell_starts``/``cell_ends`` (``dim^3`` tables, 0/0 when
    empty).
    """

    def __init__(self, dim_x=GRID_DIM, dim_y=GRID_DIM, dim_z=GRID_DIM, precision="f32", round_mode=0):
        self.dims = (int(dim_x), int(dim_y), int(dim_z))
        self.precision = precision
        self.round_mode = int(round_mode)
        self.lib = _lib(precision)
        m = self.dims[0] * self.dims[1] * self.dims[2]
        self.cell_starts = np.zeros(m, np.int32)
        self.cell_ends = np.zeros(m, np.int32)
        self.n = 0

    def build(self, x: np.ndarray, radius: float) -> "OracleGrid":
        lib = self.lib
        x = np.ascontiguousarray(x, dtype=lib._np)
        n = x.shape[0]
        self.n = n
        self.cell_of = np.empty(n, np.int32)
        self.point_ids = np.empty(n, np.int32)
        self.point_cells = np.empty(n, np.int32)
        self.radius = lib._np(radius)
        self.cell_width = lib._real(float(self.radius)).value
        self.cell_width_inv = float(lib._np(1.0) / self.radius)
        rc = lib.ora_grid_build(ctypes.c_int(n), _p(x), lib._real(float(self.radius)), *[ctypes.c_int(d) for d in self.dims],
                                ctypes.c_int(self.round_mode), _p(self.cell_of), _p(self.point_ids),
                                _p(self.point_cells), _p(self.cell_starts), _p(self.cell_ends))
        if rc != 0:
            raise MemoryError("oracle grid build failed")
        return self

    def struct(self):
        lib = self.lib
        return lib._Grid(self.dims[0], self.dims[1], self.dims[2], self.round_mode, self.n,
                         lib._real(self.cell_width).value, lib._real(self.cell_width_inv).value,
                         self.point_ids.ctypes.data, self.cell_starts.ctypes.data, self.cell_ends.ctypes.data)

    def candidates(self, x, radius, within=None, query_ids=None):
        """CSR (offsets int64[n+1], indices int32) of the query candidates in
        visit order.  ``within``: None = all candidates, "cubic" = q<=1,
        "pbf" = r<h (strict)."""
        lib = self.lib
        x = np.ascontiguousarray(x, dtype=lib._np)
        nq = x.shape[0] if query_ids is None else len(query_ids)
        qi = None if query_ids is None else np.ascontiguousarray(query_ids, np.int32)
        mode = {None: -1, "cubic": 0, "pbf": 1}[within]
        g = self.struct()
        off = np.zeros(nq + 1, np.int64)
        total = lib.ora_grid_candidates(ctypes.byref(g), _p(x), ctypes.c_int(nq), _p(qi), lib._real(float(radius)),
                                        ctypes.c_int(mode), _p(off), ctypes.c_void_p(0))
        idx = np.zeros(max(total, 1), np.int32)
        lib.ora_grid_candidates(ctypes.byref(g), _p(x), ctypes.c_int(nq), _p(qi), lib._real(float(radius)),
                                ctypes.c_int(mode), _p(off), _p(idx))
        return off, idx[:total]


# --------------------------------------------------------------------------
# Model / state containers
# --------------------------------------------------------------------------


@dataclass
class OModel:
    """Per-particle model arrays (Newton Model + ``sph:`` MODEL attributes)."""
    mass: np.ndarray
    flags: np.ndarray
    mask: np.ndarray
    world: np.ndarray
    gravity: np.ndarray                 # (W, 3)
    rest_density: np.ndarray
    surface_tension: np.ndarray
    viscosity: np.ndarray
    stiffness: np.ndarray | None = None
    exponent: np.ndarray | None = None
    radius: np.ndarray | None = None

    @property
    def n(self):
        return int(self.mass.shape[0])

    def astype(self, dt):
        f = lambda a: None if a is None else np.ascontiguousarray(a, dtype=dt)
        return replace(self, mass=f(self.mass), gravity=f(self.gravity), rest_density=f(self.rest_density),
                       surface_tension=f(self.surface_tension), viscosity=f(self.viscosity),
                       stiffness=f(self.stiffness), exponent=f(self.exponent), radius=f(self.radius),
                       flags=np.ascontiguousarray(self.flags, np.int32), mask=np.ascontiguousarray(self.mask, np.int32),
                       world=np.ascontiguousarray(self.world, np.int32))


@dataclass
class OState:
    x: np.ndarray
    v: np.ndarray
    sph: dict = field(default_factory=dict)

    def copy(self):
        return OState(self.x.copy(), self.v.copy(), {k: v.copy() for k, v in self.sph.items()})


def uniform_model(n, mass, *, rest_density=1000.0, surface_tension=0.0, viscosity=1.0e-6, stiffness=None,
                  exponent=None, gravity=(0.0, 0.0, -10.0), radius=None, dtype=np.float32) -> OModel:
    full = lambda v: np.full(n, v, dtype)
    return OModel(mass=full(mass), flags=np.ones(n, np.int32), mask=np.zeros(n, np.int32),
                  world=np.zeros(n, np.int32), gravity=np.asarray([gravity], dtype),
                  rest_density=full(rest_density), surface_tension=full(surface_tension), viscosity=full(viscosity),
                  stiffness=None if stiffness is None else full(stiffness),
                  exponent=None if exponent is None else full(exponent),
                  radius=None if radius is None else full(radius))


# --------------------------------------------------------------------------
# Configs (same fields and defaults as the reference dataclasses)
# --------------------------------------------------------------------------


@dataclass
class WcsphConfig:
    """SolverWCSPH.SphConfig  solver_wcsph.py:309-318"""
    particle_radius: float = 0.1
    particle_length: float = 0.2
    smoothing_length_coff: float = 1.35
    bound_width: float = 10.0
    bound_height: float = 10.0
    bound_length: float = 10.0
    boundary_damping: float = -0.1
    clamp_negative_pressure: bool = True


@dataclass
class DfsphConfig:
    """SolverDFSPH.SphConfig  solver_dfsph.py:446-463"""
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


@dataclass
class PbfConfig:
    """SolverPBF.PbfConfig  solver_pbf.py:521-539"""
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


def smoothing_length(cfg) -> float:
    """``smoothing_length`` property  solver_wcsph.py:408-410 (Python double)."""
    return cfg.particle_length * cfg.smoothing_length_coff


# --------------------------------------------------------------------------
# Solver steps
# --------------------------------------------------------------------------


class _Ctx:
    def __init__(self, model: OModel, precision: str, round_mode: int = 0):
        self.lib = _lib(precision)
        self.dt_np = self.lib._np
        self.m = model.astype(self.dt_np)
        self.precision = precision
        self.round_mode = round_mode
        self.R = self.lib._real

    def r(self, v):
        return self.R(float(v))

    def new_grid(self):
        return OracleGrid(precision=self.precision, round_mode=self.round_mode)


def _volume(c: _Ctx):
    vol = np.empty(c.m.n, c.dt_np)
    c.lib.ora_particle_volume(ctypes.c_int(c.m.n), _p(c.m.mass), _p(c.m.rest_density), _p(vol))
    return vol


def _particle_radius(cfg, m: OModel):
    """solver_wcsph.py:451-453: config radius, else the model's max radius."""
    if cfg.particle_radius > 0.0:
        return float(cfg.particle_radius)
    return float(np.max(m.radius)) if m.radius is not None and m.n else 0.0


def wcsph_step(model: OModel, state_in: OState, cfg: WcsphConfig, dt: float, precision="f32", round_mode=0):
    """SolverWCSPH.step  solver_wcsph.py:431-535.  Returns (state_out, extras)."""
    c = _Ctx(model, precision, round_mode)
    m, lib, n = c.m, c.lib, c.m.n
    x_in = np.ascontiguousarray(state_in.x, c.dt_np)
    v_in = np.ascontiguousarray(state_in.v, c.dt_np)
    out = OState(np.zeros_like(x_in), np.zeros_like(v_in),
                 {"rho": np.zeros(n, c.dt_np), "pressure": np.zeros(n, c.dt_np)})
    accel = np.zeros((n, 3), c.dt_np)
    if n == 0:
        return out, {"accel": accel}
    h = smoothing_length(cfg)
    pr = _particle_radius(cfg, m)
    vol = _volume(c)                                                         # :455
    grid = c.new_grid().build(x_in, h)                                       # :456
    g = grid.struct()
    lib.ora_wcsph_density_pressure(ctypes.byref(g), _p(m.flags), _p(x_in), _p(m.rest_density), _p(m.stiffness),
                                   _p(m.exponent), _p(vol), c.r(h), ctypes.c_int(int(cfg.clamp_negative_pressure)),
                                   _p(out.sph["rho"]), _p(out.sph["pressure"]))  # :458-474
    lib.ora_wcsph_acceleration(ctypes.byref(g), _p(m.flags), _p(m.mask), _p(x_in), _p(v_in), _p(out.sph["rho"]),
                               _p(out.sph["pressure"]), c.r(pr), _p(m.mass), _p(m.surface_tension),
                               _p(m.viscosity), _p(m.gravity), _p(m.world), c.r(h), _p(accel))  # :475-496
    lib.ora_wcsph_integrate_velocity(ctypes.c_int(n), _p(m.flags), _p(m.mask), _p(v_in), _p(accel), c.r(dt),
                                     _p(out.v))                              # :497-509
    lib.ora_integrate_positions(ctypes.c_int(n), _p(m.flags), _p(m.mask), _p(x_in), _p(out.v), c.r(dt),
                                _p(out.x))                                   # :510-522
    lib.ora_sph_apply_bounds(ctypes.c_int(n), _p(m.flags), _p(m.mask), _p(out.x), _p(out.v), c.r(cfg.bound_width),
                             c.r(cfg.bound_height), c.r(cfg.bound_length), c.r(cfg.boundary_damping),
                             ctypes.c_int(1))                                # :523-535
    return out, {"accel": accel, "grid": grid, "volume": vol}


def wcsph_passes_ids(model: OModel, x: np.ndarray, v: np.ndarray, cfg: WcsphConfig, ids, rho_full=None,
                     pressure_full=None, precision="f32", grid=None):
    """compute_density_pressure and acceleration_wcsph (solver_wcsph.py:119-228)
    of the listed particles only, on a grid built at ``x`` (or ``grid``).
    The density pass reads positions only; the acceleration reads the density
    and pressure of the neighbours, taken from ``rho_full``/``pressure_full``
    (whole arrays, e.g. another implementation's pass output) or from this
    call's own density of the listed particles when they are None.  Returns
    (rho[ids], pressure[ids], accel[ids], grid)."""
    c = _Ctx(model, precision)
    m, lib, n = c.m, c.lib, c.m.n
    f = c.dt_np
    x = np.ascontiguousarray(x, f)
    v = np.ascontiguousarray(v, f)
    ids = np.ascontiguousarray(ids, np.int32)
    h = smoothing_length(cfg)
    pr = _particle_radius(cfg, m)
    vol = _volume(c)
    grid = grid if grid is not None else c.new_grid().build(x, h)
    g = grid.struct()
    rho = np.zeros(n, f)
    p = np.zeros(n, f)
    lib.ora_wcsph_density_pressure_ids(ctypes.byref(g), ctypes.c_int(len(ids)), _p(ids), _p(m.flags), _p(x),
                                       _p(m.rest_density), _p(m.stiffness), _p(m.exponent), _p(vol), c.r(h),
                                       ctypes.c_int(int(cfg.clamp_negative_pressure)), _p(rho), _p(p))
    rho_in = rho if rho_full is None else np.ascontiguousarray(rho_full, f)
    p_in = p if pressure_full is None else np.ascontiguousarray(pressure_full, f)
    accel = np.zeros((n, 3), f)
    lib.ora_wcsph_acceleration_ids(ctypes.byref(g), ctypes.c_int(len(ids)), _p(ids), _p(m.flags), _p(m.mask), _p(x),
                                   _p(v), _p(rho_in), _p(p_in), c.r(pr), _p(m.mass), _p(m.surface_tension),
                                   _p(m.viscosity), _p(m.gravity), _p(m.world), c.r(h), _p(accel))
    return rho[ids], p[ids], accel[ids], grid


def wcsph_adjoint_step(model: OModel, state_in: OState, cfg: WcsphConfig, dt: float, xbar_out, vbar_out,
                       precision="f32", round_mode=0):
    """Reverse-mode derivative of one ``wcsph_step`` (surface tension must be
    0).  Recomputes the forward substep, then runs the three adjoint passes of
    ``fs_oracle.c``.  Returns (xbar_in, vbar_in, grad_stiffness[n],
    grad_viscosity[n]) -- dL/d(per-particle stiffness, viscosity) of this
    substep."""
    c = _Ctx(model, precision, round_mode)
    m, lib, n = c.m, c.lib, c.m.n
    if np.any(m.surface_tension != 0):
        raise ValueError("the WCSPH adjoint covers surface_tension = 0 only")
    x_in = np.ascontiguousarray(state_in.x, c.dt_np)
    v_in = np.ascontiguousarray(state_in.v, c.dt_np)
    out, ex = wcsph_step(model, state_in, cfg, dt, precision, round_mode)
    xbo = np.ascontiguousarray(xbar_out, c.dt_np)
    vbo = np.ascontiguousarray(vbar_out, c.dt_np)
    xb = np.zeros((n, 3), c.dt_np)
    vb = np.zeros((n, 3), c.dt_np)
    ab = np.zeros((n, 3), c.dt_np)
    rb = np.zeros(n, c.dt_np)
    gs = np.zeros(n, c.dt_np)
    gv = np.zeros(n, c.dt_np)
    if n == 0:
        return xb, vb, gs, gv
    h = smoothing_length(cfg)
    g = ex["grid"].struct()
    lib.ora_wcsph_adj_integrate(ctypes.c_int(n), _p(m.flags), _p(m.mask), _p(x_in), _p(v_in), _p(ex["accel"]),
                                c.r(dt), c.r(cfg.bound_width), c.r(cfg.bound_height), c.r(cfg.bound_length),
                                c.r(cfg.boundary_damping), _p(xbo), _p(vbo), _p(xb), _p(vb), _p(ab))
    lib.ora_wcsph_adj_force(ctypes.byref(g), _p(m.flags), _p(x_in), _p(v_in), _p(out.sph["rho"]),
                            _p(out.sph["pressure"]), _p(m.mass), _p(m.rest_density), _p(m.stiffness),
                            _p(m.exponent), _p(m.viscosity), c.r(h), ctypes.c_int(int(cfg.clamp_negative_pressure)),
                            _p(ab), _p(xb), _p(vb), _p(rb), _p(gs), _p(gv))
    lib.ora_wcsph_adj_density(ctypes.byref(g), _p(m.flags), _p(x_in), _p(m.rest_density), _p(ex["volume"]), c.r(h),
                              _p(rb), _p(xb))
    return xb, vb, gs, gv


def wcsph_rollout(model: OModel, x0, v0, cfg: WcsphConfig, dt: float, steps: int, precision="f32",
                  densities: list | None = None):
    """``steps`` WCSPH substeps from (x0, v0); returns the list of states
    [(x_0, v_0), ..., (x_steps, v_steps)].  With ``densities`` (a list), each
    substep's density array is appended to it."""
    st = OState(np.array(x0), np.array(v0), {})
    states = [(st.x, st.v)]
    for _ in range(steps):
        st, _ = wcsph_step(model, st, cfg, dt, precision)
        states.append((st.x, st.v))
        if densities is not None:
            densities.append(st.sph["rho"])
    return states


def wcsph_final_state(model: OModel, x0, v0, cfg: WcsphConfig, dt: float, steps: int, precision="f64"):
    """(x, v) after ``steps`` WCSPH substeps, keeping no intermediate states."""
    st = OState(np.array(x0), np.array(v0), {})
    for _ in range(steps):
        st, _ = wcsph_step(model, st, cfg, dt, precision)
    return st.x, st.v


def wcsph_rollout_adjoint(model: OModel, states, cfg: WcsphConfig, dt: float, xbar, vbar, precision="f32"):
    """Reverse sweep over a stored rollout: from dL/dx_T, dL/dv_T return
    (dL/dx_0, dL/dv_0, dL/dstiffness[n], dL/dviscosity[n])."""
    gs = np.zeros(model.n)
    gv = np.zeros(model.n)
    xb, vb = np.array(xbar, np.float64), np.array(vbar, np.float64)
    for x, v in reversed(states[:-1]):
        xb, vb, s, u = wcsph_adjoint_step(model, OState(x, v, {}), cfg, dt, xb, vb, precision)
        gs += s
        gv += u
    return xb, vb, gs, gv


def _dfsph_residual(values: np.ndarray, dynamic_mask: np.ndarray) -> float:
    """_density_residual / _divergence_residual  solver_dfsph.py:579-596."""
    if not np.any(dynamic_mask):
        return 0.0
    err = np.maximum(values[dynamic_mask], 0.0)
    if err.size == 0:
        return 0.0
    return float(np.mean(err))


def dfsph_step(model: OModel, state_in: OState, cfg: DfsphConfig, dt: float, precision="f32", round_mode=0):
    """SolverDFSPH.step  solver_dfsph.py:598-898.  ``state_in.sph`` must hold
    ``kappa``/``kappa_v`` when warm starting.  Returns (state_out, extras)."""
    c = _Ctx(model, precision, round_mode)
    m, lib, n = c.m, c.lib, c.m.n
    f = c.dt_np
    x_in = np.ascontiguousarray(state_in.x, f)
    v_in = np.ascontiguousarray(state_in.v, f)
    sph = {k: np.zeros(n, f) for k in ("rho", "factor", "rho_star", "d_rho_dt", "kappa", "kappa_v")}
    sph["neighbor_count"] = np.zeros(n, np.int32)
    sph["v_star"] = np.zeros((n, 3), f)
    out = OState(np.zeros_like(x_in), np.zeros_like(v_in), sph)
    accel = np.zeros((n, 3), f)
    extras = {"accel": accel, "rho_iterations": 0, "vel_iterations": 0, "rho_residuals": [],
              "vel_residuals": []}
    if n == 0:
        return out, extras
    h = smoothing_length(cfg)
    pr = _particle_radius(cfg, m)
    vol = _volume(c)                                                                      # :616
    grid = c.new_grid().build(x_in, h)                                                    # :624
    g = grid.struct()
    gp = ctypes.byref(g)
    lib.ora_density(gp, _p(m.flags), _p(x_in), _p(m.rest_density), _p(vol), c.r(h), ctypes.c_int(0),
                    _p(sph["rho"]))                                                      # :626-639
    lib.ora_dfsph_factor(gp, _p(m.flags), _p(x_in), _p(m.mass), c.r(h), ctypes.c_int(int(cfg.min_neighbors_for_factor)),
                         _p(sph["factor"]), _p(sph["neighbor_count"]))                  # :640-653
    lib.ora_accel_non_pressure(gp, _p(m.flags), _p(m.mask), _p(x_in), _p(v_in), _p(sph["rho"]), c.r(pr),
                               _p(m.mass), _p(m.surface_tension), _p(m.viscosity), _p(m.gravity), _p(m.world),
                               c.r(h), ctypes.c_int(0), _p(accel))                      # :654-674
    lib.ora_dfsph_predict_velocity(ctypes.c_int(n), _p(m.flags), _p(m.mask), _p(v_in), _p(accel), c.r(dt),
                                   _p(sph["v_star"]))                                   # :675-687
    if cfg.warm_start:                                                                    # :689-707
        sph["kappa"][:] = np.asarray(state_in.sph["kappa"], f)
        lib.ora_dfsph_correction(gp, _p(m.flags), _p(m.mask), _p(x_in), _p(sph["v_star"]), _p(m.mass),
                                 _p(sph["rho"]), _p(sph["kappa"]), c.r(h), c.r(dt))
    dynamic_mask = ((m.flags & 1) != 0) & (m.mask == 0)                                   # :709
    for density_iter in range(cfg.max_rho_iterations):                                   # :711-767
        lib.ora_dfsph_density_change_rate(gp, _p(m.flags), _p(x_in), _p(sph["v_star"]), _p(m.mass), c.r(h),
                                          _p(sph["d_rho_dt"]))
        lib.ora_dfsph_predict_density_star(ctypes.c_int(n), _p(m.flags), _p(sph["rho"]), _p(sph["d_rho_dt"]),
                                           c.r(dt), _p(sph["rho_star"]))
        extras["rho_iterations"] = density_iter + 1
        residual = _dfsph_residual(sph["rho_star"] - m.rest_density, dynamic_mask)
        extras["rho_residuals"].append(residual)
        if density_iter > 0 and residual <= cfg.eta_density:
            break
        lib.ora_dfsph_kappa_density(ctypes.c_int(n), _p(m.flags), _p(m.mask), _p(sph["rho_star"]),
                                    _p(m.rest_density), _p(sph["factor"]), c.r(dt), c.r(cfg.max_kappa),
                                    _p(sph["kappa"]))
        lib.ora_dfsph_correction(gp, _p(m.flags), _p(m.mask), _p(x_in), _p(sph["v_star"]), _p(m.mass),
                                 _p(sph["rho"]), _p(sph["kappa"]), c.r(h), c.r(dt))
        extras["rho_corrections"] = density_iter + 1
    lib.ora_integrate_positions(ctypes.c_int(n), _p(m.flags), _p(m.mask), _p(x_in), _p(sph["v_star"]), c.r(dt),
                                _p(out.x))                                               # :769-781
    grid2 = c.new_grid().build(out.x, h)                                                 # :783
    g2 = grid2.struct()
    gp2 = ctypes.byref(g2)
    lib.ora_density(gp2, _p(m.flags), _p(out.x), _p(m.rest_density), _p(vol), c.r(h), ctypes.c_int(0),
                    _p(sph["rho"]))                                                      # :785-798
    lib.ora_dfsph_factor(gp2, _p(m.flags), _p(out.x), _p(m.mass), c.r(h),
                         ctypes.c_int(int(cfg.min_neighbors_for_factor)), _p(sph["factor"]),
                         _p(sph["neighbor_count"]))                                     # :799-812
    if cfg.warm_start:                                                                    # :814-832
        sph["kappa_v"][:] = np.asarray(state_in.sph["kappa_v"], f)
        lib.ora_dfsph_correction(gp2, _p(m.flags), _p(m.mask), _p(out.x), _p(sph["v_star"]), _p(m.mass),
                                 _p(sph["rho"]), _p(sph["kappa_v"]), c.r(h), c.r(dt))
    for divergence_iter in range(cfg.max_vel_iterations):                                # :834-883
        lib.ora_dfsph_density_change_rate(gp2, _p(m.flags), _p(out.x), _p(sph["v_star"]), _p(m.mass), c.r(h),
                                          _p(sph["d_rho_dt"]))
        extras["vel_iterations"] = divergence_iter + 1
        residual = _dfsph_residual(sph["d_rho_dt"], dynamic_mask)
        extras["vel_residuals"].append(residual)
        if divergence_iter > 0 and residual <= cfg.eta_divergence:
            break
        lib.ora_dfsph_kappa_divergence(ctypes.c_int(n), _p(m.flags), _p(m.mask), _p(sph["d_rho_dt"]),
                                       _p(sph["rho"]), _p(sph["factor"]), c.r(dt), c.r(cfg.max_kappa_v),
                                       _p(sph["kappa_v"]))
        lib.ora_dfsph_correction(gp2, _p(m.flags), _p(m.mask), _p(out.x), _p(sph["v_star"]), _p(m.mass),
                                 _p(sph["rho"]), _p(sph["kappa_v"]), c.r(h), c.r(dt))
        extras["vel_corrections"] = divergence_iter + 1
    out.v[:] = sph["v_star"]                                                              # :885
    lib.ora_sph_apply_bounds(ctypes.c_int(n), _p(m.flags), _p(m.mask), _p(out.x), _p(out.v), c.r(cfg.bound_width),
                             c.r(cfg.bound_height), c.r(cfg.bound_length), c.r(cfg.boundary_damping),
                             ctypes.c_int(0))                                            # :886-898
    extras.update(grid=grid, grid2=grid2, volume=vol)
    return out, extras


def pbf_w_delta_q(cfg: PbfConfig, precision="f32") -> float:
    """_w_delta_q  solver_pbf.py:671-674: the poly6 ``@wp.func`` called from
    Python scope, i.e. evaluated with Python floats (then passed as float32)."""
    del precision
    h = smoothing_length(cfg)
    r = float(cfg.scorr_q) * h
    if not (h > 0.0 and r < h):
        return 0.0
    h2 = h * h
    rhs = h2 - r * r
    cube_h = h * h * h
    k = 315.0 / (64.0 * math.pi * (cube_h * cube_h * cube_h))
    return k * (rhs * rhs * rhs)


def pbf_step(model: OModel, state_in: OState, cfg: PbfConfig, dt: float, precision="f32", round_mode=0):
    """SolverPBF.step  solver_pbf.py:676-954.  Returns (state_out, extras)."""
    c = _Ctx(model, precision, round_mode)
    m, lib, n = c.m, c.lib, c.m.n
    f = c.dt_np
    x_in = np.ascontiguousarray(state_in.x, f)
    v_in = np.ascontiguousarray(state_in.v, f)
    sph = {"rho": np.zeros(n, f), "pbf_lambda": np.zeros(n, f), "neighbor_count": np.zeros(n, np.int32),
           "delta_p": np.zeros((n, 3), f), "step_displacement": np.zeros((n, 3), f)}
    out = OState(np.zeros_like(x_in), np.zeros_like(v_in), sph)
    accel = np.zeros((n, 3), f)
    v_pred = np.zeros((n, 3), f)
    x_pred = np.zeros((n, 3), f)
    extras = {"accel": accel}
    if n == 0:
        return out, extras
    h = smoothing_length(cfg)
    pr = _particle_radius(cfg, m)
    eps = float(cfg.boundary_epsilon)
    bx, by, bz = cfg.bound_width, cfg.bound_height, cfg.bound_length
    N = ctypes.c_int(n)
    vol = _volume(c)                                                                        # :700

    def bounds():
        lib.ora_pbf_position_bounds(N, _p(m.flags), _p(m.mask), _p(x_pred), c.r(bx), c.r(by), c.r(bz), c.r(pr),
                                    c.r(eps))

    lib.ora_pbf_predict_velocity(N, _p(m.flags), _p(m.mask), _p(v_in), _p(m.gravity), _p(m.world), c.r(dt),
                                 _p(v_pred))                                               # :702-715
    lib.ora_pbf_predict_positions(N, _p(m.flags), _p(m.mask), _p(x_in), _p(v_pred), c.r(dt), _p(x_pred))  # :716-722
    bounds()                                                                                # :723-735
    grid = c.new_grid().build(x_pred, h)                                                    # :737
    g = grid.struct()
    gp = ctypes.byref(g)
    lib.ora_density(gp, _p(m.flags), _p(x_pred), _p(m.rest_density), _p(vol), c.r(h), ctypes.c_int(1),
                    _p(sph["rho"]))                                                        # :738-751
    lib.ora_accel_non_pressure(gp, _p(m.flags), _p(m.mask), _p(x_pred), _p(v_pred), _p(sph["rho"]), c.r(pr),
                               _p(m.mass), _p(m.surface_tension), _p(m.viscosity), _p(m.gravity), _p(m.world),
                               c.r(h), ctypes.c_int(1), _p(accel))                        # :752-770
    lib.ora_pbf_apply_non_pressure(N, _p(m.flags), _p(m.mask), _p(accel), c.r(dt), _p(v_pred), _p(x_pred))  # :771-783
    bounds()                                                                                # :784-796
    w_delta_q = pbf_w_delta_q(cfg, precision)                                               # :798
    for _ in range(cfg.max_iterations):                                                     # :800-858
        lib.ora_pbf_lambda(gp, _p(m.flags), _p(m.mask), _p(x_pred), _p(m.mass), _p(m.rest_density), c.r(h),
                           c.r(cfg.lambda_epsilon), ctypes.c_int(int(cfg.min_neighbors_for_lambda)),
                           _p(sph["pbf_lambda"]), _p(sph["rho"]), _p(sph["neighbor_count"]))
        lib.ora_pbf_delta_position(gp, _p(m.flags), _p(m.mask), _p(x_pred), _p(m.mass), _p(m.rest_density),
                                   _p(sph["pbf_lambda"]), c.r(h), c.r(cfg.scorr_k), c.r(cfg.scorr_n),
                                   c.r(w_delta_q), c.r(cfg.max_delta_position), _p(sph["delta_p"]))
        lib.ora_pbf_apply_delta(N, _p(m.flags), _p(m.mask), _p(x_pred), _p(sph["delta_p"]))
        if cfg.rebuild_grid_each_iteration:
            bounds()
            grid = c.new_grid().build(x_pred, h)
            g = grid.struct()
            gp = ctypes.byref(g)
    bounds()                                                                                # :860-872
    lib.ora_density(gp, _p(m.flags), _p(x_pred), _p(m.rest_density), _p(vol), c.r(h), ctypes.c_int(1),
                    _p(sph["rho"]))                                                        # :878-891
    lib.ora_pbf_update_velocity(N, _p(m.flags), _p(m.mask), _p(x_in), _p(x_pred), _p(v_in), c.r(dt),
                                _p(out.v))                                                 # :904-917
    if cfg.xsph_c > 0.0:                                                                    # :919-937
        v_tmp = np.zeros((n, 3), f)
        lib.ora_pbf_xsph(gp, _p(m.flags), _p(m.mask), _p(x_pred), _p(out.v), _p(m.mass), _p(sph["rho"]), c.r(h),
                         c.r(cfg.xsph_c), _p(v_tmp))
        out.v[:] = v_tmp
    out.x[:] = x_pred                                                                       # :939
    lib.ora_pbf_apply_bounds(N, _p(m.flags), _p(m.mask), _p(out.x), _p(out.v), c.r(bx), c.r(by), c.r(bz), c.r(pr),
                             c.r(eps), c.r(cfg.boundary_damping))                          # :940-954
    extras.update(grid=grid, volume=vol, v_pred=v_pred, x_pred=x_pred)
    return out, extras


# --------------------------------------------------------------------------
# Dam-break driver policy  (examples/dam_break/simulation.py)
# --------------------------------------------------------------------------


def norm_maxima(v: np.ndarray, a: np.ndarray | None, precision="f32"):
    """particle_norm_maxima  simulation.py:19-36 -> (max|v|^2, max|a|^2)."""
    lib = _lib(precision)
    v = np.ascontiguousarray(v, lib._np)
    a = np.zeros((0, 3), lib._np) if a is None else np.ascontiguousarray(a, lib._np)
    out = np.zeros(2, lib._np)
    lib.ora_norm_maxima(ctypes.c_int(v.shape[0]), _p(v), ctypes.c_int(a.shape[0]), _p(a), _p(out))
    return float(out[0]), float(out[1])


def cfl_limit(method, frame_dt, substeps, speed, accel_max, *, h, spacing, stiffness=None, exponent=None,
              rest_density=1000.0, viscosity=0.0):
    """The per-substep dt limit  simulation.py:166-181."""
    limit = frame_dt / substeps
    if method == "wcsph":
        sound_speed = math.sqrt(stiffness * exponent / rest_density)
        limit = min(limit, 0.25 * h / (sound_speed + speed))
        if accel_max > 0.0:
            limit = min(limit, 0.25 * math.sqrt(h / accel_max))
        if viscosity > 0.0:
            limit = min(limit, 0.025 * h ** 2 / viscosity)
    elif speed > 0.0:
        limit = min(limit, 0.4 * spacing / speed)
    return limit


# --------------------------------------------------------------------------
# APIC particle separation
# --------------------------------------------------------------------------


@dataclass
class SeparationConfig:
    """The ``SolverAPIC.Config`` fields the separation pass reads, with the
    reference's defaults (solver/apic/solver_apic.py:557-582)."""
    bound_lo: tuple = (-1.0, -1.0, 0.0)
    bound_hi: tuple = (1.0, 1.0, 2.0)
    clamp_eps: float = 1.0e-4
    wall_margin: float = -1.0
    wall_push_band: float = -1.0
    separation_rest_dist: float = -1.0
    separation_strength: float = 0.55
    separation_iterations: int = 2


def separation_params(cfg, radius) -> dict:
    """Host-side values of the pass: wall margin and push band
    (SolverAPIC.__init__, solver_apic.py:640-650), rest distance, query
    radius, iterations and HashGrid dims (_apply_particle_separation
    :1009-1027).  ``radius`` is the model's float32 particle radii."""
    radius = np.asarray(radius, np.float32)
    n = radius.shape[0]
    if n > 0 and float(cfg.wall_margin) < 0.0:
        wall_margin = float(np.mean(radius))
    elif float(cfg.wall_margin) < 0.0:
        wall_margin = float(cfg.clamp_eps)
    else:
        wall_margin = float(cfg.wall_margin)
    wall_push_band = 1.5 * wall_margin if float(cfg.wall_push_band) < 0.0 else float(cfg.wall_push_band)
    if float(cfg.separation_rest_dist) > 0.0:
        rest = float(cfg.separation_rest_dist)
    else:
        # with no particles the reference returns before this (:1006-1007)
        rest = float(2.0 * float(np.mean(radius))) if n > 0 else 0.0
    query = max(rest, 1.0e-5)
    lo = np.asarray(cfg.bound_lo, dtype=np.float64)
    hi = np.asarray(cfg.bound_hi, dtype=np.float64)
    extent = np.maximum(hi - lo, query)
    dims = np.maximum(4, np.ceil(extent / query).astype(int))
    return dict(rest_dist=rest, query_radius=query, strength=float(cfg.separation_strength),
                iterations=max(1, int(cfg.separation_iterations)),
                wall_band=max(float(wall_push_band), float(wall_margin) * 2.0), wall_margin=wall_margin,
                dims=(int(dims[0]), int(dims[1]), int(dims[2])), bound_lo=tuple(cfg.bound_lo),
                bound_hi=tuple(cfg.bound_hi))


def particle_separation(x, flags, params: dict, precision="f32", round_mode=0):
    """``iterations`` x (grid build at the current positions, accumulate,
    apply).  Returns (positions after, delta of the last iteration)."""
    lib = _lib(precision)
    real, npt = lib._real, lib._np
    x = np.array(x, dtype=npt, copy=True)
    flags = np.ascontiguousarray(flags, np.int32)
    lo = np.asarray(params["bound_lo"], npt)
    hi = np.asarray(params["bound_hi"], npt)
    delta = np.zeros_like(x)
    grid = OracleGrid(*params["dims"], precision=precision, round_mode=round_mode)
    for _ in range(params["iterations"]):
        grid.build(x, params["query_radius"])
        g = grid.struct()
        lib.ora_separation_accumulate(ctypes.byref(g), _p(flags), _p(x), real(params["rest_dist"]),
                                      real(params["query_radius"]), real(params["strength"]), _p(lo), _p(hi),
                                      real(params["wall_band"]), _p(delta))
        lib.ora_separation_apply(ctypes.c_int(x.shape[0]), _p(flags), _p(x), _p(delta), _p(lo), _p(hi),
                                 real(params["wall_margin"]))
    return x, delta

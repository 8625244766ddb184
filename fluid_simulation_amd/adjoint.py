"""Differentiable WCSPH: a hand-written adjoint of the substep behind
``torch.autograd.Function``.

The reference would differentiate ``SolverWCSPH.step`` (``solver_wcsph.py:431-535``)
with Warp's tape.  Here ``fs_wcsph_adjoint_step`` is the reverse-mode derivative of
one substep: density, EOS pressure, pressure and viscosity forces, symplectic
Euler and the outward-damping bounds; surface tension must be 0.  It
recomputes the forward substep and replays its neighbour lists in three gather
passes.  A rollout of ``steps`` substeps keeps (x, v) every ``checkpoint_every``
substeps. Backward recomputes each segment's states from its checkpoint and
sweeps it in reverse, so memory is O(steps/k + k) states.

Gradients flow to the initial positions and velocities and to the scalar EOS
stiffness and viscosity coefficient, which every particle shares.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .newton_compat import Axis, ModelBuilder
from .solver.sph.solver_wcsph import SolverWCSPH

TANK_LO = (-1.2, -0.28, 0.0)
TANK_HI = (1.2, 0.28, 1.2)
COLUMN_LO = (-1.2, -0.28, 0.0)
COLUMN_HI = (-0.6, 0.28, 0.6)


@dataclass
class ColumnScene:
    """Cubic-lattice water column in the dam-break tank: positions, spacing
    s, radius s/2, h = 4 radius = 2s, mass 1000 s^3."""
    positions: np.ndarray
    spacing: float
    lattice: tuple

    @property
    def count(self) -> int:
        return int(self.positions.shape[0])

    @property
    def radius(self) -> float:
        return 0.5 * self.spacing

    @property
    def mass(self) -> float:
        return 1000.0 * self.spacing ** 3


def column_scene(target_count: int, jitter: float = 0.01, seed: int = 0, margin: float = 0.0) -> ColumnScene:
    """The column x in [-1.2,-0.6], y in [-0.28,0.28], z in [0,0.6] as a
    lattice of about ``target_count`` points: s0 = (V/N)^(1/3), n_a =
    round(L_a/s0) per axis, s = min(L_a/n_a), the lattice centred in the
    column, then seeded uniform jitter of +-jitter*s (float32).  10k -> 10,164
    points (22 x 21 x 22); 50k -> 50,540 (38 x 35 x 38).  ``margin`` keeps
    the column that far from the tank walls it touches (x = -1.2, y = +-0.28,
    z = 0), as the reference's wave-tank spawn does (scene.py:44-75)."""
    lo, hi = np.asarray(COLUMN_LO, np.float64), np.asarray(COLUMN_HI, np.float64)
    lo = lo + margin
    hi[1] -= margin
    ext = hi - lo
    s0 = (float(np.prod(ext)) / target_count) ** (1.0 / 3.0)
    res = tuple(max(1, int(round(e / s0))) for e in ext)
    s = float(min(e / r for e, r in zip(ext, res)))
    axes = [lo[a] + 0.5 * (ext[a] - (res[a] - 1) * s) + s * np.arange(res[a]) for a in range(3)]
    grid = np.meshgrid(*axes, indexing="ij")
    pts = np.stack([g.ravel() for g in grid], axis=1)
    rng = np.random.default_rng(seed)
    pts = pts + rng.uniform(-jitter * s, jitter * s, pts.shape)
    return ColumnScene(pts.astype(np.float32), s, res)


def centroid_p95_loss(x: torch.Tensor, centroid_target: torch.Tensor, x95_target: torch.Tensor) -> torch.Tensor:
    """|centroid - c*|^2 + (P95(x) - x*)^2, P95 = the 95th percentile of the
    particles' x coordinate (linear interpolation between order statistics)."""
    c = x.mean(dim=0)
    p95 = torch.quantile(x[:, 0], 0.95)
    return ((c - centroid_target) ** 2).sum() + (p95 - x95_target) ** 2


class DifferentiableWCSPH:
    """A WCSPH model whose ``rollout`` is differentiable.

    ``stiffness`` / ``viscosity`` are the material scalars (every particle
    shares them); ``dt`` is fixed; surface tension is 0."""

    def __init__(self, positions, mass, radius, spacing, *, dt, steps, checkpoint_every=10, device="cuda:0",
                 bounds=(1.2, 0.28, 0.6), boundary_damping=0.0, clamp_negative_pressure=True, exponent=7.0,
                 rest_density=1000.0, gravity=(0.0, 0.0, -10.0), smoothing_length_coff=2.0):
        _lib.require_device(device)
        pts = np.asarray(positions, np.float32)
        n = pts.shape[0]
        b = ModelBuilder(up_axis=Axis.Z)
        b.default_particle_radius = radius
        SolverWCSPH.register_custom_attributes(b)
        b.add_particles(pos=pts, vel=np.zeros_like(pts), mass=np.full(n, mass, np.float32),
                        radius=np.full(n, radius, np.float32),
                        custom_attributes={"sph:rest_density": np.full(n, rest_density, np.float32),
                                           "sph:surface_tension": np.zeros(n, np.float32),
                                           "sph:exponent": np.full(n, exponent, np.float32),
                                           "sph:particle_mask": np.zeros(n, np.int32)})
        self.model = b.finalize(device=device)
        self.model.set_gravity(gravity)
        cfg = SolverWCSPH.SphConfig(particle_radius=radius, particle_length=spacing,
                                    smoothing_length_coff=smoothing_length_coff, bound_width=bounds[0],
                                    bound_height=bounds[1], bound_length=bounds[2],
                                    boundary_damping=boundary_damping,
                                    clamp_negative_pressure=clamp_negative_pressure)
        self.solver = SolverWCSPH(self.model, cfg)
        self.n = n
        self.dt = float(dt)
        self.steps = int(steps)
        self.k = max(1, int(checkpoint_every))
        dev = self.model.device
        self.device = dev
        self._s = [self.model.state(), self.model.state()]
        lib = _lib.load()
        self._scratch_bytes = int(lib.fs_wcsph_adjoint_bytes(n))
        self._scratch = torch.empty(self._scratch_bytes, dtype=torch.uint8, device=dev)
        f3 = lambda: torch.empty((n, 3), dtype=torch.float32, device=dev)
        f1 = lambda: torch.empty(n, dtype=torch.float32, device=dev)
        self._fx, self._fv, self._frho, self._fp, self._fa = f3(), f3(), f1(), f1(), f3()
        self._gs, self._gv = f1(), f1()

    # ---- parameters ------------------------------------------------------------
    def _set_material(self, stiffness: torch.Tensor, viscosity: torch.Tensor) -> None:
        self.model.sph.stiffness.tensor.copy_(stiffness.detach().to(torch.float32).expand(self.n))
        self.model.sph.viscosity.tensor.copy_(viscosity.detach().to(torch.float32).expand(self.n))

    # ---- forward ---------------------------------------------------------------
    def _advance(self, x: torch.Tensor, v: torch.Tensor, steps: int, keep=None):
        """``steps`` substeps from (x, v); returns the final (x, v) as new
        tensors.  ``keep(t, x_t, v_t)`` sees every input state."""
        s_in, s_out = self._s
        s_in.particle_q.tensor.copy_(x)
        s_in.particle_qd.tensor.copy_(v)
        for t in range(steps):
            if keep is not None:
                keep(t, s_in.particle_q.tensor, s_in.particle_qd.tensor)
            self.solver.step(s_in, s_out, None, None, self.dt)
            s_in, s_out = s_out, s_in
        return s_in.particle_q.tensor.clone(), s_in.particle_qd.tensor.clone()

    def forward(self, x0, v0, stiffness, viscosity):
        """Roll out ``steps`` substeps; returns (x_T, v_T) and the checkpoints."""
        self._set_material(stiffness, viscosity)
        checkpoints = []

        def keep(t, x, v):
            if t % self.k == 0:
                checkpoints.append((x.clone(), v.clone()))

        xT, vT = self._advance(x0, v0, self.steps, keep)
        return xT, vT, checkpoints

    # ---- backward --------------------------------------------------------------
    def adjoint_step(self, x_in, v_in, xbar, vbar) -> None:
        """dL/d(x_in, v_in) from dL/d(x_out, v_out), in place in xbar/vbar;
        accumulates the per-particle stiffness/viscosity gradients."""
        sol = self.solver
        p = _lib.FsWcsphParams()
        p.smoothing_length = float(sol.smoothing_length)
        p.particle_radius = sol._particle_radius()
        p.bound[:] = sol._bounds
        p.boundary_damping = float(sol.config.boundary_damping)
        p.dt = self.dt
        p.clamp_negative_pressure = 1 if sol.config.clamp_negative_pressure else 0
        io = _lib.FsWcsphAdjointIO()
        f = io.forward
        f.x_in, f.v_in = x_in.data_ptr(), v_in.data_ptr()
        f.x_out, f.v_out = self._fx.data_ptr(), self._fv.data_ptr()
        f.rho, f.pressure, f.accel = self._frho.data_ptr(), self._fp.data_ptr(), self._fa.data_ptr()
        io.xbar_out, io.vbar_out = xbar.data_ptr(), vbar.data_ptr()
        io.xbar_in, io.vbar_in = xbar.data_ptr(), vbar.data_ptr()
        io.grad_stiffness, io.grad_viscosity = self._gs.data_ptr(), self._gv.data_ptr()
        io.scratch = self._scratch.data_ptr()
        m = sol._model_struct(with_eos=True)
        _lib.check(_lib.load().fs_wcsph_adjoint_step(ctypes.byref(self.model.particle_grid._c),
                                                     ctypes.byref(sol._work), ctypes.byref(m), ctypes.byref(p),
                                                     ctypes.byref(io), self._scratch_bytes,
                                                     ctypes.c_void_p(sol._stream())), "fs_wcsph_adjoint_step")

    def backward(self, checkpoints, gx, gv):
        """Reverse sweep: returns (dL/dx0, dL/dv0, dL/dstiffness, dL/dviscosity)."""
        xbar = gx.detach().to(torch.float32).contiguous().clone()
        vbar = gv.detach().to(torch.float32).contiguous().clone()
        self._gs.zero_()
        self._gv.zero_()
        nseg = len(checkpoints)
        for seg in reversed(range(nseg)):
            t0 = seg * self.k
            t1 = min(t0 + self.k, self.steps)
            states = []
            last = self._advance(checkpoints[seg][0], checkpoints[seg][1], t1 - t0 - 1,
                                 lambda t, x, v: states.append((x.clone(), v.clone())))
            states.append(last)
            for x_t, v_t in reversed(states):
                self.adjoint_step(x_t, v_t, xbar, vbar)
        gB = self._gs.double().sum()
        gnu = self._gv.double().sum()
        return xbar, vbar, gB, gnu

    def rollout(self, x0: torch.Tensor, v0: torch.Tensor, stiffness: torch.Tensor, viscosity: torch.Tensor):
        """Differentiable (x_T, v_T) after ``steps`` substeps."""
        return _Rollout.apply(x0, v0, stiffness, viscosity, self)


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, v0, stiffness, viscosity, sim):
        xT, vT, checkpoints = sim.forward(x0.detach(), v0.detach(), stiffness, viscosity)
        ctx.sim = sim
        ctx.checkpoints = checkpoints
        ctx.material = (stiffness.detach().clone(), viscosity.detach().clone())
        ctx.dtypes = (stiffness.dtype, viscosity.dtype)
        return xT, vT

    @staticmethod
    def backward(ctx, gx, gv):
        sim = ctx.sim
        if gx is None:
            gx = torch.zeros((sim.n, 3), dtype=torch.float32, device=sim.device)
        if gv is None:
            gv = torch.zeros((sim.n, 3), dtype=torch.float32, device=sim.device)
        sim._set_material(*ctx.material)
        gx0, gv0, gB, gnu = sim.backward(ctx.checkpoints, gx, gv)
        return gx0, gv0, gB.to(ctx.dtypes[0]), gnu.to(ctx.dtypes[1]), None


def dt_for(spacing: float, stiffness=250000.0, exponent=7.0, rest_density=1000.0, coff=2.0) -> float:
    """Fixed substep 0.25 h / c0 with c0 = sqrt(B gamma / rho0) (the
    acoustic CFL term of simulation.py:166-187)."""
    return 0.25 * spacing * coff / math.sqrt(stiffness * exponent / rest_density)

"""MI355X-native particle-fluid hot path: Warp-compatible hash grid and the
WCSPH / DFSPH / PBF neighbour passes as hand-written gfx950 HIP kernels
(``libfluidsim_hip.so``), behind the reference's ``SolverBase.step`` API.

Module map (reference counterpart in brackets):
  solver.sph.solver_wcsph.SolverWCSPH   [solver/sph/solver_wcsph.py]
  solver.sph.solver_dfsph.SolverDFSPH   [solver/sph/solver_dfsph.py]
  solver.pbf.SolverPBF                  [solver/pbf/solver_pbf.py]
  examples.dam_break.scene / simulation [examples/dam_break/scene.py, simulation.py]
  newton_compat                         [Newton Model/State/ModelBuilder/SolverBase]
  hashgrid.HashGrid                     [warp.HashGrid]
"""
from .array import Array, copy
from .newton_compat import Axis, Model, ModelBuilder, ParticleFlags, SolverBase, State

__all__ = ["Array", "copy", "Axis", "Model", "ModelBuilder", "ParticleFlags", "SolverBase", "State"]
__version__ = "0.1.0"

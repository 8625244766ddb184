from .solver_pbf import SolverPBF

__all__ = ["SolverPBF"]

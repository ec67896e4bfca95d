"""Solver classes with the reference API (GeoWarp solvers/*.py)."""
from .quasi_static_solver_2d import quasi_static_solver_2d  # noqa: F401
from .quasi_static_coupled_solver_2d import quasi_static_coupled_solver_2d  # noqa: F401
from .quasi_static_coupled_solver_3d import quasi_static_coupled_solver_3d  # noqa: F401

"""Gaussian-beam propagators (closed forms for homogeneous media)."""
from . import gb_solvers  # noqa: F401
from .gb_solvers import SolverConfig, solve_hom_diag, solve_hom_TR  # noqa: F401

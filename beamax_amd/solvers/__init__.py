"""MSGB solvers."""
from .msgb_solver import MSGBSolver  # noqa: F401
from . import faces  # noqa: F401

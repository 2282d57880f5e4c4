"""Small host utilities (metrics and parameter batching)."""
from .metrics import rel_l2, max_rel  # noqa: F401

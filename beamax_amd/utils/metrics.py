"""Error metrics (reference ``beamax/utils/arrays.py:271-289`` and
``examples/forward/2d_forward.py:70-75``)."""

import numpy as np


def _np(x):
    if hasattr(x, "detach"):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def rel_l2(a, b) -> float:
    """||a - b|| / (||a|| + 1e-30), a = reference."""
    a = _np(a).astype(np.float64)
    b = _np(b).astype(np.float64)
    return float(np.linalg.norm(a - b) / (np.linalg.norm(a) + 1e-30))


def max_rel(a, b) -> float:
    """max|a - b| / max|a|, a = reference."""
    a = _np(a).astype(np.float64)
    b = _np(b).astype(np.float64)
    return float(np.max(np.abs(a - b)) / (np.max(np.abs(a)) + 1e-300))

"""The high-frequency seam of the reference's hybrid solver.

``HybridSolver._solve_hf`` (reference ``beamax/solvers/hybrid_solver.py:
795-853``) hands the high-frequency part of a split problem to its HF solver:
an ``MSGBSolver`` gets ``forward(data, domain, sensors, ts, wpt)``,
``time_reversal(data, domain, sensors, sources, ts, data_domain, data_wpt)``
or ``adjoint(...)`` with the same arguments; any other solver is called as
``solver.<method>(data, domain, mask, ts, **kwargs)`` (forward) or
``solver.<method>(data, domain, mask, sources, ts)``.  ``solve_hf`` is that
dispatch with this package's ``MSGBSolver`` in the MSGB branch, so the
hybrid's HF step runs on the GPU when its ``hf_solver`` is this solver.  The
rest of the hybrid (frequency splitting, the low-frequency solver, k-Wave)
is outside the MSGB operator path and not part of this package.
"""

from __future__ import annotations

from .msgb_solver import MSGBSolver


def solve_hf(hf_solver, data, domain, sensors, mask, ts, wpt, method: str = "forward",
             **kwargs):
    """Solve the high-frequency component with ``hf_solver`` by ``method``
    ("forward", "time_reversal" or "adjoint"); error message as the
    reference's for an unknown method."""
    if isinstance(hf_solver, MSGBSolver):
        if method == "forward":
            return hf_solver.forward(data, domain, sensors, ts, wpt, **kwargs)
        if method == "time_reversal":
            return hf_solver.time_reversal(data, domain, sensors, kwargs["sources"], ts,
                                           kwargs["data_domain"], kwargs["data_wpt"])
        if method == "adjoint":
            return hf_solver.adjoint(data, domain, sensors, kwargs["sources"], ts,
                                     kwargs["data_domain"], kwargs["data_wpt"])
        raise ValueError(f"Unknown HF method: {method!r}")
    solver_fn = getattr(hf_solver, method)
    if method == "forward":
        return solver_fn(data, domain, mask, ts, **kwargs)
    if method in ("time_reversal", "adjoint"):
        return solver_fn(data, domain, mask, kwargs["sources"], ts)
    raise ValueError(f"Unknown HF method: {method!r}")

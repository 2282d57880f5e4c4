"""Closed-form Gaussian-beam propagators (``beamax/gb/gb_solvers.py``).

``solve_hom_diag`` keeps the reference ``SolverFn`` signature
``(x0, p0, M0, a0, mode, ts, c, lam=None, config=None) -> (xt, pt, Mt, At)``
with shapes (b, Nt, d), (b, Nt, d), (b, Nt, d, d), (b, Nt, 1)
(gb_solvers.py:24-93, 356-422) and runs the HIP kernel ``bh_solve_hom_diag``.
Inside ``MSGBSolver`` the same closed form is fused into the summation
kernel; these functions also serve as markers selecting that path.
``solve_hom_TR`` (gb_solvers.py:484-594) is fused into ``bh_adjoint_params``.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

from .. import _lib


def _c0(c):
    if callable(c):
        import numpy as _np
        val = _np.asarray(c(_np.zeros((1, 1))))
        return float(val.reshape(-1)[0])
    return float(c)


def solve_hom_diag(x0, p0, M0, a0, mode, ts, c, lam=None, config=None):
    """Homogeneous medium, diagonal M0 (gb_solvers.py:356-422) on the GPU."""
    import torch
    dev = x0.device if isinstance(x0, torch.Tensor) else torch.device("cuda")
    f64 = lambda v: torch.as_tensor(v, device=dev).to(torch.float64).contiguous()
    c128 = lambda v: torch.as_tensor(v, device=dev).to(torch.complex128).contiguous()
    x0, p0, mode, ts = f64(x0), f64(p0), f64(mode).reshape(-1), f64(ts).reshape(-1)
    M0 = c128(M0)
    a0 = c128(a0).reshape(-1)
    b, d = x0.shape
    alpha = torch.diagonal(M0, dim1=1, dim2=2).contiguous()
    nt = ts.numel()
    xt = torch.empty((b, nt, d), dtype=torch.float64, device=dev)
    pt = torch.empty_like(xt)
    Mt = torch.empty((b, nt, d, d), dtype=torch.complex128, device=dev)
    At = torch.empty((b, nt, 1), dtype=torch.complex128, device=dev)
    _lib.call("bh_solve_hom_diag", b, d, _lib.ptr(x0), _lib.ptr(p0), _lib.ptr(alpha),
              _lib.ptr(a0), _lib.ptr(mode), _c0(c), _lib.ptr(ts), nt, _lib.ptr(xt),
              _lib.ptr(pt), _lib.ptr(Mt), _lib.ptr(At), _lib.stream_handle())
    return xt, pt, Mt, At


def solve_hom_TR(xT, pT, mT, aT, mode, ts, c, lam=None, config=None):
    """Terminal-value closed form (gb_solvers.py:484-594).

    Inside ``MSGBSolver.adjoint`` this propagation is fused into the adjoint
    parameter kernel; called directly it is not provided on the GPU.
    """
    raise NotImplementedError(
        "solve_hom_TR runs fused inside MSGBSolver.adjoint (bh_adjoint_params)")


@dataclass(frozen=True)
class SolverConfig:
    """Numerical ODE settings (gb_solvers.py:597-703); kept for API parity.

    The closed-form propagators used on the GPU take no tolerances.
    """

    rtol: float = 1e-7
    atol: float = 1e-9
    dt0: Optional[float] = None
    max_steps: int = 4096

    @classmethod
    def from_precision(cls, x64: bool = True) -> "SolverConfig":
        return cls() if x64 else cls(rtol=1e-4, atol=1e-6)

"""Reconstruction with the MSGB operator: CG on the normal equations.

``cgnr`` minimises ||F x - d||_2 for a linear operator F with an exact
transpose (``solvers.linear.FrozenForward``) by conjugate gradients on
F^T F x = F^T d (CGNR; Hestenes and Stiefel 1952, Saad "Iterative Methods for
Sparse Linear Systems" alg. 8.4), with an optional symmetric positive
definite preconditioner M ~ (F^T F)^-1 applied as z = M(g).  The stopping
test is the plain normal-equation residual ||F^T (d - F x)|| / ||F^T d||,
whatever the preconditioner.  The reference has no reconstruction solver;
this is the operator use its adjoint/time-reversal examples lead to
(``examples/reconstruction/2d_time_reversal_and_adjoint.py``).
"""

from __future__ import annotations

import time

import numpy as np


def _dot(a, b):
    return float((a.reshape(-1).double() * b.reshape(-1).double()).sum().item())


def _sync(t):
    import torch
    if t.is_cuda:
        torch.cuda.synchronize(t.device)


def cgnr(op, data, *, x0=None, tol: float = 1e-3, maxiter: int = 500, precond=None,
         callback=None):
    """Returns (x, info); info holds iterations, residual history, seconds and
    whether ``tol`` was reached.  ``op`` needs ``apply(x)``, ``transpose(r)``
    and ``shape = ((Nt, S), N)``."""
    import torch
    _sync(data)
    t0 = time.perf_counter()
    d = data.reshape(op.shape[0]).to(torch.float64)
    if x0 is None:
        x = torch.zeros(op.shape[1], dtype=torch.float64, device=d.device)
        r = d.clone()
    else:
        x = x0.to(torch.float64).reshape(op.shape[1]).clone()
        r = d - op.apply(x).to(torch.float64)
    g = op.transpose(r)                    # F^T r
    if x0 is None:
        g0 = np.sqrt(_dot(g, g))
    else:
        # residuals stay relative to ||F^T d||, as from a zero start
        gd = op.transpose(d)
        g0 = np.sqrt(_dot(gd, gd))
    hist = [np.sqrt(_dot(g, g)) / g0 if g0 > 0 else 0.0]
    if g0 == 0.0:
        return x, {"iterations": 0, "residuals": hist, "seconds": 0.0, "converged": True}
    if hist[0] <= tol:
        _sync(x)
        return x, {"iterations": 0, "residuals": hist, "seconds": time.perf_counter() - t0,
                   "converged": True}
    z = precond(g) if precond is not None else g
    p = z.clone()
    gz = _dot(g, z)
    it = 0
    converged = False
    while it < maxiter:
        w = op.apply(p).to(torch.float64)  # F p
        ww = _dot(w, w)
        if ww <= 0.0:
            break
        alpha = gz / ww
        x.add_(p, alpha=alpha)
        r.add_(w, alpha=-alpha)
        g = op.transpose(r)
        it += 1
        res = np.sqrt(_dot(g, g)) / g0
        hist.append(res)
        if callback is not None:
            callback(it, x, res)
        if res <= tol:
            converged = True
            break
        z = precond(g) if precond is not None else g
        gz_new = _dot(g, z)
        beta = gz_new / gz
        gz = gz_new
        p.mul_(beta).add_(z)
    _sync(x)
    return x, {"iterations": it, "residuals": hist, "seconds": time.perf_counter() - t0,
               "converged": converged}

"""CGNR against a direct least-squares solve (CPU, small dense operators)."""

import numpy as np
import pytest

from beamax_amd.reconstruct import cgnr


class _MatOp:
    def __init__(self, A, shape_out, shape_in):
        import torch
        self.A = torch.as_tensor(A)
        self.shape = (shape_out, shape_in)

    def apply(self, x):
        return (self.A @ x.reshape(-1)).reshape(self.shape[0])

    def transpose(self, r):
        return (self.A.T @ r.reshape(-1)).reshape(self.shape[1])


@pytest.mark.parametrize("m,n", [(60, 20), (40, 40), (30, 45)])
def test_cgnr_reaches_least_squares(m, n):
    import torch
    rng = np.random.default_rng(m + n)
    A = rng.standard_normal((m, n)) @ np.diag(np.linspace(1.0, 8.0, n))
    d = rng.standard_normal(m)
    op = _MatOp(A, (m,), (n,))
    x, info = cgnr(op, torch.as_tensor(d), tol=1e-10, maxiter=10 * n)
    assert info["converged"]
    g = A.T @ (d - A @ x.numpy())
    assert np.linalg.norm(g) <= 1e-9 * np.linalg.norm(A.T @ d)
    # minimum-norm least squares solution (x0 = 0 keeps CGNR in range(A^T))
    ref = np.linalg.lstsq(A, d, rcond=None)[0]
    assert np.linalg.norm(x.numpy() - ref) <= 1e-7 * np.linalg.norm(ref)
    # the residual history is the plain normal-equation residual
    assert info["residuals"][0] == 1.0 and info["residuals"][-1] <= 1e-10


def test_cgnr_preconditioned_matches_plain():
    import torch
    rng = np.random.default_rng(7)
    m, n = 80, 30
    A = rng.standard_normal((m, n)) @ np.diag(np.logspace(0, 2, n))
    d = rng.standard_normal(m)
    op = _MatOp(A, (m,), (n,))
    D = torch.as_tensor(1.0 / np.sum(A * A, axis=0))     # Jacobi on A^T A
    x1, i1 = cgnr(op, torch.as_tensor(d), tol=1e-9, maxiter=500)
    x2, i2 = cgnr(op, torch.as_tensor(d), tol=1e-9, maxiter=500, precond=lambda g: D * g)
    assert i1["converged"] and i2["converged"]
    assert i2["iterations"] < i1["iterations"]
    assert np.linalg.norm(x1.numpy() - x2.numpy()) <= 1e-6 * np.linalg.norm(x1.numpy())


def test_fcgnr_with_a_fixed_spd_preconditioner_matches_cgnr():
    """With a fixed SPD preconditioner, flexible CG on the normal equations
    takes the same steps as preconditioned CGNR (up to rounding) and reaches
    the same least-squares solution; with a step-dependent preconditioner
    it still converges."""
    import torch
    from beamax_amd.reconstruct import fcgnr
    rng = np.random.default_rng(8)
    m, n = 70, 25
    A = rng.standard_normal((m, n)) @ np.diag(np.logspace(0, 1.5, n))
    d = rng.standard_normal(m)
    op = _MatOp(A, (m,), (n,))
    D = torch.as_tensor(1.0 / np.sum(A * A, axis=0))
    x1, i1 = cgnr(op, torch.as_tensor(d), tol=1e-10, maxiter=500, precond=lambda g: D * g)
    x2, i2 = fcgnr(op, torch.as_tensor(d), lambda g: D * g, tol=1e-10, maxiter=500, window=3)
    assert i1["converged"] and i2["converged"]
    assert abs(i1["iterations"] - i2["iterations"]) <= 2
    assert np.linalg.norm(x1.numpy() - x2.numpy()) <= 1e-7 * np.linalg.norm(x1.numpy())
    k = [0]

    def varying(g):
        k[0] += 1
        return (1.0 + 0.5 * np.sin(k[0])) * D * g
    x3, i3 = fcgnr(op, torch.as_tensor(d), varying, tol=1e-10, maxiter=500, window=3)
    assert i3["converged"]
    assert np.linalg.norm(x3.numpy() - x1.numpy()) <= 1e-7 * np.linalg.norm(x1.numpy())

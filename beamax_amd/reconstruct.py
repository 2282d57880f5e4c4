"""Reconstruction with the MSGB operator: CG on the normal equations.

``cgnr`` minimises ||F x - d||_2 for a linear operator F with an exact
transpose (``solvers.linear.FrozenForward``) by conjugate gradients on
F^T F x = F^T d (CGNR; Hestenes and Stiefel 1952, Saad "Iterative Methods for
Sparse Linear Systems" alg. 8.4), with an optional symmetric positive
definite preconditioner M ~ (F^T F)^-1 applied as z = M(g).  ``fcgnr`` is the
flexible variant (Notay 2000, "Flexible conjugate gradients"): each new
direction is made F^T F-conjugate to the last ``window`` directions by
explicit projection, so the preconditioner may change from step to step
(a network, for instance).  The stopping test of both is the plain
normal-equation residual ||F^T (d - F x)|| / ||F^T d||, whatever the
preconditioner.  ``FourierPreconditioner`` is a fixed SPD preconditioner for
2-D images (a negative power of |k| in the Fourier domain).

Every scalar of an iteration (step length, conjugation weights, residual
norm) stays a device tensor: an iteration issues no host synchronisation
except the one read of the residual norm for the stopping test, and with
``graph=True`` the whole iteration (operator, transpose, preconditioner,
vector updates) is captured once as a HIP graph and replayed.  The
reference has no reconstruction solver; this is the operator use its
adjoint/time-reversal examples lead to
(``examples/reconstruction/2d_time_reversal_and_adjoint.py``).
"""

from __future__ import annotations

import time

import numpy as np


def _dot(a, b):
    """<a, b> as a 0-d FP64 device tensor (no host read)."""
    import torch
    return torch.dot(a.reshape(-1), b.reshape(-1))


def _sync(t):
    import torch
    if t.is_cuda:
        torch.cuda.synchronize(t.device)


class _State:
    """The iteration's device tensors (fixed addresses, so a captured graph
    reads and writes the same memory on every replay)."""

    def __init__(self, op, d, x0):
        import torch
        self.op = op
        self.d = d
        if x0 is None:
            self.x = torch.zeros(op.shape[1], dtype=torch.float64, device=d.device)
            self.r = d.clone()
        else:
            self.x = x0.to(torch.float64).reshape(op.shape[1]).clone()
            self.r = d - op.apply(self.x).to(torch.float64)
        self.g = op.transpose(self.r).to(torch.float64)
        self.gg = _dot(self.g, self.g)


class FourierPreconditioner:
    """M g = Re F^-1[w F g] for a 2-D image g: the Fourier multiplier
    w(k) = max(|k|, floor)^power / max_k w (|k| in cycles per sample), a
    symmetric positive definite map, so it can precondition ``cgnr``.  With
    a negative power it damps the high spatial frequencies, which the
    normal operator F^T F of the MSGB forward amplifies; the 2-D FFTs are
    the package's FFT kernel (``bh_fft``, rows then columns, FP64), applied
    in place in a buffer allocated once (no allocation per call, so a
    captured CG iteration can include it)."""

    def __init__(self, shape, power: float = -0.5, floor: float = 0.03, device=None):
        import torch
        from . import _lib
        self._lib = _lib
        self.shape = tuple(int(n) for n in shape)
        if len(self.shape) != 2:
            raise ValueError(f"FourierPreconditioner: a 2-D image shape, got {self.shape}")
        n0, n1 = self.shape
        dev = torch.device("cuda") if device is None else torch.device(device)
        k0 = torch.as_tensor(np.fft.fftfreq(n0), dtype=torch.float64, device=dev)
        k1 = torch.as_tensor(np.fft.fftfreq(n1), dtype=torch.float64, device=dev)
        kk = torch.sqrt(k0[:, None] ** 2 + k1[None, :] ** 2)
        w = torch.clamp(kk, min=float(floor)) ** float(power)
        self.w = (w / w.max()).unsqueeze(-1).contiguous()      # (n0, n1, 1)
        self.buf = torch.zeros((n0, n1, 2), dtype=torch.float64, device=dev)
        lib = _lib.load()
        nb = max(lib.bh_fft_workspace_size(n1, n0), lib.bh_fft_workspace_size(n0, n1))
        self.ws = _lib.workspace(max(int(nb), 8), dev)

    def _fft2(self, direction: int, scale: float):
        lib, n0, n1 = self._lib, *self.shape
        st = lib.stream_handle()
        # rows: transform t starts at t * n1, unit stride
        lib.call("bh_fft", lib.ptr(self.buf), n1, n0, 1, 0, n1, 1, direction, 1.0,
                 lib.ptr(self.ws), self.ws.numel(), st)
        # columns: transform t starts at t, stride n1
        lib.call("bh_fft", lib.ptr(self.buf), n0, n1, n1, 1, 0, n1, direction, scale,
                 lib.ptr(self.ws), self.ws.numel(), st)

    def __call__(self, g):
        n0, n1 = self.shape
        self.buf[..., 0].copy_(g.reshape(n0, n1))
        self.buf[..., 1].zero_()
        self._fft2(-1, 1.0)
        self.buf.mul_(self.w)
        self._fft2(1, 1.0 / (n0 * n1))
        return self.buf[..., 0].reshape(g.shape)


def cgnr(op, data, *, x0=None, tol: float = 1e-3, maxiter: int = 500, precond=None,
         callback=None, graph: bool = False):
    """Returns (x, info); info holds iterations, residual history, seconds and
    whether ``tol`` was reached.  ``op`` needs ``apply(x)``, ``transpose(r)``
    and ``shape = ((Nt, S), N)``; ``precond(g)`` returns M g for the FP64
    gradient image g."""
    import torch
    _sync(data)
    t0 = time.perf_counter()
    d = data.reshape(op.shape[0]).to(torch.float64)
    st = _State(op, d, x0)
    if x0 is None:
        g0 = float(st.gg.sqrt().item())
    else:
        # residuals stay relative to ||F^T d||, as from a zero start
        gd = op.transpose(d).to(torch.float64)
        g0 = float(_dot(gd, gd).sqrt().item())
    hist = [float(st.gg.sqrt().item()) / g0 if g0 > 0 else 0.0]
    if g0 == 0.0 or hist[0] <= tol:
        _sync(st.x)
        return st.x, {"iterations": 0, "residuals": hist,
                      "seconds": time.perf_counter() - t0, "converged": True}
    z = precond(st.g) if precond is not None else st.g
    p = z.to(torch.float64).clone()
    gz = _dot(st.g, z)

    def step():
        w = op.apply(p).to(torch.float64)               # F p
        alpha = gz / _dot(w, w)
        st.x.addcmul_(p, alpha)
        st.r.addcmul_(w, alpha, value=-1.0)
        st.g.copy_(op.transpose(st.r))
        st.gg.copy_(_dot(st.g, st.g))
        zn = precond(st.g) if precond is not None else st.g
        gz_new = _dot(st.g, zn)
        p.mul_(gz_new / gz).add_(zn)
        gz.copy_(gz_new)

    run = step
    if graph and d.is_cuda:
        step()                                           # warm: workspaces, caches
        hist.append(float(st.gg.sqrt().item()) / g0)
        if hist[-1] > tol:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                step()
            run = g.replay
    it = len(hist) - 1
    converged = hist[-1] <= tol
    while not converged and it < maxiter:
        run()
        it += 1
        res = float(st.gg.sqrt().item()) / g0            # the one host read per iteration
        hist.append(res)
        if callback is not None:
            callback(it, st.x, res)
        converged = res <= tol
    _sync(st.x)
    return st.x, {"iterations": it, "residuals": hist, "seconds": time.perf_counter() - t0,
                  "converged": converged}


def fcgnr(op, data, precond, *, x0=None, tol: float = 1e-3, maxiter: int = 200,
          window: int = 4, callback=None):
    """Flexible CG on the normal equations: z_k = precond(g_k) may be a
    nonlinear or step-dependent map; p_k = z_k minus its F^T F-projections on
    the last ``window`` directions (Notay's FCG(m)); the F^T F p products come
    from gradient differences, so an iteration costs one ``apply``, one
    ``transpose`` and one ``precond``.  Same return value as ``cgnr``."""
    import torch
    _sync(data)
    t0 = time.perf_counter()
    d = data.reshape(op.shape[0]).to(torch.float64)
    st = _State(op, d, x0)
    if x0 is None:
        g0 = float(st.gg.sqrt().item())
    else:
        gd = op.transpose(d).to(torch.float64)
        g0 = float(_dot(gd, gd).sqrt().item())
    hist = [float(st.gg.sqrt().item()) / g0 if g0 > 0 else 0.0]
    dirs = []                                            # (p, F^T F p, |F p|^2)
    it = 0
    converged = g0 == 0.0 or hist[0] <= tol
    while not converged and it < maxiter:
        z = precond(st.g).to(torch.float64).reshape(st.g.shape)
        p = z.clone()
        for pj, apj, nj in dirs:
            p.addcmul_(pj, _dot(z, apj) / nj, value=-1.0)
        w = op.apply(p).to(torch.float64)
        ww = _dot(w, w)
        alpha = _dot(p, st.g) / ww
        st.x.addcmul_(p, alpha)
        st.r.addcmul_(w, alpha, value=-1.0)
        gn = op.transpose(st.r).to(torch.float64)
        ap = (st.g - gn) / alpha                          # F^T F p
        st.g.copy_(gn)
        st.gg.copy_(_dot(gn, gn))
        dirs.append((p, ap, ww))
        if len(dirs) > window:
            dirs.pop(0)
        it += 1
        res = float(st.gg.sqrt().item()) / g0
        hist.append(res)
        if callback is not None:
            callback(it, st.x, res)
        converged = res <= tol
    _sync(st.x)
    return st.x, {"iterations": it, "residuals": hist, "seconds": time.perf_counter() - t0,
                  "converged": converged}

"""Closed-form Gaussian-beam propagators (``beamax/gb/gb_solvers.py``).

``solve_hom_diag`` keeps the reference ``SolverFn`` signature
``(x0, p0, M0, a0, mode, ts, c, lam=None, config=None) -> (xt, pt, Mt, At)``
with shapes (b, Nt, d), (b, Nt, d), (b, Nt, d, d), (b, Nt, 1)
(gb_solvers.py:24-93, 356-422) and runs the HIP kernel ``bh_solve_hom_diag``.
Inside ``MSGBSolver`` the same closed form is fused into the summation
kernel; these functions also serve as markers selecting that path.
``solve_hom_TR`` (gb_solvers.py:484-594) runs ``bh_solve_hom_TR``; inside
``MSGBSolver`` it is fused into ``bh_adjoint_params``.
``solve_ODE_base`` (gb_solvers.py:941-1016) integrates the beam ODEs in a
heterogeneous (grid) medium on the GPU (``bh_ode_forward``); inside
``MSGBSolver`` it selects the trajectory-fed summation.  ``solve_ODE_batch_t``
(gb_solvers.py:1017-1113) integrates per-beam intervals (``bh_ode_static``);
``MSGBSolver`` runs it for the adjoint and time reversal in a heterogeneous
medium.
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


def _tr_time_grid(ts, b):
    """Per-beam time rows (b, nt) from ts as the reference reads it
    (gb_solvers.py:547-564): a scalar, a vector of b per-beam times, a common
    timeline, (b, nt), (nt, b) or anything else flattened into one timeline."""
    import torch
    t = ts if isinstance(ts, torch.Tensor) else torch.as_tensor(np.asarray(ts, dtype=np.float64))
    t = t.to(torch.float64)
    if t.ndim == 0:
        return t.reshape(1, 1).expand(b, 1)
    if t.ndim == 1:
        if t.shape[0] == b:
            return t.reshape(b, 1)
        return t.reshape(1, -1).expand(b, t.shape[0])
    if t.shape[0] == b:
        return t.reshape(b, -1)
    if t.ndim == 2 and t.shape[1] == b:
        return t.transpose(0, 1)
    flat = t.reshape(-1)
    return flat.reshape(1, -1).expand(b, flat.shape[0])


def solve_hom_TR(xT, pT, mT, aT, mode, ts, c, lam=None, config=None):
    """Terminal-value closed form (gb_solvers.py:484-594) on the GPU
    (``bh_solve_hom_TR``): states at each beam's times measured from its
    first time, x = xT + c0 p^ dt, M = mT (I + c0 dt P_perp mT / |p|)^-1,
    A = aT / sqrt(det(...)).  Returns (xt, pt, Mt, At) with shapes (b, Nt, d),
    (b, Nt, d), (b, Nt, d, d), (b, Nt, 1).  Inside ``MSGBSolver.adjoint`` the
    same propagation to t = 0 is fused into the adjoint parameter kernel."""
    import torch
    dev = xT.device if isinstance(xT, torch.Tensor) else torch.device("cuda")
    f64 = lambda v: torch.as_tensor(v, device=dev).to(torch.float64).contiguous()
    c128 = lambda v: torch.as_tensor(v, device=dev).to(torch.complex128).contiguous()
    xT, pT = f64(xT), f64(pT)
    b, d = pT.shape
    mT = c128(mT).reshape(b, d, d)
    aT = c128(aT).reshape(b, -1)[:, 0].contiguous()
    mode = f64(mode).reshape(b)
    tsb = _tr_time_grid(ts if not isinstance(ts, torch.Tensor) else ts.to(dev), b)
    tsb = tsb.to(device=dev, dtype=torch.float64).contiguous()
    nt = tsb.shape[1]
    xt = torch.empty((b, nt, d), dtype=torch.float64, device=dev)
    pt = torch.empty_like(xt)
    Mt = torch.empty((b, nt, d, d), dtype=torch.complex128, device=dev)
    At = torch.empty((b, nt, 1), dtype=torch.complex128, device=dev)
    _lib.call("bh_solve_hom_TR", b, d, _lib.ptr(xT), _lib.ptr(pT), _lib.ptr(mT), _lib.ptr(aT),
              _lib.ptr(mode), _c0(c), _lib.ptr(tsb), nt, _lib.ptr(xt), _lib.ptr(pt),
              _lib.ptr(Mt), _lib.ptr(At), _lib.stream_handle())
    return xt, pt, Mt, At


def speed_of(c):
    """The device speed model of a SolverFn ``c`` argument: a float, a
    ``GridSpeed``, or a ``Domain.c_fn`` / ``Domain``."""
    from ..geometry import Domain
    from ..utils.interp import GridSpeed
    if isinstance(c, GridSpeed):
        return c
    owner = getattr(c, "__self__", None)
    if isinstance(c, Domain) or isinstance(owner, Domain):
        dom = c if isinstance(c, Domain) else owner
        if dom.is_homogeneous:
            return GridSpeed(np.full((1,) * dom.ndim, float(dom.c)))
        return dom.speed_model
    if isinstance(c, (int, float, np.floating)):
        return None
    raise NotImplementedError("the sound speed must be a number, a grid (GridSpeed) or a "
                              "Domain's c_fn to run on the GPU")


def solve_ODE_base(x0, p0, M0, a0, mode, ts, c, lam=0.0, solver_config=None):
    """Beam ODEs in a (heterogeneous) medium, adaptive Dormand-Prince 5(4)
    (gb_solvers.py:870-1016), diagonal M0.  Returns (xt, pt, Mt, At) with
    shapes (b, Nt, d), (b, Nt, d), (b, Nt, d, d), (b, Nt, 1)."""
    import ctypes as C
    import torch
    from ..solvers.msgb_solver import ForwardBeams
    dev = x0.device if isinstance(x0, torch.Tensor) else torch.device("cuda")
    f64 = lambda v: torch.as_tensor(v, device=dev).to(torch.float64).contiguous()
    c128 = lambda v: torch.as_tensor(v, device=dev).to(torch.complex128).contiguous()
    x0, p0, ts = f64(x0), f64(p0), f64(ts).reshape(-1)
    mode = f64(mode).reshape(-1)
    M0 = c128(M0)
    a0 = c128(a0).reshape(-1)
    b, d = x0.shape
    alpha = torch.diagonal(M0, dim1=1, dim2=2)
    if bool((M0 - torch.diag_embed(alpha)).abs().max() > 0) if b else False:
        raise NotImplementedError("solve_ODE_base on the GPU takes a diagonal M0")
    sp = speed_of(c)
    if sp is None:
        from ..utils.interp import GridSpeed
        sp = GridSpeed(np.full((1,) * d, float(c)))
    cfg = solver_config if solver_config is not None else SolverConfig()
    nt = ts.numel()
    ns = 2 * d + 2 * d * d + 2
    xt = torch.empty((b, nt, d), dtype=torch.float64, device=dev)
    pt = torch.empty_like(xt)
    Mt = torch.empty((b, nt, d, d), dtype=torch.complex128, device=dev)
    At = torch.empty((b, nt, 1), dtype=torch.complex128, device=dev)
    for sign in (1.0, -1.0):
        sel = torch.nonzero(mode == sign).reshape(-1)
        k = sel.numel()
        if k == 0:
            continue
        beams = ForwardBeams(k, d, dev)
        beams.x0.copy_(x0[sel]); beams.p0.copy_(p0[sel]); beams.alpha.copy_(alpha[sel])
        beams.a0.copy_(a0[sel]); beams.omega.zero_()
        traj = torch.empty((nt, k, ns), dtype=torch.float64, device=dev)
        status = torch.empty(k, dtype=torch.int32, device=dev)
        sps = sp.struct(dev, lam=float(lam or 0.0))
        _lib.call("bh_ode_forward", C.byref(beams.struct()), d, 0 if sign > 0 else -1,
                  C.byref(sps), _lib.ptr(ts), nt, float(cfg.rtol), float(cfg.atol),
                  int(cfg.max_steps), _lib.ptr(traj), _lib.ptr(status), _lib.stream_handle())
        if bool((status != 0).any()):
            raise RuntimeError("beam ODE integration failed: more than max_steps steps or a "
                               "non-finite state")
        tr = traj.permute(1, 0, 2)                     # (k, nt, ns)
        xt[sel] = tr[..., :d]
        pt[sel] = tr[..., d:2 * d]
        Mt[sel] = torch.view_as_complex(tr[..., 2 * d:2 * d + 2 * d * d].contiguous()
                                        .reshape(k, nt, d * d, 2)).reshape(k, nt, d, d)
        At[sel] = torch.view_as_complex(tr[..., 2 * d + 2 * d * d:].contiguous()
                                        .reshape(k, nt, 1, 2))
    return xt, pt, Mt, At


def solve_ODE_batch_t(x0, p0, M0, a0, mode, ts, c, lam=None, solver_config=None):
    """Beam ODEs over per-beam time grids (gb_solvers.py:1017-1113) without
    absorption (coupled_rhs, 870-908), full complex M0, on the GPU
    (``bh_ode_static``).  ts: (b, 2) intervals (t0_i, t1_i), forward or
    backward in time.  Returns (xt, pt, Mt, At) with shapes (b, 2, d),
    (b, 2, d), (b, 2, d, d), (b, 2, 1).  Inside ``MSGBSolver`` it is the
    propagation of the adjoint and time-reversal beams in a heterogeneous
    medium."""
    import ctypes as C
    import torch
    from ..solvers.msgb_solver import StaticBeams
    dev = x0.device if isinstance(x0, torch.Tensor) else torch.device("cuda")
    f64 = lambda v: torch.as_tensor(v, device=dev).to(torch.float64).contiguous()
    c128 = lambda v: torch.as_tensor(v, device=dev).to(torch.complex128).contiguous()
    x0, p0, mode, ts = f64(x0), f64(p0), f64(mode).reshape(-1), f64(ts)
    b, d = x0.shape
    if ts.ndim != 2 or ts.shape[0] != b or ts.shape[1] != 2:
        raise NotImplementedError("solve_ODE_batch_t on the GPU takes ts of shape (b, 2): one "
                                  "(t0, t1) interval per beam")
    sp = speed_of(c)
    if sp is None:
        from ..utils.interp import GridSpeed
        sp = GridSpeed(np.full((1,) * d, float(c)))
    cfg = solver_config if solver_config is not None else SolverConfig()
    beams = StaticBeams(b, d, dev)
    beams.x.copy_(x0)
    beams.p.copy_(p0)
    beams.M.copy_(c128(M0).reshape(b, d, d))
    beams.A.copy_(c128(a0).reshape(-1))
    beams.omega.fill_(1.0)
    t0 = ts[:, 0].contiguous()
    t1 = ts[:, 1].contiguous()
    status = torch.empty(max(b, 1), dtype=torch.int32, device=dev)
    sps = sp.struct(dev)
    st = beams.struct()
    _lib.call("bh_ode_static", C.byref(st), d, C.byref(sps), _lib.ptr(t0), _lib.ptr(t1),
              _lib.ptr(mode), 0.0, float(cfg.rtol), float(cfg.atol), float(cfg.dt0 or 0.0),
              int(cfg.max_steps), _lib.ptr(status), _lib.stream_handle())
    if b and bool((status[:b] != 0).any()):
        raise RuntimeError("beam ODE integration failed: more than max_steps steps or a "
                           "non-finite state")
    x_0, p_0 = f64(x0), f64(p0)
    xt = torch.stack([x_0, beams.x], dim=1)
    pt = torch.stack([p_0, beams.p], dim=1)
    Mt = torch.stack([c128(M0).reshape(b, d, d), beams.M], dim=1)
    At = torch.stack([c128(a0).reshape(-1), beams.A], dim=1).reshape(b, 2, 1)
    return xt, pt, Mt, At


@dataclass(frozen=True)
class SolverConfig:
    """Numerical ODE settings (gb_solvers.py:597-703): rtol, atol and
    max_steps drive ``solve_ODE_base``; the closed forms take none."""

    rtol: float = 1e-7
    atol: float = 1e-9
    dt0: Optional[float] = None
    max_steps: int = 4096

    @classmethod
    def from_precision(cls, x64: bool = True) -> "SolverConfig":
        return cls() if x64 else cls(rtol=1e-4, atol=1e-6)

"""Grid-valued sound speed (``beamax/utils/interp.py``).

``make_c_function_from_grid`` keeps the reference signature and returns a
``GridSpeed``: the tensor-product interpolant (n-linear on the grid values,
or cubic B-spline on coefficients prefiltered once with
``scipy.ndimage.spline_filter``, optional Gaussian pre-smoothing, boundary
extension nearest/wrap/reflect/mirror) whose value, gradient and Hessian the
HIP kernels evaluate (``bh_speed_eval``, and inside the beam ODE integrator
``bh_ode_forward``).  Calling it evaluates on the GPU.
"""

from __future__ import annotations

import ctypes as C
from collections.abc import Sequence
from typing import Optional

import numpy as np

from .. import _lib

_BOUNDARY = {"clamp": "nearest", "nearest": "nearest", "edge": "nearest", "reflect": "reflect",
             "mirror": "mirror", "wrap": "wrap", "periodic": "wrap"}
_BND_CODE = {"nearest": 0, "wrap": 1, "reflect": 2, "mirror": 3}
_METHOD_CODE = {"linear": 0, "bspline3": 1}


def _canonical_boundary(boundary: str) -> str:
    try:
        return _BOUNDARY[boundary]
    except KeyError as exc:
        valid = ", ".join(sorted(_BOUNDARY))
        raise ValueError(f"Unsupported boundary='{boundary}'. Expected one of: {valid}") from exc


class GridSpeed:
    """Interpolated grid sound speed (interp.py:158-270)."""

    def __init__(self, c_map, spacing=None, origin=None, *, method: str = "linear",
                 boundary: str = "clamp", smooth_sigma=None):
        from scipy import ndimage
        method = method.lower()
        self.boundary = _canonical_boundary(boundary)
        values = np.asarray(c_map, dtype=np.float64)
        if values.ndim == 0:
            raise ValueError("c_map must have at least one dimension.")
        if smooth_sigma is not None:
            sigma = np.asarray(smooth_sigma, dtype=float)
            if sigma.ndim == 0:
                sigma = np.full(values.ndim, float(sigma))
            if sigma.shape != (values.ndim,):
                raise ValueError(f"smooth_sigma must be scalar or length {values.ndim}, got "
                                 f"shape {sigma.shape}.")
            if np.any(sigma < 0):
                raise ValueError("smooth_sigma entries must be non-negative.")
            values = ndimage.gaussian_filter(values, sigma=tuple(float(s) for s in sigma),
                                             mode=self.boundary)
        if method not in _METHOD_CODE:
            raise ValueError("method must be 'linear' or 'bspline3'.")
        self.method = method
        self.grid_values = values
        self.table = (values if method == "linear" else
                      ndimage.spline_filter(values, order=3, mode=self.boundary))
        d = values.ndim
        if d > 3:
            raise NotImplementedError("grid sound speed is built for 1-3 dimensions")
        sp = np.asarray(spacing if spacing is not None else (1.0,) * d, dtype=float)
        org = np.asarray(origin if origin is not None else (0.0,) * d, dtype=float)
        if sp.shape != (d,):
            raise ValueError(f"spacing must be length {d}, got shape {sp.shape}.")
        if org.shape != (d,):
            raise ValueError(f"origin must be length {d}, got shape {org.shape}.")
        if not np.all(np.isfinite(sp)) or np.any(sp <= 0):
            raise ValueError("spacing entries must be finite and strictly positive.")
        if not np.all(np.isfinite(org)):
            raise ValueError("origin entries must be finite.")
        self.ndim = d
        self.spacing = tuple(float(v) for v in sp)
        self.origin = tuple(float(v) for v in org)
        self._dev = {}

    def device_table(self, device):
        import torch
        key = str(device)
        if key not in self._dev:
            self._dev[key] = torch.as_tensor(np.ascontiguousarray(self.table), device=device)
        return self._dev[key]

    def struct(self, device, lam: float = 0.0) -> "_lib.Speed":
        t = self.device_table(device)
        s = _lib.Speed()
        s.ndim = self.ndim
        s.method = _METHOD_CODE[self.method]
        s.boundary = _BND_CODE[self.boundary]
        for a in range(3):
            s.N[a] = self.table.shape[a] if a < self.ndim else 1
            s.spacing[a] = self.spacing[a] if a < self.ndim else 1.0
            s.origin[a] = self.origin[a] if a < self.ndim else 0.0
        s.values = t.data_ptr()
        s.lam = float(lam)
        return s

    def evaluate(self, x, device=None):
        """(c, grad, hessian) at points x (..., d), on the GPU; torch tensors."""
        import torch
        dev = device if device is not None else torch.device("cuda")
        xs = torch.as_tensor(np.asarray(x, dtype=np.float64) if not isinstance(x, torch.Tensor)
                             else x, device=dev).to(torch.float64)
        shp = tuple(xs.shape[:-1])
        if xs.shape[-1] != self.ndim:
            raise ValueError(f"coords must have final dimension {self.ndim}; got shape "
                             f"{tuple(xs.shape)}.")
        pts = xs.reshape(-1, self.ndim).contiguous()
        n = pts.shape[0]
        c = torch.empty(n, dtype=torch.float64, device=dev)
        g = torch.empty((n, self.ndim), dtype=torch.float64, device=dev)
        h = torch.empty((n, self.ndim, self.ndim), dtype=torch.float64, device=dev)
        sp = self.struct(dev)
        _lib.call("bh_speed_eval", C.byref(sp), _lib.ptr(pts), n, _lib.ptr(c), _lib.ptr(g),
                  _lib.ptr(h), _lib.stream_handle())
        return (c.reshape(shp), g.reshape(shp + (self.ndim,)),
                h.reshape(shp + (self.ndim, self.ndim)))

    def __call__(self, x):
        return self.evaluate(x)[0]

    def grad(self, x):
        return self.evaluate(x)[1]

    def hessian(self, x):
        return self.evaluate(x)[2]


def make_c_function_from_grid(c_map, spacing=None, origin=None, *, method: str = "linear",
                              boundary: str = "clamp",
                              smooth_sigma: Optional[float | Sequence[float]] = None):
    """interp.py:158-270; returns a ``GridSpeed``."""
    return GridSpeed(c_map, spacing, origin, method=method, boundary=boundary,
                     smooth_sigma=smooth_sigma)

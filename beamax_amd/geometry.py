"""Domain and Sensor (host geometry).

API of ``beamax.geometry`` (reference ``beamax/geometry.py:15-616``).  A
scalar ``c`` is the homogeneous medium (closed-form propagators); a grid of
shape N becomes the reference's default interpolant (n-linear, clamp, or wrap
when every axis is periodic: geometry.py:176-221), as does a ``GridSpeed``
passed directly; other callables are accepted for API compatibility and
rejected by the solver.
"""

from __future__ import annotations

from typing import Optional, Tuple

import numpy as np


def _np_of(x):
    try:
        import torch
        if isinstance(x, torch.Tensor):
            return x.detach().cpu().numpy()
    except ImportError:
        pass
    return x


class Domain:
    """Axis-aligned rectangular domain (geometry.py:15-400)."""

    def __init__(self, N: Tuple[int, ...], dx: Tuple[float, ...], periodic: Tuple[bool, ...],
                 cfl: float = 0.3, c=1500.0, density=1.0, alpha_coeff=None, lam: float = 0.0,
                 alpha_power=None) -> None:
        Nv = np.asarray(N)
        if Nv.ndim != 1 or Nv.size == 0:
            raise ValueError("N must contain at least one spatial dimension.")
        if any(isinstance(n, (bool, np.bool_)) or not isinstance(n, (int, np.integer)) for n in Nv):
            raise ValueError(f"N entries must be integers; got {N}.")
        N = tuple(int(n) for n in Nv)
        if any(n <= 0 for n in N):
            raise ValueError(f"N entries must be positive; got {N}.")
        if len(dx) != len(N):
            raise ValueError(f"dx must have length {len(N)}, got {len(dx)}.")
        dxa = np.asarray(dx, dtype=float)
        if not np.all(np.isfinite(dxa)) or np.any(dxa <= 0):
            raise ValueError(f"dx entries must be finite and positive; got {dx}.")
        if len(periodic) != len(N):
            raise ValueError(f"periodic must have length {len(N)}, got {len(periodic)}.")
        if any(not isinstance(p, (bool, np.bool_)) for p in periodic):
            raise ValueError(f"periodic entries must be boolean; got {periodic}.")
        if not np.isfinite(cfl) or cfl <= 0:
            raise ValueError(f"cfl must be finite and positive; got {cfl}.")
        if not np.isfinite(lam) or lam < 0:
            raise ValueError(f"lam must be finite and non-negative; got {lam}.")

        def _param(name, value, allow_none, positive):
            if value is None:
                if allow_none:
                    return None
                raise ValueError(f"{name} cannot be None.")
            if callable(value):
                return value
            arr = np.asarray(value)
            if arr.ndim != 0 and tuple(arr.shape) != N:
                raise ValueError(f"{name} must be scalar, callable, or have shape {N}; got {arr.shape}.")
            if not np.all(np.isfinite(arr)):
                raise ValueError(f"{name} must contain only finite values.")
            if positive and np.any(arr <= 0):
                raise ValueError(f"{name} must be strictly positive.")
            if not positive and np.any(arr < 0):
                raise ValueError(f"{name} must be non-negative.")
            return float(arr) if arr.ndim == 0 else arr

        self.N = N
        self.dx = tuple(float(x) for x in dxa)
        self.periodic = tuple(bool(p) for p in periodic)
        self.cfl = float(cfl)
        self.c = _param("c", c, False, True)
        self.density = _param("density", density, True, True)
        self.alpha_coeff = _param("alpha_coeff", alpha_coeff, True, False)
        self.lam = float(lam)
        self.alpha_power = _param("alpha_power", alpha_power, True, False)

    @property
    def ndim(self) -> int:
        return len(self.N)

    @property
    def grid_size(self):
        """N * dx per axis (geometry.py:226-236)."""
        return np.asarray(self.N, dtype=float) * np.asarray(self.dx, dtype=float)

    @property
    def xmax(self) -> float:
        return float(np.linalg.norm(self.grid_size))

    @property
    def k_max(self) -> float:
        return float(np.pi * np.min(1.0 / np.asarray(self.dx)))

    def generate_meshgrid(self):
        sp = [np.arange(0, self.N[i], 1) * self.dx[i] for i in range(self.ndim)]
        fo = [np.arange(-self.N[i] // 2, self.N[i] // 2, 1) for i in range(self.ndim)]
        return np.meshgrid(*sp, indexing="ij"), np.meshgrid(*fo, indexing="ij")

    @property
    def grid(self):
        return np.stack(self.generate_meshgrid()[0], axis=-1)

    @property
    def is_homogeneous(self) -> bool:
        return isinstance(self.c, float)

    @property
    def speed_model(self):
        """The ``GridSpeed`` the HIP kernels evaluate (None for a scalar c)."""
        from .utils.interp import GridSpeed
        if isinstance(self.c, float):
            return None
        if isinstance(self.c, GridSpeed):
            return self.c
        if callable(self.c):
            raise NotImplementedError("a Python callable sound speed cannot run on the GPU; "
                                      "pass a grid (array of shape N) or a GridSpeed")
        if not hasattr(self, "_speed"):
            self._speed = GridSpeed(np.asarray(self.c, dtype=float), spacing=self.dx,
                                    origin=(0.0,) * self.ndim,
                                    boundary="wrap" if all(self.periodic) else "clamp")
        return self._speed

    @property
    def sound_speed_array(self):
        from .utils.interp import GridSpeed
        if isinstance(self.c, GridSpeed) and self.c.grid_values.shape == self.N \
                and self.c.spacing == self.dx and not any(self.c.origin):
            # the interpolant takes the (smoothed) grid values at the nodes
            return np.asarray(self.c.grid_values, dtype=float)
        if callable(self.c):
            return np.asarray(_np_of(self.c(self.grid)), dtype=float)
        return np.broadcast_to(np.asarray(self.c, dtype=float), self.N)

    def c_fn(self, x):
        """Sound speed at positions x (..., d)."""
        if isinstance(self.c, float):
            return np.full(np.asarray(x).shape[:-1], self.c)
        if callable(self.c):
            return self.c(x)
        return self.speed_model(x)

    def compute_max_speed(self) -> float:
        return float(np.max(self.sound_speed_array))

    def compute_min_speed(self) -> float:
        return float(np.min(self.sound_speed_array))

    def compute_max_freq(self) -> float:
        return self.compute_max_speed() / (2 * min(self.dx))

    def generate_time_domain(self):
        """dt = cfl*min(dx)/max(c); arange(0, |N dx| / min(c), dt) (geometry.py:387-400)."""
        dt = self.cfl * min(self.dx) / self.compute_max_speed()
        tmax = self.xmax / self.compute_min_speed()
        return np.arange(0, tmax, dt)


class Sensor:
    """Receiver/source geometry from positions or a binary mask (geometry.py:403-616).

    Mask-derived positions are the row-major nonzero indices times dx.
    """

    def __init__(self, domain: Domain, positions=None, binary_mask=None) -> None:
        if positions is None and binary_mask is None:
            raise ValueError("Either positions or binary_mask must be provided")
        if positions is not None and binary_mask is not None:
            raise ValueError("Cannot provide both positions and binary_mask")
        self.domain = domain
        if positions is not None:
            pos = self._normalize_positions(positions)
            self._positions = pos
            self._binary_mask = self._positions_to_mask(pos)
        else:
            mask = self._normalize_mask(binary_mask)
            self._binary_mask = mask
            self._positions = self._mask_to_positions(mask)

    def _normalize_positions(self, positions):
        pos = np.asarray(_to_numpy(positions), dtype=float)
        if pos.ndim == 1:
            pos = pos[None, :]
        if pos.ndim != 2 or pos.shape[1] != self.domain.ndim:
            raise ValueError("positions must have shape (Ns, ndim) or (ndim,), "
                             f"got {pos.shape} for ndim={self.domain.ndim}.")
        if not np.all(np.isfinite(pos)):
            raise ValueError("positions must contain only finite values.")
        if np.any(pos < 0) or np.any(pos >= self.domain.grid_size):
            raise ValueError("positions must lie inside the half-open domain [0, N*dx).")
        idx = np.clip(np.rint(pos / np.asarray(self.domain.dx)).astype(int), 0,
                      np.asarray(self.domain.N) - 1)
        if np.unique(idx, axis=0).shape[0] != idx.shape[0]:
            raise ValueError("positions must map to distinct grid points after quantisation.")
        return pos

    def _normalize_mask(self, mask):
        m = np.asarray(_to_numpy(mask))
        if tuple(m.shape) != tuple(self.domain.N):
            raise ValueError(f"binary_mask must have shape {self.domain.N}, got {m.shape}.")
        if not np.all(np.isfinite(m)):
            raise ValueError("binary_mask must contain only finite values.")
        if not np.all((m == 0) | (m == 1)):
            raise ValueError("binary_mask must contain only binary values 0 and 1.")
        if not bool(np.any(m == 1)):
            raise ValueError("binary_mask must contain at least one positive active entry.")
        return m

    def _positions_to_mask(self, pos):
        idx = np.clip(np.round(pos / np.asarray(self.domain.dx)).astype(int), 0,
                      np.asarray(self.domain.N) - 1)
        mask = np.zeros(self.domain.N)
        mask[tuple(idx.T)] = 1
        return mask

    def _mask_to_positions(self, mask):
        idx = np.nonzero(mask > 0)
        return np.stack(idx, axis=1) * np.asarray(self.domain.dx)

    @property
    def positions(self):
        return self._positions

    @property
    def binary_mask(self):
        return self._binary_mask


def _to_numpy(x):
    if hasattr(x, "detach"):
        return x.detach().cpu().numpy()
    return x

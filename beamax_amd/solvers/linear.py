"""The MSGB forward operator with a frozen beam set, and its exact transpose.

The reference forward (``msgb_solver.py:858-982``) re-selects the ``top_n``
coefficients of every input, so p0 -> data is not linear.  Freezing the
selected index set I makes it linear:

    F p0 = S(a_I * (1/2 * half_mask * W p0)[I])

with W the MSWPT analysis (``transforms.py:696-733``), a_I the unit-coefficient
beam amplitudes of ``compute_forward_parameters`` (``forward_solver_utils.py:
255-373``) and S the mirrored real beam sum (``gb/core.py:194-329``).  Its
transpose for the real inner products is

    F^T r = Re W^H (1/2 * scatter_I(S^T r)),

which the reference only reaches through JAX autodiff; here it is three HIP
entry points (``bh_forward_transpose``, ``bh_coeff_scatter``,
``bh_mswpt_adjoint``).  F^T is the operator CG on the normal equations needs
(``reconstruct.cgnr``).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib
from ..geometry import Domain, Sensor
from ..transforms import MSWPT
from .msgb_solver import ForwardBeams, MSGBSolver, _np


class FrozenForward:
    """Linear forward operator on a frozen index set ``idx`` (int64, device).

    ``apply(p0)`` -> (Nt, S) sensor data; ``transpose(r)`` -> field of shape
    ``domain.N``.  Both run the solver's precision (``fp64`` or
    ``fp32_split``) and envelope cut (``cull_log2``).
    """

    def __init__(self, solver: MSGBSolver, domain: Domain, sensors, ts, wpt: MSWPT, idx):
        import torch
        if any(domain.periodic):
            raise NotImplementedError("the frozen-index transpose supports free-space domains only")
        if not domain.is_homogeneous:
            raise NotImplementedError("heterogeneous sound speed is not on the GPU path")
        solver._check_supported()
        self.solver, self.domain, self.wpt = solver, domain, wpt
        dev = wpt.device
        d = domain.ndim
        pos = solver._sensor_positions(sensors)
        if pos is None:
            raise ValueError("Unsupported sensor type")
        self.points = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, d)
        self.ts = torch.as_tensor(np.ascontiguousarray(_np(ts), dtype=np.float64), device=dev)
        self.idx = torch.as_tensor(idx, dtype=torch.int64, device=dev).contiguous()
        self.k = int(self.idx.numel())
        self.seg = solver._segments(self.points, domain.dx, dev)
        # the same term set as sum_forward: culled only where it culls
        self._cull = solver.cull_log2 if (solver.cull_log2 > 0 and self.seg.efficient) else 0.0
        self._gs = _lib.dbl_array(domain.grid_size)
        ones = torch.ones(self.k, dtype=torch.complex128, device=dev)
        self.unit = ForwardBeams(self.k, d, dev)
        _lib.call("bh_forward_params", C.byref(wpt.tiling), self._gs, _lib.ptr(self.idx),
                  _lib.ptr(ones), self.k, self.unit.struct(), _lib.stream_handle())
        self._vals = torch.empty(self.k, dtype=torch.complex128, device=dev)
        self._beams = ForwardBeams(self.k, d, dev)
        self._grad = torch.empty(self.k, dtype=torch.complex128, device=dev)
        self._coeffs = torch.empty(wpt.total_coeffs, dtype=torch.complex128, device=dev)
        lib = _lib.load()
        self._tws = _lib.workspace(
            lib.bh_forward_transpose_workspace_size(self.k, self.ts.numel(), self.seg.ngroups), dev)
        self._mws = _lib.workspace(lib.bh_mswpt_workspace_size(C.byref(wpt.tiling)), dev)

    @classmethod
    def from_field(cls, solver: MSGBSolver, p, domain: Domain, sensors, ts, wpt: MSWPT):
        """Freeze the index set the reference's forward would select for ``p``."""
        cpos = wpt.forward(p, solver.input_type, scale=0.5, half_mask=True)
        idx, _ = solver._select(cpos, wpt, half_frame=True)
        return cls(solver, domain, sensors, ts, wpt, idx)

    @property
    def shape(self):
        return (self.ts.numel(), self.seg.S), tuple(self.domain.N)

    def apply(self, p0, out=None):
        """F p0: (Nt, S) data (float32 in fp32_split, float64 in fp64)."""
        wpt = self.wpt
        cpos = wpt.forward(p0, self.solver.input_type, scale=0.5, half_mask=True)
        _lib.call("bh_coeff_gather", _lib.ptr(cpos), wpt.total_coeffs, _lib.ptr(self.idx), self.k,
                  _lib.ptr(self._vals), _lib.stream_handle())
        _lib.call("bh_forward_params", C.byref(wpt.tiling), self._gs, _lib.ptr(self.idx),
                  _lib.ptr(self._vals), self.k, self._beams.struct(), _lib.stream_handle())
        return self.solver.sum_forward(self._beams, self.domain, self.points, self.ts, out=out)

    def transpose(self, r, out=None):
        """F^T r: real field of shape ``domain.N`` (float64)."""
        import torch
        nt, S = self.shape[0]
        r = r.reshape(nt, S)
        if r.dtype not in (torch.float32, torch.float64):
            r = r.to(torch.float64)
        r = r.contiguous()
        wpt = self.wpt
        b = self.unit.struct()
        ss = self.seg.struct()
        _lib.call("bh_forward_transpose", C.byref(b), self.domain.ndim, 1, float(self.domain.c),
                  _lib.ptr(self.ts), nt, C.byref(ss), self._cull,
                  _lib.PRECISION[self.solver.precision], _lib.ptr(r),
                  int(r.dtype == torch.float64), _lib.ptr(self._grad), _lib.ptr(self._tws),
                  self._tws.numel(), _lib.stream_handle())
        _lib.call("bh_coeff_scatter", _lib.ptr(self.idx), _lib.ptr(self._grad), self.k, 0.5,
                  wpt.total_coeffs, _lib.ptr(self._coeffs), _lib.stream_handle())
        return wpt.adjoint(self._coeffs, real=True, out=out, ws=self._mws)

    def normal(self, p, out=None):
        """F^T F p."""
        return self.transpose(self.apply(p), out=out)
